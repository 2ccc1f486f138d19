// vm_engine.h -- host runtime of the HIP optimizer: a batch of instances in
// pooled device arrays, driven by the five per-iteration launches of
// vm_kernels.hip, one iteration captured as a HIP graph.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "vm_kernels.h"

namespace vmrt {

struct VmError : std::runtime_error {
    int code;
    VmError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hipCheck(hipError_t e, const char *what, const char *file, int line);

// Runs f, mapping exceptions to a vm_status code and the thread's last error.
int         guardedCall(const std::function<void()> &f);
const char *lastError();
#define VM_HIP(x) ::vmrt::hipCheck((x), #x, __FILE__, __LINE__)

template <class T> struct DBuf {
    T     *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Reallocates (contents lost) when the capacity is too small.
    void reserve(size_t cnt)
    {
        if (cnt <= n && p) return;
        release();
        size_t c = cnt ? cnt : 1;
        VM_HIP(hipMalloc(reinterpret_cast<void **>(&p), c * sizeof(T)));
        n = c;
    }
};

// Pinned host staging buffer (hipHostMalloc), so device-to-host copies run as
// asynchronous DMA.  Pinning is slow, so a buffer that has to grow takes at
// least half as much again as it held (contents lost).
template <class T> struct HBuf {
    T     *p = nullptr;
    size_t n = 0;
    HBuf() = default;
    HBuf(const HBuf &) = delete;
    HBuf &operator=(const HBuf &) = delete;
    ~HBuf() { release(); }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
    void reserve(size_t cnt)
    {
        if (cnt <= n && p) return;
        const size_t c = std::max<size_t>(cnt ? cnt : 1, n + n / 2);
        release();
        VM_HIP(hipHostMalloc(reinterpret_cast<void **>(&p), c * sizeof(T), hipHostMallocDefault));
        n = c;
    }
};

// Host description of one instance.  Positions and moments are kept in double
// on the host so either device precision round-trips exactly.
struct InstSpec {
    float               W = 1, H = 1;
    std::vector<double> pos, m, v;   // xy interleaved, live sites
    std::vector<double> tgt;         // xy interleaved, live targets
    std::vector<int>    tslot;       // slot number of each live target (subsampling hash index)
    std::vector<uint8_t> frozen;     // empty: all sites move
    double lr = 1e-3, gamma = 0.5, beta1 = 0.9, beta2 = 0.999, eps = 1e-8;
    int    milestones[4] = {80, 150, 200, 250};
    float  fraction = 1.0f;
    uint32_t seed   = 1;
    int    K = 16;
    float  cell = 0.0f;      // hash-grid cell
    int    stageR = 2;
    float  chunkExt = 1.6f;  // target chunk diagonal, in mean site spacings
    int    tpl      = vmk::kTargetsPerLane; // loss kernel: targets per lane (chunks of up to 64 * tpl)
    int    useBox = 0;
    float  box[4] = {0, 0, 0, 0};
    int    maxIter = 0;
    float  tol = 0.0f;
    vmk::InstState state{};  // host copy of the device state
    // Slot layouts (the plain optimizer's): pos holds every generator slot,
    // empty ones with x < 0; liveS counts the live ones (-1: every entry is
    // live), which set the mean spacing.  rectW > 0: only slots in
    // [rect[0], rect[2]) x [rect[1], rect[3]) of a rectW-wide slot grid move.
    int    liveS = -1;
    int    rectW = 0;
    int    rect[4] = {0, 0, 0, 0};
};

struct DiagOut {
    std::vector<double> grad;      // [2S]
    double              lossSum = 0;
    int64_t             count   = 0;
    std::vector<int>    knn;       // [T*K] per live target (input order)
    std::vector<vmk::TermDbg> terms; // [T] per live target (input order)
};

// Boundary mesh of the inside cells: vertices welded by label triple, edges as
// vertex-id pairs, watertight iff every vertex has one in- and one out-edge.
struct Boundary {
    std::vector<float>   vertices; // [V * 2]
    std::vector<int32_t> edges;    // [E * 2]
    bool                 watertight = false;
    int                  loops      = 0;
};

// The clip's per-cell outputs are views of the engine's pinned staging
// buffers, valid until its next clip.
struct ClipOut {
    int            S      = 0;
    const int     *start  = nullptr; // [S+1]: cell i's vertices are [start[i], start[i+1])
    const float   *verts  = nullptr; // [start[S]*2]
    const int     *labels = nullptr; // [start[S]]
    const uint8_t *inside = nullptr; // [S]
    int            unsafe = 0, overflow = 0;
    float          gpuMs  = 0;
    Boundary       bound; // assembled on the device (site-index labels)
};

class Engine {
public:
    virtual ~Engine() = default;
    static std::unique_ptr<Engine> create(int precision, hipStream_t stream);

    std::vector<InstSpec> specs;

    int  add(InstSpec s)
    {
        specs.push_back(std::move(s));
        dirty = true;
        return int(specs.size()) - 1;
    }
    // Brings host specs up to date with the device (positions, moments, state).
    virtual void pull() = 0;
    // Lays out, uploads and bins every instance; rebuilds the iteration graph.
    virtual void commit() = 0;
    virtual float step(int iterations) = 0;                 // GPU ms
    // Mean ms per phase over `iterations` iterations: scan, scatter, loss, exact selection, adam.
    static constexpr int kPhases = 5;
    virtual void  profile(int iterations, float *phaseMs) = 0;
    virtual void  readStates() = 0;                          // device -> specs[i].state
    virtual double evaluate(int id, int64_t *count) = 0;     // loss sum over all live targets of `id`
    virtual void  diag(int id, bool wantGrad, bool wantDbg, DiagOut &out) = 0;
    // One evaluation loss pass over every instance: gradient sums [2 * sum S] and per-target terms [sum T]
    // (instance by instance, live targets in input order; site indices instance-local).
    virtual void  diagAll(bool wantGrad, bool wantTerms, DiagOut &out) = 0;
    virtual void  knnQuery(int id, const double *q, int nq, int K, int *idx, double *d2) = 0;
    virtual void  clip(int id, const uint64_t *dMasks, int tilesX, int tilesY, float voxel, ClipOut &out) = 0;
    virtual size_t deviceBytes() const = 0;
    // One evaluation-mode loss pass with the loss kernel's path counters.
    virtual void lossStats(uint64_t out[8]) = 0;

    // ---- local edits of instance 0 of a committed engine (the plain optimizer) ----
    // Room kept past the targets at commit: `tailTargets` targets and
    // `tailWork` chunks that appendTargets() lays out without a re-layout.
    int64_t tailTargets = 0;
    int     tailWork    = 0;
    // Sorted position -> input index of instance `id`'s targets at the last commit.
    virtual const std::vector<int> &permOf(int id) const = 0;
    // Retires pooled targets (sorted positions): their entries become dead points.
    virtual void killTargets(const std::vector<int64_t> &pos) = 0;
    // Lays out extra targets of instance 0 (xy interleaved, with their slot
    // numbers) as chunks of their own in the reserved tail; pos[k] receives the
    // pooled position of target k.  False (nothing changed) when the tail is full.
    virtual bool appendTargets(const std::vector<double> &tgt, const std::vector<int> &tslot,
                               std::vector<int64_t> &pos) = 0;
    // Generator slots [x0, x1) x [y0, y1) of instance 0 (slot layout, slotsW wide)
    // from the row-major slot array xy (null: positions kept); zeroMoments also
    // clears their Adam moments.
    virtual void uploadSiteRect(int slotsW, int x0, int y0, int x1, int y1, const float *xy, bool zeroMoments) = 0;
    // Instance id's record from its spec without a re-layout (learning rate,
    // schedule, target box, movable rectangle; gradient shift for liveTargets)
    // and, when asked, a zero step state.
    virtual void updateRecord(int id, int liveTargets, bool resetState) = 0;
    // Bins every site again (after positions were replaced).
    virtual void rebin() = 0;

    bool        dirty = true;
    bool        deviceNewer = false;
    hipStream_t stream = nullptr;
    // Iteration-graph bookkeeping: captures taken by an existing exec in place
    // (hipGraphExecUpdate) and execs instantiated (first use, or a new shape).
    int64_t     graphUpdates = 0, graphInstantiations = 0;
};

// Boundary mesh of clipped cells (voro_optimizer.cu:798-854): every edge of an
// inside cell whose other side is an outside cell or the box.  Cells are CSR
// (start[n + 1], verts, labels: >= 0 the neighbour cell, -1..-4 the box sides)
// in one index space; vertices are identified by the sorted triple of cells /
// box sides they are equidistant to and numbered in order of first use, so the
// mesh is watertight by construction.  watertight: every vertex has exactly
// one outgoing and one incoming edge; loops then counts the cycles.

// Lays out one instance at the given pooled offsets: its Inst record, site
// chunks, target chunks (one wavefront each) and the sort permutation of its
// targets (sorted position -> input index).
// extentCap bounds the chunk extent (mean spacings) below the spec's choice.
void layoutInstance(const InstSpec &s, int id, int64_t siteOff, int64_t cellOff, int64_t targetOff, vmk::Inst &in,
                    std::vector<vmk::SiteWork> &siteWork, std::vector<vmk::TargetWork> &targetWork,
                    std::vector<int> &perm, float extentCap = 1e30f);

// A layout with fewer target chunks than this (one wavefront each) is re-chunked
// with extents of at most kSmallLayoutExtent spacings, so small fits still spread
// over the chip.
constexpr int   kMinLayoutChunks   = 8192;
constexpr float kSmallLayoutExtent = 0.7f;

// Fixed-point shift of the FP32 gradient sums of an instance with T targets in
// a domain of extent `ext`.
int gradShiftFor(int numTargets, float ext);

// The tuner's per-size-class choice (vm_tune_table.h) for a fit of numSites
// sites with k = K: hash-grid cell, staged rings, target chunk extent (mean
// spacings) and targets per lane (kept to what is compiled for K).
void tunedChoice(int numSites, float W, float H, int K, float *cell, int *stageR, float *chunkExt, int *tpl);
// The tuner's k for the size class of a fit of numSites sites (the smallest k
// whose loss terms equalled those of k = 32 on the class's tuning fits).
int tunedK(int numSites);

} // namespace vmrt
