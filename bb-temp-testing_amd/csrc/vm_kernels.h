// vm_kernels.h -- kernel argument blocks and host-side launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "vm_types.h"

namespace vmk {

// Point / vector in the kernel's real type (layout of float2 / double2).
template <class R> struct V2 {
    R x, y;
};

// Gradient accumulator element (two per site, x then y).  FP32 runs sum the
// gradient terms in 64-bit fixed point, round(term * 2^gradShift), so the sum
// does not depend on the order the atomics land in; FP64 runs sum in double.
template <class R> struct GradAcc {
    using T = double;
};
template <> struct GradAcc<float> {
    using T = long long;
};
template <class R> using GradT = typename GradAcc<R>::T;

// Per-target diagnostics written by the loss kernel when requested.
struct TermDbg {
    int    i0, j;  // nearest site and bisector partner (instance-local), -1 if unused
    double loss;   // VoroLoss term (in the kernel's real type, widened)
};

template <class R> struct LossArgs {
    const Inst       *insts;
    InstState        *state;
    const TargetWork *work;
    const int        *nWork;
    const V2<R>      *pos;    // current positions (instance-local order)
    const V2<R>      *spos;   // positions sorted by cell
    const int        *sidx;   // instance-local site index of each sorted entry
    const int        *start;  // cell starts
    const V2<R>      *tpos;   // targets sorted by cell
    const int        *tslot;  // their slot numbers (subsampling hash index)
    GradT<R>         *grad;   // accumulated gradient [2 per site], null: loss only
    LossPartial      *partial; // [instances * kLossPartials]: step loss sums (iterate mode)
    int               mode;   // kModeIterate / kModeEvaluate
    int              *dbgKnn; // optional k-NN lists of instance dbgInst's targets ([T * K], instance-local indices)
    int               dbgInst;
    TermDbg          *dbgTerm;
    unsigned long long *stats; // optional [64 x 8] path counters: fast, ring (staged), ring (global),
                               // idle, candidates, staged sites, fast waves with an unscreened lane,
                               // unscreened targets in those waves (summed over the 64 lines)
    int               exactOnly; // 1: every lane takes the exact selection (parity checks of that path)
    ExactItem        *exactQ;   // targets queued for the exact selection; capacity >= the group's targets
    int              *exactLen; // queue length (zeroed by the scan kernel)
    int               exactGrid; // workgroups of the exact-selection launch (4 wavefronts each)
};

template <class R> struct AdamArgs {
    const Inst          *insts;
    InstState           *state;
    const SiteWork      *work;
    const int           *nWork;
    V2<R>               *pos, *m, *v;
    GradT<R>            *grad;
    const unsigned char *frozen;
    const LossPartial   *partial; // folded into InstState (count, loss) before the update
    int                 *cellOf, *rank, *counts;
    int                  doAdam; // 0: only bin the current positions
};

template <class R> struct ClipArgs {
    const Inst     *insts;
    int             inst;
    const V2<R>    *pos, *spos;
    const int      *sidx, *start;
    const uint64_t *masks;
    int             tilesX, tilesY;
    float           invVoxel;
    float          *verts;   // [S*24*2]
    int            *labels;  // [S*24]
    int            *count;   // [S]
    unsigned char  *inside;  // [S]
    int            *flags;   // {unsafe, overflow}
};

// Boundary of the inside cells, assembled on the device from the clip outputs
// (voro_optimizer.cu:798-854): every edge of an inside cell facing an outside
// cell or the box, its end points welded by their sorted label triple.  Vertex
// ids follow the first appearance of each triple in the order (cell, edge,
// start before end), as a sequential pass would number them.
struct BndKey {
    uint32_t a, b, c, pad; // sorted label triple, box sides shifted by kBoxLabels
};
struct BoundaryArgs {
    int                  S;      // cells (the instance's live sites)
    const int           *count;  // clip outputs: [S], [S*24], [S*24*2], [S]
    const int           *labels;
    const float         *verts;
    const unsigned char *inside;
    int                  cap;    // end point capacity of the arrays below (2 per edge)
    int                 *ecount; // [S] boundary edges per cell
    int                 *eoff;   // [S] their exclusive offsets
    BndKey              *key;    // [cap] end point keys: 2e = start, 2e+1 = end of edge e
    float               *pos;    // [cap*2] end point positions
    uint32_t            *table;  // [tableMask+1 >= 2 cap] open-addressing table
    uint32_t             tableMask;
    int                 *slot;   // [cap] table slot of each end point
    int                 *rep;    // [cap] first end point with the same key
    int                 *vid;    // [cap] vertex id of each representative
    float               *vertices; // [cap*2] out: V vertices
    int                 *edges;    // [cap] out: 2E vertex ids
    int                 *outDeg, *inDeg, *nxt; // [cap] (V <= 2E)
    int                 *lab0, *lab1, *nxt1;   // [cap] loop-count scratch
    int                 *result; // [5] out: V, watertight, loops, overflow (more than cap end points), E
};
// One workgroup: count, emit, weld, degrees, watertight test and loop count.
hipError_t launchBoundary(hipStream_t s, const BoundaryArgs &a);

// Cells per scan tile, and the tiles of an instance with `cells` cells (gw*gh+1).
constexpr int kScanTileCells = 16384;
inline int scanTilesFor(int64_t cells) { return int((cells + kScanTileCells - 1) / kScanTileCells); }

// Scatter folded into the scan when every instance has a single tile
// (posBytes = sizeof(V2<R>); 0: the scatter kernel follows).
struct ScanScatter {
    const void *pos;
    const int  *cellOf, *rank;
    void       *spos;
    int        *sidx;
    int         posBytes;
};
hipError_t launchScan(int grid, hipStream_t s, const Inst *insts, InstState *state, LossPartial *partial,
                      const ScanTile *tiles, ScanCtl *ctl, unsigned long long *flags, int *counts, int *start,
                      int mode, int *exactLen, int nQueues, const ScanScatter &sc);
template <class R>
hipError_t launchScatter(int grid, hipStream_t s, const Inst *insts, const InstState *state, const SiteWork *work,
                         const int *nWork, const void *pos, const int *cellOf, const int *rank, const int *start,
                         void *spos, int *sidx, int mode, ScanCtl *ctl);
// One loss launch over a group of chunks of fits sharing targets per lane (each
// fit's k comes from its instance record).
template <class R> hipError_t launchLoss(int tpl, int grid, hipStream_t s, const LossArgs<R> &a);
// The exact selection for the targets the loss launch queued; always follows it.
template <class R> hipError_t launchExact(hipStream_t s, const LossArgs<R> &a);
template <class R> hipError_t launchAdamBin(int grid, hipStream_t s, const AdamArgs<R> &a);
template <class R>
hipError_t launchKnnQuery(int K, hipStream_t s, const Inst *insts, int inst, const void *spos, const int *sidx,
                          const int *start, const void *q, int nq, int *outIdx, void *outD2);
template <class R> hipError_t launchClip(int nSites, hipStream_t s, const ClipArgs<R> &a);
// Per-cell clip output (kMaxCellVerts slots per cell) -> CSR: start[n + 1] and the
// cells' vertices / labels back to back.  tileSum holds csrTiles(n) ints.
int        csrTiles(int n);
hipError_t launchClipCsr(hipStream_t s, const float *verts, const int *labels, const int *count, int n, int *start,
                         int *tileSum, float *cverts, int *clabels);
hipError_t launchPoissonClear(int grid, hipStream_t s, const PoiInst *insts, const PoiWork *work, int nWork, float *g);
hipError_t launchPoissonPhase(int grid, hipStream_t s, const PoiInst *insts, const PoiWork *work, int nWork, float *g,
                              int px, int py, int round);
hipError_t launchFillInt(hipStream_t s, int *p, int64_t n, int v);
template <class R> hipError_t launchKillTargets(hipStream_t s, V2<R> *tpos, int *tslot, const int64_t *idx, int n);
// One-launch resample of a region whose 2-cell halo fits kPoiRegionCells cells of LDS.
constexpr int kPoiRegionCells   = 8000; // 64,000 B of float2
constexpr int kPoiRegionThreads = 1024;
bool          poissonRegionFits(const PoiInst &pi);
hipError_t    launchPoissonRegion(hipStream_t s, const PoiInst &pi, float *grid, int trials);

// Workgroups for the exact-selection launch of a group with `targets` targets:
// every target takes it in FP64 runs, with VM_LOSS_EXACT and for the k-NN list
// diagnostics (one wavefront per target); otherwise only the targets of chunks
// that could not be staged do (the screen's open targets take the selection
// inside their chunk), about one wavefront per 8192 targets.
inline int exactGridFor(int64_t targets, bool allExact)
{
    const int64_t g = allExact ? (targets + 3) / 4 : (targets + 32767) / 32768;
    return int(g < 16 ? 16 : (g > 4096 ? 4096 : g));
}

inline bool lossKSupported(int K) { return K == 2 || K == 4 || K == 6 || K == 8 || K == 12 || K == 16 || K == 24 || K == 32; }
// Targets per lane compiled: 2, 3 and 4 for FP32 runs (the kernels take k from each instance's record); 3 for
// FP64 runs (they send every target to the exact selection, so the chunk shape matters little there).
inline bool lossTplSupported(int tpl, bool fp64)
{
    return tpl == kTargetsPerLane || (!fp64 && (tpl == 2 || tpl == 4));
}

} // namespace vmk
