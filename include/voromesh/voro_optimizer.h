#pragma once
// voro_optimizer.h -- voromesh::VoroOptimizer on MI355X (gfx950).
//
// Drop-in for the reference's C++ API (include/voromesh/voro_optimizer.h:34-138):
// same types, same methods, same semantics, same fail-fast error behaviour (a
// failing call prints the error on stderr and exits with status 1).  The class
// is header-only and talks to libvoromesh_hip.so through its C ABI
// (voromesh_hip.h); link with -lvoromesh_hip.
#include <array>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "voromesh/shape2d.h"
#include "voromesh_hip.h"

namespace voromesh {

struct OptimizerParams
{
    int                iterations      = 400;
    int                localIterations = 100;
    float              lrPerSpacing    = 0.09f;
    float              sampleFraction  = 0.2f;
    float              gamma           = 0.5f;
    std::array<int, 4> milestones      = {80, 150, 200, 250};
    unsigned           seed            = 1;
    // Knobs the reference fixes at compile time (voro_optimizer.cu:54-57), appended
    // so aggregate initialisation of the fields above is unchanged.
    int   lossK      = 12;   // neighbours per target
    float gridCell   = 0.0f; // hash-grid cell in world units (0: library choice)
    int   stageRings = 0;    // cell rings the loss kernel stages in LDS (0: library choice)
    int   precision  = 0;    // 0: FP32 sites, gradients and Adam; 1: all FP64
};

struct VoroMesh2D
{
    std::vector<Vec2>    generators;
    std::vector<uint8_t> inside;
    std::vector<int>     cellStart;
    std::vector<Vec2>    cellVerts;

    std::vector<Vec2>               vertices;
    std::vector<std::array<int, 2>> edges;

    bool  watertight    = false;
    int   loops         = 0;
    int   unsafeCells   = 0;
    int   overflowCells = 0;
    float gpuMs         = 0.0f;
    float hostMs        = 0.0f;
};

namespace detail {
inline void check(int status, const char *what)
{
    if (status != VM_OK) {
        std::fprintf(stderr, "voromesh: %s failed (status %d): %s\n", what, status, vm_last_error());
        std::exit(1);
    }
}
} // namespace detail

class VoroOptimizer
{
public:
    VoroOptimizer(float width, float height, float spacing, const OptimizerParams &params = {})
        : mParams(params)
    {
        vm_params p;
        vm_default_params(&p);
        p.iterations       = params.iterations;
        p.local_iterations = params.localIterations;
        p.lr_per_spacing   = params.lrPerSpacing;
        p.sample_fraction  = params.sampleFraction;
        p.gamma            = params.gamma;
        for (int k = 0; k < 4; ++k) p.milestones[k] = params.milestones[k];
        p.seed        = params.seed;
        p.loss_k      = params.lossK;
        p.grid_cell   = params.gridCell;
        p.stage_rings = params.stageRings;
        p.precision   = params.precision;
        detail::check(vm_create(width, height, spacing, &p, &mH), "VoroOptimizer");
    }
    ~VoroOptimizer() { vm_destroy(mH); }

    VoroOptimizer(const VoroOptimizer &) = delete;
    VoroOptimizer &operator=(const VoroOptimizer &) = delete;

    void SetTargets(const std::vector<Vec2> &slots)
    {
        detail::check(vm_set_targets(mH, reinterpret_cast<const vm_vec2 *>(slots.data()), int(slots.size())),
                      "SetTargets");
    }
    void UpdateTargets(const std::vector<int> &dirty, const std::vector<Vec2> &slots)
    {
        detail::check(vm_update_targets(mH, dirty.data(), int(dirty.size()),
                                        reinterpret_cast<const vm_vec2 *>(slots.data()), int(slots.size())),
                      "UpdateTargets");
    }

    void SetOccupancy(const std::vector<uint64_t> &masks, int tilesX, int tilesY, float voxelSize)
    {
        detail::check(vm_set_occupancy(mH, masks.data(), tilesX, tilesY, voxelSize), "SetOccupancy");
    }
    void UpdateOccupancy(const std::vector<uint64_t> &masks, int tx0, int ty0, int tx1, int ty1)
    {
        detail::check(vm_update_occupancy(mH, masks.data(), tx0, ty0, tx1, ty1), "UpdateOccupancy");
    }

    void SetGenerators(const Vec2 *slots, int slotsW, int slotsH, float slotCellSize)
    {
        detail::check(vm_set_generators(mH, reinterpret_cast<const vm_vec2 *>(slots), slotsW, slotsH, slotCellSize),
                      "SetGenerators");
        mSlotsW = slotsW, mSlotsH = slotsH;
    }
    void UpdateGenerators(const Vec2 *slots, int cx0, int cy0, int cx1, int cy1)
    {
        detail::check(vm_update_generators(mH, reinterpret_cast<const vm_vec2 *>(slots), cx0, cy0, cx1, cy1),
                      "UpdateGenerators");
    }

    void BeginSchedule(int iterations, int cx0, int cy0, int cx1, int cy1)
    {
        detail::check(vm_begin_schedule(mH, iterations, cx0, cy0, cx1, cy1), "BeginSchedule");
    }
    void BeginSchedule(int iterations) { BeginSchedule(iterations, 0, 0, mSlotsW, mSlotsH); }

    float Step(int iterations)
    {
        float ms = 0.0f;
        detail::check(vm_step(mH, iterations, &ms), "Step");
        return ms;
    }
    float Optimize() { return Step(ScheduleLength() - Iteration()); }

    float LastBatchLoss() const { return state().lastBatchLoss; }
    float EvaluateLoss()
    {
        double mean = 0.0;
        detail::check(vm_evaluate_loss(mH, &mean), "EvaluateLoss");
        return float(mean);
    }

    VoroMesh2D Extract()
    {
        vm_mesh m;
        detail::check(vm_extract(mH, &m), "Extract");
        VoroMesh2D out;
        const int n = m.num_slots;
        auto      v = [](const vm_vec2 *p, size_t k) { return Vec2{p[k].x, p[k].y}; };
        out.generators.resize(n);
        for (int i = 0; i < n; ++i) out.generators[i] = v(m.generators, i);
        out.inside.assign(m.inside, m.inside + n);
        out.cellStart.assign(m.cell_start, m.cell_start + n + 1);
        out.cellVerts.resize(out.cellStart[n]);
        for (int i = 0; i < out.cellStart[n]; ++i) out.cellVerts[i] = v(m.cell_verts, i);
        out.vertices.resize(m.num_vertices);
        for (int i = 0; i < m.num_vertices; ++i) out.vertices[i] = v(m.vertices, i);
        out.edges.resize(m.num_edges);
        for (int i = 0; i < m.num_edges; ++i) out.edges[i] = {m.edges[2 * i], m.edges[2 * i + 1]};
        out.watertight    = m.watertight != 0;
        out.loops         = m.loops;
        out.unsafeCells   = m.unsafe_cells;
        out.overflowCells = m.overflow_cells;
        out.gpuMs         = m.gpu_ms;
        out.hostMs        = m.host_ms;
        return out;
    }

    int                    Iteration() const { return state().iteration; }
    int                    ScheduleLength() const { return state().scheduleLength; }
    int                    NumSlots() const { return mSlotsW * mSlotsH; }
    const OptimizerParams &Params() const { return mParams; }

    // Not in the reference: how many FP32 gradient terms had a component clamped before the fixed-point sum
    // (0 unless two neighbouring sites nearly coincide).
    int64_t GradSaturation() const
    {
        int64_t n = 0;
        detail::check(vm_grad_saturation(mH, &n), "GradSaturation");
        return n;
    }

private:
    struct State
    {
        int   iteration, scheduleLength;
        float lastBatchLoss;
    };
    State state() const
    {
        State s{};
        detail::check(vm_get_state(mH, &s.iteration, &s.scheduleLength, &s.lastBatchLoss, nullptr), "state");
        return s;
    }

    OptimizerParams mParams;
    vm_handle      *mH     = nullptr;
    int             mSlotsW = 0, mSlotsH = 0;
};

} // namespace voromesh
