// dropin_demo.cpp -- the reference demo's call sequences against the drop-in
// headers: Editor::Reset (apps/voromesh_demo.cpp:103-113) and Editor::Stamp
// (:117-156) on a synthetic canvas (a star outline rasterised into target
// slots and 8x8-voxel occupancy tiles), headless.  Prints one JSON line and
// exits 0 when both extracted boundaries are watertight with the expected loop
// counts and the loss went down.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "poisson/poisson_disk.h"
#include "voromesh/voro_optimizer.h"

using voromesh::Vec2;

namespace {

constexpr float kDomain = 1.0f;
constexpr int   kVox    = 256; // voxels per side

struct Canvas
{
    std::vector<Vec2>     slots; // target slots, {-1,-1} = free
    std::vector<uint64_t> masks; // 8x8-voxel tiles, bit (j&7)*8 + (i&7)
    int                   tiles = kVox / 8;
    float                 voxel = kDomain / kVox;

    bool inside(int i, int j) const { return (masks[(j >> 3) * tiles + (i >> 3)] >> (((j & 7) << 3) | (i & 7))) & 1ull; }
    void set(int i, int j, bool v)
    {
        uint64_t &m = masks[(j >> 3) * tiles + (i >> 3)];
        uint64_t  b = 1ull << (((j & 7) << 3) | (i & 7));
        m           = v ? (m | b) : (m & ~b);
    }
};

std::vector<Vec2> star()
{
    std::vector<Vec2> p;
    for (int k = 0; k < 10; ++k) {
        float a = 1.5707963f + float(k) * 0.62831853f, r = (k % 2) ? 0.18f : 0.40f;
        p.push_back({0.5f + r * std::cos(a), 0.5f + r * std::sin(a)});
    }
    return p;
}

bool inPolygon(const std::vector<Vec2> &poly, float x, float y)
{
    bool in = false;
    for (size_t s = 0, t = poly.size() - 1; s < poly.size(); t = s++) {
        Vec2 a = poly[s], b = poly[t];
        if ((a.y > y) != (b.y > y) && x < a.x + (y - a.y) * (b.x - a.x) / (b.y - a.y)) in = !in;
    }
    return in;
}

Canvas rasterize(const std::vector<Vec2> &poly)
{
    Canvas c;
    c.masks.assign(size_t(c.tiles) * c.tiles, 0);
    for (int j = 0; j < kVox; ++j)
        for (int i = 0; i < kVox; ++i) c.set(i, j, inPolygon(poly, (i + 0.5f) * c.voxel, (j + 0.5f) * c.voxel));
    for (size_t s = 0; s < poly.size(); ++s) {
        Vec2  a = poly[s], b = poly[(s + 1) % poly.size()];
        float len = std::hypot(b.x - a.x, b.y - a.y);
        int   n   = std::max(1, int(std::ceil(len / (0.5f * c.voxel))));
        for (int k = 0; k < n; ++k) {
            float t = float(k) / n;
            c.slots.push_back({a.x + t * (b.x - a.x), a.y + t * (b.y - a.y)});
        }
    }
    return c;
}

} // namespace

int main()
{
    const float radius = 0.02f;
    voromesh::OptimizerParams params;
    params.lossK = 16;
    voromesh::VoroOptimizer opt(kDomain, kDomain, radius, params);

    // ---- Editor::Reset -------------------------------------------------------
    std::vector<Vec2> poly   = star();
    Canvas            canvas = rasterize(poly);
    opt.SetTargets(canvas.slots);
    opt.SetOccupancy(canvas.masks, canvas.tiles, canvas.tiles, canvas.voxel);
    poisson::Sampler sampler(kDomain, kDomain, radius, 30, 1);
    bool samplesOk = sampler.VerifyAll();
    auto slots     = [&] { return reinterpret_cast<const Vec2 *>(sampler.Cells().data()); };
    opt.SetGenerators(slots(), sampler.GridW(), sampler.GridH(), sampler.CellSize());
    float loss0    = opt.EvaluateLoss();
    float globalMs = opt.Optimize();
    float loss1    = opt.EvaluateLoss();
    voromesh::VoroMesh2D mesh = opt.Extract();
    bool  ok1 = mesh.watertight && mesh.loops == 1 && mesh.overflowCells == 0;

    // ---- Editor::Stamp: carve a disc out of the right arm ---------------------
    Vec2  c{0.80f, 0.60f};
    float brush = 0.06f;
    std::vector<int> dirty;
    for (size_t s = 0; s < canvas.slots.size(); ++s) {
        Vec2 p = canvas.slots[s];
        if (p.x >= 0 && std::hypot(p.x - c.x, p.y - c.y) < brush) {
            canvas.slots[s] = {-1.0f, -1.0f};
            dirty.push_back(int(s));
        }
    }
    for (int k = 0; k < 256; ++k) { // rim points that lie inside the object become surface
        float a = 6.2831853f * float(k) / 256.0f;
        Vec2  p{c.x + brush * std::cos(a), c.y + brush * std::sin(a)};
        if (inPolygon(poly, p.x, p.y)) {
            dirty.push_back(int(canvas.slots.size()));
            canvas.slots.push_back(p);
        }
    }
    int tx0 = canvas.tiles, ty0 = canvas.tiles, tx1 = 0, ty1 = 0;
    for (int j = 0; j < kVox; ++j)
        for (int i = 0; i < kVox; ++i) {
            float x = (i + 0.5f) * canvas.voxel, y = (j + 0.5f) * canvas.voxel;
            if (std::hypot(x - c.x, y - c.y) < brush && canvas.inside(i, j)) {
                canvas.set(i, j, false);
                tx0 = std::min(tx0, i >> 3), ty0 = std::min(ty0, j >> 3);
                tx1 = std::max(tx1, (i >> 3) + 1), ty1 = std::max(ty1, (j >> 3) + 1);
            }
        }
    opt.UpdateTargets(dirty, canvas.slots);
    opt.UpdateOccupancy(canvas.masks, tx0, ty0, tx1, ty1);
    float margin = brush + 2.0f * radius, cs = sampler.CellSize();
    int   rx0 = std::max(int(std::floor((c.x - margin) / cs)), 0);
    int   ry0 = std::max(int(std::floor((c.y - margin) / cs)), 0);
    int   rx1 = std::min(int(std::ceil((c.x + margin) / cs)), sampler.GridW());
    int   ry1 = std::min(int(std::ceil((c.y + margin) / cs)), sampler.GridH());
    float poissonMs = sampler.Resample(rx0, ry0, rx1, ry1);
    samplesOk       = samplesOk && sampler.Verify(rx0, ry0, rx1, ry1);
    opt.UpdateGenerators(slots(), rx0, ry0, rx1, ry1);
    opt.BeginSchedule(opt.Params().localIterations, rx0, ry0, rx1, ry1);
    float localMs = opt.Optimize();
    voromesh::VoroMesh2D mesh2 = opt.Extract();
    bool  ok2 = mesh2.watertight && mesh2.overflowCells == 0 && mesh2.loops >= 1;
    int   live = 0;
    for (Vec2 g : mesh2.generators) live += g.x >= 0;

    long long clamped = (long long)opt.GradSaturation();
    bool ok = ok1 && ok2 && samplesOk && loss1 < loss0 && opt.Iteration() == opt.ScheduleLength() && clamped == 0;
    std::printf("{\"ok\": %s, \"slots\": %d, \"live\": %d, \"loss0\": %.6e, \"loss1\": %.6e, \"global_ms\": %.3f, "
                "\"watertight\": [%d, %d], \"loops\": [%d, %d], \"unsafe\": [%d, %d], \"overflow\": [%d, %d], "
                "\"boundary_edges\": [%zu, %zu], \"poisson_ms\": %.3f, \"local_ms\": %.3f, \"extract_gpu_ms\": %.3f, "
                "\"extract_host_ms\": %.3f, \"iteration\": %d, \"grad_clamped\": %lld}\n",
                ok ? "true" : "false", opt.NumSlots(), live, loss0, loss1, globalMs, mesh.watertight,
                mesh2.watertight, mesh.loops, mesh2.loops, mesh.unsafeCells, mesh2.unsafeCells, mesh.overflowCells,
                mesh2.overflowCells, mesh.edges.size(), mesh2.edges.size(), poissonMs, localMs, mesh2.gpuMs,
                mesh2.hostMs, opt.Iteration(), clamped);
    return ok ? 0 : 1;
}
