// shot_demo.cpp -- headless run of the reference demo's scripted edit sequence
// (apps/voromesh_demo.cpp:526-570, `voromesh_demo --shot`) through the drop-in
// C++ API, without a window: the heart is rasterized on a canvas, Poisson
// generators are fitted to it (global optimisation) and extracted, then three
// brush drags run the whole local pipeline per stamp -- canvas edit, upload of
// the changed target slots and occupancy tiles, Poisson resampling of the
// generator slots around the disc, local re-optimisation, extraction -- with
// the demo's timing breakdown and mesh diagnostics (VoroLoss, watertightness,
// surface-to-mesh accuracy in voxels).  Exits 0 when the final mesh is
// watertight, like the demo.
//
// Usage: shot_demo [radius=24] [raster=1024] [localIterations=100] [--json]
//   --json  also print one JSON line with the figures (last line of stdout)
#include "poisson/poisson_disk.h"
#include "voromesh/canvas.h"
#include "voromesh/shape2d.h"
#include "voromesh/voro_optimizer.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace {

using voromesh::Vec2;
using Clock = std::chrono::steady_clock;

constexpr float kDomain = 1024.0f;

double msBetween(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct EditTiming {
    double canvas = 0, upload = 0, poisson = 0, optimize = 0, extract = 0, total = 0;
    float  poissonGpu = 0, optimizeGpu = 0, clipGpu = 0, assembleHost = 0;
    int    removed = 0, added = 0, slots = 0;
};

struct Region {
    int cx0 = 0, cy0 = 0, cx1 = 0, cy1 = 0;
};

// Canvas + Poisson sampler + optimizer and the local edit pipeline of the demo
// (apps/voromesh_demo.cpp:89-156).
class Editor
{
public:
    Editor(float radius, float voxel, const voromesh::OptimizerParams &params)
        : mRadius(radius), mCanvas(kDomain, kDomain, voxel), mOpt(kDomain, kDomain, radius, params)
    {
    }

    float Reset(unsigned seed)
    {
        mCanvas.RasterizeShape(voromesh::MakeHeart(0.5f * kDomain, 0.5f * kDomain, 0.8f * kDomain, 4096));
        mOpt.SetTargets(mCanvas.TargetSlots());
        mOpt.SetOccupancy(mCanvas.InsideMasks(), mCanvas.TilesX(), mCanvas.TilesY(), mCanvas.VoxelSize());
        mSampler = std::make_unique<poisson::Sampler>(kDomain, kDomain, mRadius, 30, seed);
        mOpt.SetGenerators(slots(), mSampler->GridW(), mSampler->GridH(), mSampler->CellSize());
        const float ms = mOpt.Optimize();
        mMesh          = mOpt.Extract();
        return ms;
    }

    bool Stamp(Vec2 c, float brush, bool add, EditTiming &t)
    {
        const auto                   t0 = Clock::now();
        const voromesh::Canvas::Edit e  = mCanvas.StampDisc(c, brush, add);
        const auto                   t1 = Clock::now();
        if (!e.Changed()) return false;

        mOpt.UpdateTargets(e.dirtySlots, mCanvas.TargetSlots());
        mOpt.UpdateOccupancy(mCanvas.InsideMasks(), e.tx0, e.ty0, e.tx1, e.ty1);
        const auto t2 = Clock::now();

        // Generators within two Poisson radii of the disc are replaced.
        const float margin = brush + 2.0f * mRadius, cs = mSampler->CellSize();
        Region      r;
        r.cx0 = std::max(int(std::floor((c.x - margin) / cs)), 0);
        r.cy0 = std::max(int(std::floor((c.y - margin) / cs)), 0);
        r.cx1 = std::min(int(std::ceil((c.x + margin) / cs)), mSampler->GridW());
        r.cy1 = std::min(int(std::ceil((c.y + margin) / cs)), mSampler->GridH());
        t.poissonGpu = mSampler->Resample(r.cx0, r.cy0, r.cx1, r.cy1);
        mOpt.UpdateGenerators(slots(), r.cx0, r.cy0, r.cx1, r.cy1);
        const auto t3 = Clock::now();

        mOpt.BeginSchedule(mOpt.Params().localIterations, r.cx0, r.cy0, r.cx1, r.cy1);
        t.optimizeGpu = mOpt.Optimize();
        const auto t4 = Clock::now();

        mMesh         = mOpt.Extract();
        const auto t5 = Clock::now();

        t.clipGpu      = mMesh.gpuMs;
        t.assembleHost = mMesh.hostMs;
        t.canvas   = msBetween(t0, t1);
        t.upload   = msBetween(t1, t2);
        t.poisson  = msBetween(t2, t3);
        t.optimize = msBetween(t3, t4);
        t.extract  = msBetween(t4, t5);
        t.total    = msBetween(t0, t5);
        t.removed  = e.removed;
        t.added    = e.added;
        t.slots    = (r.cx1 - r.cx0) * (r.cy1 - r.cy0);
        return true;
    }

    std::vector<Vec2> SurfacePoints() const
    {
        std::vector<Vec2> pts;
        for (Vec2 p : mCanvas.TargetSlots())
            if (p.x >= 0.0f) pts.push_back(p);
        return pts;
    }

    const voromesh::Canvas     &Canvas() const { return mCanvas; }
    const voromesh::VoroMesh2D &Mesh() const { return mMesh; }
    voromesh::VoroOptimizer    &Optimizer() { return mOpt; }
    const poisson::Sampler     &Sampler() const { return *mSampler; }

private:
    const Vec2 *slots() const { return reinterpret_cast<const Vec2 *>(mSampler->Cells().data()); }

    float                             mRadius;
    voromesh::Canvas                  mCanvas;
    voromesh::VoroOptimizer           mOpt;
    std::unique_ptr<poisson::Sampler> mSampler;
    voromesh::VoroMesh2D              mMesh;
};

float segmentDistance(Vec2 p, Vec2 a, Vec2 b)
{
    const float ex = b.x - a.x, ey = b.y - a.y, l2 = ex * ex + ey * ey;
    const float t = l2 > 0.0f ? std::clamp(((p.x - a.x) * ex + (p.y - a.y) * ey) / l2, 0.0f, 1.0f) : 0.0f;
    return std::hypot(p.x - a.x - t * ex, p.y - a.y - t * ey);
}

// Boundary accuracy against the canvas surface (apps/voromesh_demo.cpp:183-209):
// surface point -> nearest boundary edge, boundary vertex -> nearest point.
struct MeshReport {
    double loss = 0, rmsVoxels = 0, meanToMesh = 0, maxToMesh = 0, meanToSurface = 0, maxToSurface = 0;
    int    live = 0, inside = 0, vertices = 0, edges = 0, loops = 0, unsafe = 0, overflow = 0;
    bool   watertight = false;
};

MeshReport report(Editor &ed)
{
    const voromesh::VoroMesh2D &m     = ed.Mesh();
    const float                 voxel = ed.Canvas().VoxelSize();
    const std::vector<Vec2>     pts   = ed.SurfacePoints();
    MeshReport                  r;
    r.loss      = ed.Optimizer().EvaluateLoss();
    r.rmsVoxels = std::sqrt(r.loss) / voxel;
    for (size_t i = 0; i < m.generators.size(); ++i) r.live += m.generators[i].x >= 0.0f, r.inside += m.inside[i];
    r.vertices = int(m.vertices.size()), r.edges = int(m.edges.size()), r.loops = m.loops;
    r.unsafe = m.unsafeCells, r.overflow = m.overflowCells, r.watertight = m.watertight;
    if (!m.edges.empty() && !pts.empty()) {
        for (Vec2 p : pts) {
            float best = 1e30f;
            for (const auto &e : m.edges) best = std::min(best, segmentDistance(p, m.vertices[e[0]], m.vertices[e[1]]));
            r.meanToMesh += best, r.maxToMesh = std::max(r.maxToMesh, double(best));
        }
        r.meanToMesh /= double(pts.size());
        for (Vec2 v : m.vertices) {
            float best = 1e30f;
            for (Vec2 p : pts) best = std::min(best, std::hypot(p.x - v.x, p.y - v.y));
            r.meanToSurface += best, r.maxToSurface = std::max(r.maxToSurface, double(best));
        }
        r.meanToSurface /= double(m.vertices.size());
    }
    r.meanToMesh /= voxel, r.maxToMesh /= voxel, r.meanToSurface /= voxel, r.maxToSurface /= voxel;
    return r;
}

void printReport(const char *label, Editor &ed, const MeshReport &r)
{
    std::printf("[%s]\n", label);
    std::printf("  canvas                   %zu surface points in %zu tiles\n", ed.Canvas().SurfaceCount(),
                ed.Canvas().SurfaceTileCount());
    std::printf("  VoroLoss (all targets)   %.4f  (RMS %.3f voxels)\n", r.loss, r.rmsVoxels);
    std::printf("  cells                    %d live of %zu slots (%d inside), unsafe %d, overflow %d\n", r.live,
                ed.Mesh().generators.size(), r.inside, r.unsafe, r.overflow);
    std::printf("  boundary                 %d vertices, %d edges, %d loop(s), %s\n", r.vertices, r.edges, r.loops,
                r.watertight ? "WATERTIGHT" : "NOT WATERTIGHT");
    std::printf("  surface -> mesh          mean %.3f  max %.3f voxels\n", r.meanToMesh, r.maxToMesh);
    std::printf("  mesh vertex -> surface   mean %.3f  max %.3f voxels\n", r.meanToSurface, r.maxToSurface);
}

struct Stat {
    double mean = 0, median = 0, p95 = 0, max = 0;
};

Stat stat(const std::vector<EditTiming> &v, double EditTiming::*f)
{
    std::vector<double> x;
    for (const EditTiming &e : v) x.push_back(e.*f);
    Stat s;
    if (x.empty()) return s;
    std::sort(x.begin(), x.end());
    for (double y : x) s.mean += y;
    s.mean /= double(x.size());
    s.median = x[x.size() / 2];
    s.p95    = x[std::min(x.size() - 1, size_t(double(x.size()) * 0.95))];
    s.max    = x.back();
    return s;
}

std::string jsonStat(const char *name, const Stat &s)
{
    char b[256];
    std::snprintf(b, sizeof(b), "\"%s\": {\"mean\": %.4f, \"median\": %.4f, \"p95\": %.4f, \"max\": %.4f}", name, s.mean,
                  s.median, s.p95, s.max);
    return b;
}

std::string jsonReport(const MeshReport &r)
{
    char b[512];
    std::snprintf(b, sizeof(b),
                  "{\"loss\": %.6g, \"rms_voxels\": %.4f, \"live\": %d, \"inside\": %d, \"vertices\": %d, \"edges\": "
                  "%d, \"loops\": %d, \"watertight\": %s, \"unsafe\": %d, \"overflow\": %d, \"surface_to_mesh_mean\": "
                  "%.4f, \"surface_to_mesh_max\": %.4f, \"vertex_to_surface_mean\": %.4f, \"vertex_to_surface_max\": "
                  "%.4f}",
                  r.loss, r.rmsVoxels, r.live, r.inside, r.vertices, r.edges, r.loops, r.watertight ? "true" : "false",
                  r.unsafe, r.overflow, r.meanToMesh, r.maxToMesh, r.meanToSurface, r.maxToSurface);
    return b;
}

} // namespace

int main(int argc, char **argv)
{
    std::vector<const char *> pos;
    bool                      json = false;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--json")) json = true;
        else pos.push_back(argv[i]);
    }
    const float               radius = pos.size() > 0 ? float(std::atof(pos[0])) : 24.0f;
    const int                 raster = pos.size() > 1 ? std::atoi(pos[1]) : 1024;
    voromesh::OptimizerParams params;
    params.localIterations = pos.size() > 2 ? std::atoi(pos[2]) : 100;
    const float voxel      = kDomain / float(raster);

    Editor     ed(radius, voxel, params);
    const auto r0       = Clock::now();
    const float globalMs = ed.Reset(1);
    const double resetWall = msBetween(r0, Clock::now());
    std::printf("Canvas %d x %d (voxel %.3f), Poisson r = %.1f -> %d x %d generator slots\n", raster, raster, voxel,
                radius, ed.Sampler().GridW(), ed.Sampler().GridH());
    std::printf("Reset + global optimisation (%d it): %.3f ms GPU, %.3f ms wall\n", params.iterations, globalMs,
                resetWall);
    const MeshReport before = report(ed);
    printReport("heart", ed, before);

    // The demo's drags: a bite into the right lobe, a hole in the middle, a bump
    // on the left; stamps every quarter radius along each drag.
    struct Drag {
        Vec2  from, to;
        float r;
        bool  add;
    };
    const Drag drags[] = {
        {{640, 120}, {830, 230}, 48, false},
        {{512, 560}, {512, 560}, 70, false},
        {{120, 300}, {80, 440}, 40, true},
    };
    std::vector<EditTiming> edits;
    std::printf("Scripted edits (%d local iterations each):\n", params.localIterations);
    for (const Drag &d : drags) {
        const float len = std::hypot(d.to.x - d.from.x, d.to.y - d.from.y);
        const int   n   = std::max(1, int(std::ceil(len / (0.25f * d.r))));
        for (int k = 0; k <= n; ++k) {
            const float t = float(k) / float(n);
            const Vec2  c{d.from.x + t * (d.to.x - d.from.x), d.from.y + t * (d.to.y - d.from.y)};
            EditTiming  et;
            if (ed.Stamp(c, d.r, d.add, et)) {
                edits.push_back(et);
                std::printf("  %-5s (%4.0f,%4.0f): -%3d +%3d pts, %3d slots | canvas %.3f  upload %.3f  poisson %.3f "
                            "(%.3f GPU)  opt %.3f (%.3f GPU, %d it)  extract %.3f (clip %.3f GPU, host %.3f) | total "
                            "%.3f ms\n",
                            d.add ? "add" : "carve", c.x, c.y, et.removed, et.added, et.slots, et.canvas, et.upload,
                            et.poisson, double(et.poissonGpu), et.optimize, double(et.optimizeGpu),
                            params.localIterations, et.extract, double(et.clipGpu), double(et.assembleHost), et.total);
            }
            if (len == 0.0f) break;
        }
    }
    const Stat sCanvas = stat(edits, &EditTiming::canvas), sUpload = stat(edits, &EditTiming::upload);
    const Stat sPoisson = stat(edits, &EditTiming::poisson), sOpt = stat(edits, &EditTiming::optimize);
    const Stat sExtract = stat(edits, &EditTiming::extract), sTotal = stat(edits, &EditTiming::total);
    std::printf("Edit timings over %zu stamps (wall clock, incl. sync):\n", edits.size());
    const std::pair<const char *, Stat> rows[] = {{"canvas stamp (CPU)", sCanvas}, {"upload targets+occ", sUpload},
                                                   {"poisson resample", sPoisson},  {"local optimisation", sOpt},
                                                   {"extract", sExtract},           {"TOTAL", sTotal}};
    for (const auto &row : rows)
        std::printf("  %-22s mean %7.3f  median %7.3f  p95 %7.3f  max %7.3f ms\n", row.first, row.second.mean,
                    row.second.median, row.second.p95, row.second.max);
    const MeshReport after = report(ed);
    printReport("after edits", ed, after);

    if (json) {
        double optGpu = 0;
        for (const EditTiming &e : edits) optGpu += e.optimizeGpu;
        std::printf("{\"radius\": %.3f, \"raster\": %d, \"local_iterations\": %d, \"generator_slots\": %d, "
                    "\"global_gpu_ms\": %.4f, \"reset_wall_ms\": %.4f, \"stamps\": %zu, \"local_opt_gpu_ms_mean\": "
                    "%.4f, %s, %s, %s, %s, %s, %s, \"heart\": %s, \"after\": %s}\n",
                    double(radius), raster, params.localIterations, ed.Sampler().GridW() * ed.Sampler().GridH(),
                    double(globalMs), resetWall, edits.size(), edits.empty() ? 0.0 : optGpu / double(edits.size()),
                    jsonStat("canvas_ms", sCanvas).c_str(), jsonStat("upload_ms", sUpload).c_str(),
                    jsonStat("poisson_ms", sPoisson).c_str(), jsonStat("optimize_ms", sOpt).c_str(),
                    jsonStat("extract_ms", sExtract).c_str(), jsonStat("total_ms", sTotal).c_str(),
                    jsonReport(before).c_str(), jsonReport(after).c_str());
    }
    std::fflush(stdout);
    return ed.Mesh().watertight ? 0 : 1;
}
