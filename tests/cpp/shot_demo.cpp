// shot_demo.cpp -- headless benchmark of the local edit pipeline, through the
// drop-in C++ API only (no window, no GL).
//
// The session follows the reference demo's `--shot` run (apps/voromesh_demo.cpp,
// Editor::Reset / Editor::Stamp and the scripted drags): the heart shape is
// rasterized on a voxel canvas, Poisson-disk generators are fitted to its
// surface samples and the mesh is extracted; then a script of brush strokes is
// replayed.  Every stamp runs the same calls in the same order as the demo:
//   Canvas::StampDisc -> VoroOptimizer::UpdateTargets + UpdateOccupancy
//   -> Sampler::Resample (slots within brush + 2 sampling radii of the stamp)
//   -> VoroOptimizer::UpdateGenerators -> BeginSchedule(localIterations, box)
//   -> Optimize -> Extract
// and is timed stage by stage.  The mesh figures use the demo's definitions:
// VoroLoss over all targets (and its RMS in voxels), live / inside cells,
// boundary vertices, edges and loops, watertightness, the mean and maximum
// distance from each surface sample to the nearest boundary edge and from
// each boundary vertex to the nearest surface sample (in voxels).
//
// Usage: shot_demo [radius=24] [raster=1024] [localIterations=100] [--json]
//   --json  the last line of stdout is one JSON object with the figures.
// Exit status 0 when the final mesh is watertight.
#include "poisson/poisson_disk.h"
#include "voromesh/canvas.h"
#include "voromesh/shape2d.h"
#include "voromesh/voro_optimizer.h"

#include <algorithm>
#include <array>
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

constexpr float kWorld = 1024.0f; // the session's square domain (world units = voxels at raster 1024)

// ---- stage timing ---------------------------------------------------------
enum Stage { kCanvasStage, kUploadStage, kResampleStage, kOptimizeStage, kExtractStage, kNumStages };
constexpr const char *kStageKey[kNumStages] = {"canvas", "upload", "poisson", "optimize", "extract"};
constexpr const char *kStageLabel[kNumStages] = {"canvas stamp (host)", "upload targets+occupancy",
                                                 "poisson resample", "local optimisation", "extract"};

// Consecutive laps of one stamp, in milliseconds.
class Laps
{
public:
    Laps() : mStart(Clock::now()), mLast(mStart) {}
    void close(Stage s)
    {
        const Clock::time_point now = Clock::now();
        mMs[s] = std::chrono::duration<double, std::milli>(now - mLast).count();
        mLast  = now;
    }
    double stage(Stage s) const { return mMs[s]; }
    double total() const { return std::chrono::duration<double, std::milli>(mLast - mStart).count(); }

private:
    Clock::time_point                mStart, mLast;
    std::array<double, kNumStages>   mMs{};
};

struct StampRecord {
    Vec2   center{};
    float  brush = 0;
    bool   add   = false;
    Laps   laps;
    float  resampleGpu = 0, optimizeGpu = 0, clipGpu = 0, assembleHost = 0;
    int    removed = 0, added = 0, slots = 0;
};

// ---- the script -----------------------------------------------------------
// Brush strokes as data (world units): a bite dragged into the right lobe, a
// hole punched in the middle (a click: start == end), a bump dragged out on
// the left.  A stroke stamps every quarter brush radius along its length.
struct Stroke {
    Vec2  from, to;
    float brush;
    bool  add;
};
constexpr Stroke kScript[] = {
    {{640.0f, 120.0f}, {830.0f, 230.0f}, 48.0f, false},
    {{512.0f, 560.0f}, {512.0f, 560.0f}, 70.0f, false},
    {{120.0f, 300.0f}, {80.0f, 440.0f}, 40.0f, true},
};

std::vector<Vec2> stampCentres(const Stroke &s)
{
    const float dx = s.to.x - s.from.x, dy = s.to.y - s.from.y;
    const float length = std::hypot(dx, dy);
    if (length == 0.0f) return {s.from};
    const int         steps = std::max(1, int(std::ceil(length / (0.25f * s.brush))));
    std::vector<Vec2> out;
    out.reserve(size_t(steps) + 1);
    for (int k = 0; k <= steps; ++k) {
        const float t = float(k) / float(steps);
        out.push_back(Vec2{s.from.x + t * dx, s.from.y + t * dy});
    }
    return out;
}

// Generator slots [x0, x1) x [y0, y1) of the Poisson cells within `reach` of c.
struct SlotBox {
    int x0, y0, x1, y1;
    int area() const { return (x1 - x0) * (y1 - y0); }
};
SlotBox slotsWithin(Vec2 c, float reach, float cell, int gw, int gh)
{
    auto below = [cell](float v) { return std::max(int(std::floor(v / cell)), 0); };
    auto above = [cell](float v, int n) { return std::min(int(std::ceil(v / cell)), n); };
    return SlotBox{below(c.x - reach), below(c.y - reach), above(c.x + reach, gw), above(c.y + reach, gh)};
}

// ---- the session ----------------------------------------------------------
class Session
{
public:
    Session(float sampleRadius, float voxel, const voromesh::OptimizerParams &prm)
        : mSampleRadius(sampleRadius), mCanvas(kWorld, kWorld, voxel), mOpt(kWorld, kWorld, sampleRadius, prm)
    {
    }

    // Heart outline -> targets and occupancy; fresh generators; global fit; mesh.
    float fitHeart(unsigned seed)
    {
        mCanvas.RasterizeShape(voromesh::MakeHeart(0.5f * kWorld, 0.5f * kWorld, 0.8f * kWorld, 4096));
        mOpt.SetTargets(mCanvas.TargetSlots());
        mOpt.SetOccupancy(mCanvas.InsideMasks(), mCanvas.TilesX(), mCanvas.TilesY(), mCanvas.VoxelSize());
        mSampler = std::make_unique<poisson::Sampler>(kWorld, kWorld, mSampleRadius, 30, seed);
        mOpt.SetGenerators(generatorSlots(), mSampler->GridW(), mSampler->GridH(), mSampler->CellSize());
        const float gpuMs = mOpt.Optimize();
        mMesh             = mOpt.Extract();
        return gpuMs;
    }

    // One stamp through the local pipeline; false when the canvas did not change.
    bool stamp(Vec2 c, float brush, bool add, StampRecord &rec)
    {
        rec       = StampRecord{};
        rec.center = c, rec.brush = brush, rec.add = add;
        const voromesh::Canvas::Edit edit = mCanvas.StampDisc(c, brush, add);
        rec.laps.close(kCanvasStage);
        if (!edit.Changed()) return false;
        rec.removed = edit.removed, rec.added = edit.added;

        mOpt.UpdateTargets(edit.dirtySlots, mCanvas.TargetSlots());
        mOpt.UpdateOccupancy(mCanvas.InsideMasks(), edit.tx0, edit.ty0, edit.tx1, edit.ty1);
        rec.laps.close(kUploadStage);

        const SlotBox box = slotsWithin(c, brush + 2.0f * mSampleRadius, mSampler->CellSize(), mSampler->GridW(),
                                        mSampler->GridH());
        rec.slots         = box.area();
        rec.resampleGpu   = mSampler->Resample(box.x0, box.y0, box.x1, box.y1);
        mOpt.UpdateGenerators(generatorSlots(), box.x0, box.y0, box.x1, box.y1);
        rec.laps.close(kResampleStage);

        mOpt.BeginSchedule(mOpt.Params().localIterations, box.x0, box.y0, box.x1, box.y1);
        rec.optimizeGpu = mOpt.Optimize();
        rec.laps.close(kOptimizeStage);

        mMesh = mOpt.Extract();
        rec.laps.close(kExtractStage);
        rec.clipGpu = mMesh.gpuMs, rec.assembleHost = mMesh.hostMs;
        return true;
    }

    const voromesh::Canvas     &canvas() const { return mCanvas; }
    const voromesh::VoroMesh2D &mesh() const { return mMesh; }
    voromesh::VoroOptimizer    &optimizer() { return mOpt; }
    const poisson::Sampler     &sampler() const { return *mSampler; }

private:
    const Vec2 *generatorSlots() const { return reinterpret_cast<const Vec2 *>(mSampler->Cells().data()); }

    float                             mSampleRadius;
    voromesh::Canvas                  mCanvas;
    voromesh::VoroOptimizer           mOpt;
    std::unique_ptr<poisson::Sampler> mSampler;
    voromesh::VoroMesh2D              mMesh;
};

// ---- mesh figures ---------------------------------------------------------
// Buckets of a square grid over the world, for nearest-primitive queries by
// growing rings of buckets (exact: a ring is searched only while it can hold
// something nearer than the best so far).
class Buckets
{
public:
    explicit Buckets(float size) : mSize(size), mN(std::max(1, int(std::ceil(kWorld / size)))), mItems(size_t(mN) * mN) {}
    int  clampIndex(float v) const { return std::min(std::max(int(std::floor(v / mSize)), 0), mN - 1); }
    void insertBox(int item, float x0, float y0, float x1, float y1)
    {
        for (int j = clampIndex(y0); j <= clampIndex(y1); ++j)
            for (int i = clampIndex(x0); i <= clampIndex(x1); ++i) mItems[size_t(j) * mN + i].push_back(item);
    }
    // Smallest dist(item) over all items, given that every item within distance d
    // of p lies in the buckets a ring search of radius d reaches.
    template <class Dist> float nearest(Vec2 p, Dist dist) const
    {
        const int cx = clampIndex(p.x), cy = clampIndex(p.y);
        float     best = INFINITY;
        for (int ring = 0; ring < mN; ++ring) {
            for (int j = cy - ring; j <= cy + ring; ++j) {
                if (j < 0 || j >= mN) continue;
                const bool edgeRow = j == cy - ring || j == cy + ring;
                for (int i = cx - ring; i <= cx + ring; i += edgeRow ? 1 : 2 * ring) {
                    if (i >= 0 && i < mN)
                        for (int item : mItems[size_t(j) * mN + i]) best = std::min(best, dist(item));
                    if (ring == 0) break;
                }
            }
            // Everything outside the searched square is farther than its inner margin.
            const float margin = std::min(std::min(p.x - float(cx - ring) * mSize, float(cx + ring + 1) * mSize - p.x),
                                          std::min(p.y - float(cy - ring) * mSize, float(cy + ring + 1) * mSize - p.y));
            if (best <= margin) break;
        }
        return best;
    }

private:
    float                         mSize;
    int                           mN;
    std::vector<std::vector<int>> mItems;
};

float distanceToSegment(Vec2 p, Vec2 a, Vec2 b)
{
    const float ex = b.x - a.x, ey = b.y - a.y, l2 = ex * ex + ey * ey;
    const float t = l2 > 0.0f ? std::clamp(((p.x - a.x) * ex + (p.y - a.y) * ey) / l2, 0.0f, 1.0f) : 0.0f;
    return std::hypot(p.x - a.x - t * ex, p.y - a.y - t * ey);
}

struct MeshFigures {
    double loss = 0, rmsVoxels = 0;
    int    live = 0, slots = 0, inside = 0, vertices = 0, edges = 0, loops = 0, unsafe = 0, overflow = 0;
    bool   watertight = false;
    double toMeshMean = 0, toMeshMax = 0, toSurfaceMean = 0, toSurfaceMax = 0; // voxels
};

MeshFigures measure(Session &s)
{
    const voromesh::VoroMesh2D &m     = s.mesh();
    const float                 voxel = s.canvas().VoxelSize();
    MeshFigures                 f;
    f.loss      = s.optimizer().EvaluateLoss();
    f.rmsVoxels = std::sqrt(f.loss) / voxel;
    f.slots     = int(m.generators.size());
    for (size_t i = 0; i < m.generators.size(); ++i) f.live += m.generators[i].x >= 0.0f ? 1 : 0;
    for (uint8_t v : m.inside) f.inside += v;
    f.vertices = int(m.vertices.size()), f.edges = int(m.edges.size()), f.loops = m.loops;
    f.unsafe = m.unsafeCells, f.overflow = m.overflowCells, f.watertight = m.watertight;

    std::vector<Vec2> surface;
    for (Vec2 p : s.canvas().TargetSlots())
        if (p.x >= 0.0f) surface.push_back(p);
    if (m.edges.empty() || surface.empty()) return f;
    const float bucket = 16.0f * voxel;
    Buckets     edgeBuckets(bucket), pointBuckets(bucket);
    for (size_t e = 0; e < m.edges.size(); ++e) {
        const Vec2 a = m.vertices[size_t(m.edges[e][0])], b = m.vertices[size_t(m.edges[e][1])];
        edgeBuckets.insertBox(int(e), std::min(a.x, b.x), std::min(a.y, b.y), std::max(a.x, b.x), std::max(a.y, b.y));
    }
    for (size_t k = 0; k < surface.size(); ++k)
        pointBuckets.insertBox(int(k), surface[k].x, surface[k].y, surface[k].x, surface[k].y);
    for (Vec2 p : surface) {
        const float d = edgeBuckets.nearest(p, [&](int e) {
            return distanceToSegment(p, m.vertices[size_t(m.edges[size_t(e)][0])], m.vertices[size_t(m.edges[size_t(e)][1])]);
        });
        f.toMeshMean += d, f.toMeshMax = std::max(f.toMeshMax, double(d));
    }
    for (Vec2 v : m.vertices) {
        const float d = pointBuckets.nearest(v, [&](int k) { return std::hypot(surface[size_t(k)].x - v.x, surface[size_t(k)].y - v.y); });
        f.toSurfaceMean += d, f.toSurfaceMax = std::max(f.toSurfaceMax, double(d));
    }
    f.toMeshMean /= double(surface.size()) * voxel, f.toMeshMax /= voxel;
    f.toSurfaceMean /= double(m.vertices.size()) * voxel, f.toSurfaceMax /= voxel;
    return f;
}

void showFigures(const char *tag, const Session &s, const MeshFigures &f)
{
    std::printf("mesh %-6s loss %.5g (rms %.3f vx) | cells %d/%d live, %d inside, unsafe %d, overflow %d | boundary "
                "%dv %de %d loop%s %s | surface->mesh %.3f/%.3f vx, vertex->surface %.3f/%.3f vx (mean/max) | "
                "canvas %zu samples in %zu tiles\n",
                tag, f.loss, f.rmsVoxels, f.live, f.slots, f.inside, f.unsafe, f.overflow, f.vertices, f.edges, f.loops,
                f.loops == 1 ? "" : "s", f.watertight ? "watertight" : "OPEN", f.toMeshMean, f.toMeshMax,
                f.toSurfaceMean, f.toSurfaceMax, s.canvas().SurfaceCount(), s.canvas().SurfaceTileCount());
}

// ---- summaries and JSON ---------------------------------------------------
struct Summary {
    double mean = 0, median = 0, p95 = 0, max = 0;
};
Summary summarize(std::vector<double> v)
{
    Summary s;
    if (v.empty()) return s;
    std::sort(v.begin(), v.end());
    double sum = 0;
    for (double x : v) sum += x;
    s.mean   = sum / double(v.size());
    s.median = v[v.size() / 2];
    s.p95    = v[std::min(v.size() - 1, size_t(0.95 * double(v.size())))];
    s.max    = v.back();
    return s;
}

// Minimal JSON object writer (numbers, booleans and nested objects).
class Json
{
public:
    Json &num(const char *k, double v, int digits = 4)
    {
        char b[64];
        std::snprintf(b, sizeof(b), "%.*f", digits, v);
        return raw(k, b);
    }
    Json &integer(const char *k, long long v) { return raw(k, std::to_string(v)); }
    Json &boolean(const char *k, bool v) { return raw(k, v ? "true" : "false"); }
    Json &object(const char *k, const Json &o) { return raw(k, o.str()); }
    std::string str() const { return "{" + mBody + "}"; }

private:
    Json &raw(const char *k, const std::string &v)
    {
        if (!mBody.empty()) mBody += ", ";
        mBody += std::string("\"") + k + "\": " + v;
        return *this;
    }
    std::string mBody;
};

Json summaryJson(const Summary &s)
{
    return Json().num("mean", s.mean).num("median", s.median).num("p95", s.p95).num("max", s.max);
}

Json figuresJson(const MeshFigures &f)
{
    Json j;
    j.num("loss", f.loss, 6).num("rms_voxels", f.rmsVoxels).integer("live", f.live).integer("inside", f.inside);
    j.integer("vertices", f.vertices).integer("edges", f.edges).integer("loops", f.loops).boolean("watertight", f.watertight);
    j.integer("unsafe", f.unsafe).integer("overflow", f.overflow);
    j.num("surface_to_mesh_mean", f.toMeshMean).num("surface_to_mesh_max", f.toMeshMax);
    j.num("vertex_to_surface_mean", f.toSurfaceMean).num("vertex_to_surface_max", f.toSurfaceMax);
    return j;
}

} // namespace

int main(int argc, char **argv)
{
    bool                      json = false;
    std::vector<const char *> args;
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "--json") == 0) json = true;
        else args.push_back(argv[i]);
    }
    auto arg = [&](size_t k, double fallback) { return k < args.size() ? std::atof(args[k]) : fallback; };
    const float               sampleRadius = float(arg(0, 24.0));
    const int                 raster       = int(arg(1, 1024.0));
    voromesh::OptimizerParams prm;
    prm.localIterations = int(arg(2, 100.0));

    Session           session(sampleRadius, kWorld / float(raster), prm);
    const Clock::time_point t0 = Clock::now();
    const float       fitGpu   = session.fitHeart(1);
    const double      fitWall  = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    const int         gw = session.sampler().GridW(), gh = session.sampler().GridH();
    std::printf("shot: raster %d^2 (voxel %.3f), sampling radius %.1f, %d x %d generator slots\n", raster,
                double(session.canvas().VoxelSize()), double(sampleRadius), gw, gh);
    std::printf("fit:  %d iterations in %.3f ms GPU (%.3f ms wall with set-up and extract)\n", prm.iterations,
                double(fitGpu), fitWall);
    const MeshFigures heart = measure(session);
    showFigures("heart", session, heart);

    std::vector<StampRecord> records;
    for (const Stroke &s : kScript)
        for (Vec2 c : stampCentres(s)) {
            StampRecord rec;
            if (session.stamp(c, s.brush, s.add, rec)) records.push_back(rec);
        }

    std::printf("stamps (%d local iterations each), ms:\n", prm.localIterations);
    std::printf("  %3s %-5s %7s %7s %4s %5s %5s %5s | %7s %7s %7s %7s %7s %7s | %8s %8s %8s %8s\n", "#", "op", "x", "y",
                "r", "-pts", "+pts", "slots", kStageKey[0], kStageKey[1], kStageKey[2], kStageKey[3], kStageKey[4],
                "total", "poi-gpu", "opt-gpu", "clip-gpu", "host");
    for (size_t k = 0; k < records.size(); ++k) {
        const StampRecord &r = records[k];
        std::printf("  %3zu %-5s %7.1f %7.1f %4.0f %5d %5d %5d | %7.3f %7.3f %7.3f %7.3f %7.3f %7.3f | %8.3f %8.3f %8.3f "
                    "%8.3f\n",
                    k + 1, r.add ? "add" : "carve", double(r.center.x), double(r.center.y), double(r.brush), r.removed,
                    r.added, r.slots, r.laps.stage(kCanvasStage), r.laps.stage(kUploadStage),
                    r.laps.stage(kResampleStage), r.laps.stage(kOptimizeStage), r.laps.stage(kExtractStage),
                    r.laps.total(), double(r.resampleGpu), double(r.optimizeGpu), double(r.clipGpu),
                    double(r.assembleHost));
    }
    std::array<Summary, kNumStages> stageSummary;
    std::vector<double>             totals, optGpu;
    for (const StampRecord &r : records) totals.push_back(r.laps.total()), optGpu.push_back(r.optimizeGpu);
    std::printf("per-stamp wall clock over %zu stamps (mean / median / p95 / max ms):\n", records.size());
    for (int s = 0; s < kNumStages; ++s) {
        std::vector<double> v;
        for (const StampRecord &r : records) v.push_back(r.laps.stage(Stage(s)));
        stageSummary[size_t(s)] = summarize(v);
        const Summary &x        = stageSummary[size_t(s)];
        std::printf("  %-26s %8.3f %8.3f %8.3f %8.3f\n", kStageLabel[s], x.mean, x.median, x.p95, x.max);
    }
    const Summary totalSummary = summarize(totals);
    std::printf("  %-26s %8.3f %8.3f %8.3f %8.3f\n", "whole stamp", totalSummary.mean, totalSummary.median,
                totalSummary.p95, totalSummary.max);
    const MeshFigures after = measure(session);
    showFigures("edited", session, after);

    if (json) {
        Json j;
        j.num("radius", sampleRadius, 3).integer("raster", raster).integer("local_iterations", prm.localIterations);
        j.integer("generator_slots", gw * gh).num("global_gpu_ms", fitGpu).num("reset_wall_ms", fitWall);
        j.integer("stamps", (long long)records.size()).num("local_opt_gpu_ms_mean", summarize(optGpu).mean);
        for (int s = 0; s < kNumStages; ++s)
            j.object((std::string(kStageKey[s]) + "_ms").c_str(), summaryJson(stageSummary[size_t(s)]));
        j.object("total_ms", summaryJson(totalSummary));
        j.object("heart", figuresJson(heart)).object("after", figuresJson(after));
        std::printf("%s\n", j.str().c_str());
    }
    std::fflush(stdout);
    return session.mesh().watertight ? 0 : 1;
}
