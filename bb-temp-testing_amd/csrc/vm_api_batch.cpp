// vm_api_batch.cpp -- C ABI: Poisson sampler (poisson::Sampler) and its batched
// form, and batches of independent fits.
#include "voromesh_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "vm_engine.h"

using vmrt::DBuf;
using vmrt::HBuf;
using vmrt::VmError;

namespace {

#define VM_REQUIRE(cond, msg)                                                                                      \
    do {                                                                                                           \
        if (!(cond)) throw VmError(VM_EINVAL, msg);                                                                \
    } while (0)

void poissonDims(float W, float H, float r, int *gw, int *gh, float *cs)
{
    // poisson_disk.cu:119-124: cell r/sqrt(2), dims ceil(W/cell) (FP32 as the reference).
    float c = r / sqrtf(2.0f);
    *cs     = c;
    *gw     = int(ceilf(W / c));
    *gh     = int(ceilf(H / c));
}

// Runs the dart thrower over a table of instance regions on `stream`.
struct PoissonRunner {
    DBuf<vmk::PoiInst> dInst;
    DBuf<vmk::PoiWork> dClear, dPhase;

    void run(hipStream_t stream, const std::vector<vmk::PoiInst> &insts, float *dGrid, int trials)
    {
        std::vector<vmk::PoiWork> clear, phase;
        for (int i = 0; i < int(insts.size()); ++i) {
            const vmk::PoiInst &p = insts[i];
            int rw = p.x1 - p.x0, rh = p.y1 - p.y0;
            if (rw <= 0 || rh <= 0) continue;
            for (int b = 0; b < rw * rh; b += 256) clear.push_back({i, b});
            int np = ((rw + 2) / 3) * ((rh + 2) / 3);
            for (int b = 0; b < np; b += 256) phase.push_back({i, b});
        }
        if (clear.empty()) return;
        dInst.reserve(insts.size()), dClear.reserve(clear.size()), dPhase.reserve(phase.size());
        VM_HIP(hipMemcpyAsync(dInst.p, insts.data(), insts.size() * sizeof(vmk::PoiInst), hipMemcpyHostToDevice, stream));
        VM_HIP(hipMemcpyAsync(dClear.p, clear.data(), clear.size() * sizeof(vmk::PoiWork), hipMemcpyHostToDevice, stream));
        VM_HIP(hipMemcpyAsync(dPhase.p, phase.data(), phase.size() * sizeof(vmk::PoiWork), hipMemcpyHostToDevice, stream));
        int gc = int(std::min<size_t>(clear.size(), 8192)), gp = int(std::min<size_t>(phase.size(), 8192));
        VM_HIP(vmk::launchPoissonClear(gc, stream, dInst.p, dClear.p, int(clear.size()), dGrid));
        for (int round = 0; round < trials; ++round)
            for (int py = 0; py < 3; ++py)
                for (int px = 0; px < 3; ++px)
                    VM_HIP(vmk::launchPoissonPhase(gp, stream, dInst.p, dPhase.p, int(phase.size()), dGrid, px, py, round));
        // Keep the host tables alive until the copies have been consumed.
        VM_HIP(hipStreamSynchronize(stream));
    }
};

} // namespace

// ===========================================================================
// poisson::Sampler
// ===========================================================================
struct vm_sampler {
    float                W = 1, H = 1, r = 0.1f, cs = 0.1f;
    int                  gw = 0, gh = 0, trials = 30;
    unsigned             seed = 1, roundBase = 0;
    hipStream_t          stream = nullptr;
    hipEvent_t           e0 = nullptr, e1 = nullptr;
    DBuf<float>          dGrid;
    HBuf<vm_vec2>        hRows; // pinned staging of resampled grid rows
    std::vector<vm_vec2> cells;
    PoissonRunner        runner;
    ~vm_sampler()
    {
        dGrid.release();
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

int vm_poisson_grid_dims(float width, float height, float radius, int *gw, int *gh, float *cs)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(radius > 0 && width > 0 && height > 0 && gw && gh && cs, "invalid arguments");
        poissonDims(width, height, radius, gw, gh, cs);
    });
}

int vm_sampler_resample(vm_sampler *s, int cx0, int cy0, int cx1, int cy1, float *gpuMs)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(s, "null sampler");
        if (gpuMs) *gpuMs = 0.0f;
        int x0 = std::max(cx0, 0), y0 = std::max(cy0, 0), x1 = std::min(cx1, s->gw), y1 = std::min(cy1, s->gh);
        if (x0 >= x1 || y0 >= y1) return;
        vmk::PoiInst p{};
        p.W = s->W, p.H = s->H, p.radius = s->r, p.cellSize = s->cs, p.gw = s->gw, p.gh = s->gh;
        p.x0 = x0, p.y0 = y0, p.x1 = x1, p.y1 = y1, p.seed = s->seed, p.roundBase = s->roundBase, p.gridOff = 0;
        VM_HIP(hipEventRecord(s->e0, s->stream));
        if (vmk::poissonRegionFits(p)) VM_HIP(vmk::launchPoissonRegion(s->stream, p, s->dGrid.p, s->trials));
        else s->runner.run(s->stream, {p}, s->dGrid.p, s->trials);
        VM_HIP(hipEventRecord(s->e1, s->stream));
        s->roundBase += unsigned(s->trials);
        // Rows [y0, y1) come down in one DMA into pinned memory; the region's
        // columns are then copied into the host grid.
        const size_t gw = size_t(s->gw), n = size_t(y1 - y0) * gw;
        s->hRows.reserve(n);
        VM_HIP(hipMemcpyAsync(s->hRows.p, reinterpret_cast<vm_vec2 *>(s->dGrid.p) + size_t(y0) * gw, n * sizeof(vm_vec2),
                              hipMemcpyDeviceToHost, s->stream));
        VM_HIP(hipStreamSynchronize(s->stream));
        for (int y = y0; y < y1; ++y)
            std::memcpy(s->cells.data() + size_t(y) * gw + x0, s->hRows.p + size_t(y - y0) * gw + x0,
                        size_t(x1 - x0) * sizeof(vm_vec2));
        float ms = 0;
        VM_HIP(hipEventElapsedTime(&ms, s->e0, s->e1));
        if (gpuMs) *gpuMs = ms;
    });
}

int vm_sampler_create(float width, float height, float radius, int trials, unsigned seed, vm_sampler **out)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(out && width > 0 && height > 0 && radius > 0 && trials >= 0, "invalid arguments");
        auto s = std::make_unique<vm_sampler>();
        s->W = width, s->H = height, s->r = radius, s->trials = trials, s->seed = seed;
        poissonDims(width, height, radius, &s->gw, &s->gh, &s->cs);
        size_t n = size_t(s->gw) * s->gh;
        s->cells.assign(n, vm_vec2{-1.0f, -1.0f});
        VM_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        VM_HIP(hipEventCreate(&s->e0));
        VM_HIP(hipEventCreate(&s->e1));
        s->dGrid.reserve(2 * n);
        float ms;
        int   rc = vm_sampler_resample(s.get(), 0, 0, s->gw, s->gh, &ms);
        if (rc) throw VmError(rc, vm_last_error());
        *out = s.release();
    });
}

int vm_sampler_destroy(vm_sampler *s)
{
    return vmrt::guardedCall([&] { delete s; });
}

int vm_sampler_verify(vm_sampler *s, int cx0, int cy0, int cx1, int cy1, int *ok)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(s && ok, "invalid arguments");
        // poisson_disk.cu:179-199, host check over the region expanded by 2 cells.
        float r2  = s->r * s->r;
        int   vx0 = std::max(cx0 - 2, 0), vx1 = std::min(cx1 + 2, s->gw);
        int   vy0 = std::max(cy0 - 2, 0), vy1 = std::min(cy1 + 2, s->gh);
        *ok       = 1;
        for (int cy = vy0; cy < vy1 && *ok; ++cy)
            for (int cx = vx0; cx < vx1 && *ok; ++cx) {
                vm_vec2 a = s->cells[size_t(cy) * s->gw + cx];
                if (a.x < 0) continue;
                for (int ny = std::max(cy - 2, 0); ny <= std::min(cy + 2, s->gh - 1); ++ny)
                    for (int nx = std::max(cx - 2, 0); nx <= std::min(cx + 2, s->gw - 1); ++nx) {
                        if (nx == cx && ny == cy) continue;
                        vm_vec2 b = s->cells[size_t(ny) * s->gw + nx];
                        if (b.x < 0) continue;
                        float dx = a.x - b.x, dy = a.y - b.y;
                        if (dx * dx + dy * dy < r2) *ok = 0;
                    }
            }
    });
}

int vm_sampler_load_cells(vm_sampler *s, const vm_vec2 *cells, int n)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(s && cells && size_t(n) == s->cells.size(), "n must be grid_w * grid_h");
        std::copy(cells, cells + n, s->cells.begin());
        VM_HIP(hipMemcpyAsync(s->dGrid.p, s->cells.data(), size_t(n) * sizeof(vm_vec2), hipMemcpyHostToDevice,
                              s->stream));
        VM_HIP(hipStreamSynchronize(s->stream));
    });
}

int vm_sampler_dims(vm_sampler *s, int *gw, int *gh, float *cs)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(s, "null sampler");
        if (gw) *gw = s->gw;
        if (gh) *gh = s->gh;
        if (cs) *cs = s->cs;
    });
}

int vm_sampler_cells(vm_sampler *s, vm_vec2 *cells, int n)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(s && cells && size_t(n) == s->cells.size(), "n must be grid_w * grid_h");
        std::copy(s->cells.begin(), s->cells.end(), cells);
    });
}

int vm_sampler_cells_region(vm_sampler *s, int cx0, int cy0, int cx1, int cy1, vm_vec2 *cells, int n)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(s && cells && size_t(n) == s->cells.size(), "n must be grid_w * grid_h");
        const int x0 = std::max(cx0, 0), y0 = std::max(cy0, 0), x1 = std::min(cx1, s->gw), y1 = std::min(cy1, s->gh);
        for (int y = y0; y < y1; ++y)
            std::copy(s->cells.begin() + size_t(y) * s->gw + x0, s->cells.begin() + size_t(y) * s->gw + x1,
                      cells + size_t(y) * s->gw + x0);
    });
}

int vm_poisson_batch(int n, const float *w, const float *h, const float *r, const unsigned *seed, int trials,
                     const int64_t *offsets, vm_vec2 *cells, float *gpuMs)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(n >= 0 && (n == 0 || (w && h && r && seed && offsets && cells)), "invalid arguments");
        if (gpuMs) *gpuMs = 0.0f;
        if (n == 0) return;
        std::vector<vmk::PoiInst> insts(n);
        int64_t                   total = 0;
        for (int i = 0; i < n; ++i) {
            vmk::PoiInst &p = insts[i];
            std::memset(&p, 0, sizeof(p));
            VM_REQUIRE(w[i] > 0 && h[i] > 0 && r[i] > 0, "invalid instance");
            poissonDims(w[i], h[i], r[i], &p.gw, &p.gh, &p.cellSize);
            p.W = w[i], p.H = h[i], p.radius = r[i], p.seed = seed[i], p.roundBase = 0;
            p.x0 = 0, p.y0 = 0, p.x1 = p.gw, p.y1 = p.gh;
            p.gridOff = offsets[i];
            VM_REQUIRE(offsets[i] >= 0, "negative offset");
            total = std::max(total, offsets[i] + int64_t(p.gw) * p.gh);
        }
        hipStream_t stream;
        VM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        hipEvent_t e0, e1;
        VM_HIP(hipEventCreate(&e0));
        VM_HIP(hipEventCreate(&e1));
        {
            DBuf<float>   grid;
            PoissonRunner runner;
            grid.reserve(2 * size_t(total));
            VM_HIP(hipEventRecord(e0, stream));
            runner.run(stream, insts, grid.p, trials);
            VM_HIP(hipEventRecord(e1, stream));
            VM_HIP(hipMemcpyAsync(cells, grid.p, size_t(total) * sizeof(vm_vec2), hipMemcpyDeviceToHost, stream));
            VM_HIP(hipStreamSynchronize(stream));
            float ms = 0;
            VM_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (gpuMs) *gpuMs = ms;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(stream);
    });
}

} // extern "C"

// ===========================================================================
// Batches of independent fits
// ===========================================================================
struct vm_batch {
    int                           precision = 0;
    hipStream_t                   stream    = nullptr;
    std::unique_ptr<vmrt::Engine> eng;
    bool                          committed = false;
    // vm_batch_extract: occupancy upload and the returned mesh (valid until the next call).
    vmrt::DBuf<uint64_t> dMasks;
    std::vector<vm_vec2> xGen;
    std::vector<uint8_t> xInside;
    std::vector<int32_t> xStart, xLabels;
    std::vector<float>   xVerts;
    vmrt::Boundary       xBound;
    ~vm_batch()
    {
        dMasks.release();
        eng.reset();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace vmrt {

// Instance spec from a public fit description (live sites/targets only).
InstSpec specFromDesc(const vm_fit_desc &d)
{
    VM_REQUIRE(d.width > 0 && d.height > 0, "domain must be positive");
    VM_REQUIRE(d.num_sites >= 0 && d.num_targets >= 0, "negative sizes");
    VM_REQUIRE((d.sites || d.num_sites == 0) && (d.targets || d.num_targets == 0), "null arrays");
    InstSpec s;
    s.W = d.width, s.H = d.height;
    s.pos.reserve(2 * size_t(d.num_sites)), s.tgt.reserve(2 * size_t(d.num_targets));
    s.tslot.reserve(size_t(d.num_targets));
    for (int i = 0; i < d.num_sites; ++i)
        if (d.sites[i].x >= 0.0f) s.pos.push_back(d.sites[i].x), s.pos.push_back(d.sites[i].y);
    for (int i = 0; i < d.num_targets; ++i)
        if (d.targets[i].x >= 0.0f) {
            s.tgt.push_back(d.targets[i].x), s.tgt.push_back(d.targets[i].y);
            s.tslot.push_back(i);
        }
    s.lr = double(d.lr), s.gamma = double(d.gamma), s.beta1 = 0.9, s.beta2 = 0.999, s.eps = 1e-8;
    for (int k = 0; k < 4; ++k) s.milestones[k] = d.milestones[k];
    s.fraction = d.sample_fraction, s.seed = d.seed;
    const int S = std::max(int(s.pos.size() / 2), 1);
    s.K         = d.loss_k > 0 ? d.loss_k : tunedK(S);
    VM_REQUIRE(vmk::lossKSupported(s.K), "unsupported loss_k");
    float cell, ce;
    int   rs, tpl;
    tunedChoice(S, d.width, d.height, s.K, &cell, &rs, &ce, &tpl);
    s.cell     = d.grid_cell > 0 ? d.grid_cell : cell;
    s.stageR   = d.stage_rings > 0 ? d.stage_rings : rs;
    s.chunkExt = d.chunk_extent > 0 ? d.chunk_extent : ce;
    VM_REQUIRE(d.targets_per_lane >= 0 && d.targets_per_lane <= vmk::kMaxTargetsPerLane, "targets_per_lane must be 0-4");
    s.tpl = d.targets_per_lane > 0 ? d.targets_per_lane : tpl;
    VM_REQUIRE(vmk::lossTplSupported(s.tpl, false), "targets_per_lane must be 2, 3 or 4");
    s.maxIter = d.max_iterations, s.tol = d.tol;
    return s;
}

} // namespace vmrt

extern "C" {

int vm_batch_create(int precision, vm_batch **out)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(out && (precision == 0 || precision == 1), "invalid arguments");
        auto b       = std::make_unique<vm_batch>();
        b->precision = precision;
        VM_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        b->eng = vmrt::Engine::create(precision, b->stream);
        *out   = b.release();
    });
}

int vm_batch_destroy(vm_batch *b)
{
    return vmrt::guardedCall([&] { delete b; });
}

int vm_batch_add(vm_batch *b, const vm_fit_desc *d, int *id)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(b && d, "invalid arguments");
        VM_REQUIRE(!b->committed, "batch already committed");
        int i = b->eng->add(vmrt::specFromDesc(*d));
        if (id) *id = i;
    });
}

int vm_batch_commit(vm_batch *b)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(b, "null batch");
        b->eng->commit();
        b->committed = true;
    });
}

int vm_batch_step(vm_batch *b, int iterations, float *gpuMs)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(b && b->committed, "batch not committed");
        float ms = b->eng->step(iterations);
        if (gpuMs) *gpuMs = ms;
    });
}

int vm_batch_profile(vm_batch *b, int iterations, float *phaseMs)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(b && b->committed && phaseMs && iterations > 0, "invalid arguments");
        b->eng->profile(iterations, phaseMs);
    });
}

int vm_batch_get_sites(vm_batch *b, int id, vm_vec2 *out, int n)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(b && out && id >= 0 && id < int(b->eng->specs.size()), "invalid arguments");
        b->eng->pull();
        const vmrt::InstSpec &s = b->eng->specs[id];
        VM_REQUIRE(size_t(n) == s.pos.size() / 2, "n must equal the instance's live site count");
        for (int i = 0; i < n; ++i) out[i] = vm_vec2{float(s.pos[2 * i]), float(s.pos[2 * i + 1])};
    });
}

int vm_batch_extract(vm_batch *b, int id, const uint64_t *masks, int tiles_x, int tiles_y, float voxel_size,
                     vm_mesh *out)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(b && out && id >= 0 && id < int(b->eng->specs.size()), "invalid arguments");
        VM_REQUIRE(!masks || (tiles_x > 0 && tiles_y > 0 && voxel_size > 0.0f),
                   "occupancy masks need positive tile counts and voxel size");
        std::memset(out, 0, sizeof(*out));
        const uint64_t *dm = nullptr;
        if (masks) {
            const size_t nt = size_t(tiles_x) * size_t(tiles_y);
            b->dMasks.reserve(nt);
            VM_HIP(hipMemcpyAsync(b->dMasks.p, masks, nt * sizeof(uint64_t), hipMemcpyHostToDevice, b->stream));
            dm = b->dMasks.p;
        }
        vmrt::ClipOut c;
        b->eng->clip(id, dm, tiles_x, tiles_y, voxel_size, c);
        const auto t0 = std::chrono::steady_clock::now();
        b->eng->pull();
        const vmrt::InstSpec &s = b->eng->specs[id];
        const int             n = int(s.pos.size() / 2);
        VM_REQUIRE(c.S == n, "clip output does not match the instance");
        b->xGen.resize(size_t(n));
        for (int i = 0; i < n; ++i) b->xGen[size_t(i)] = vm_vec2{float(s.pos[2 * i]), float(s.pos[2 * i + 1])};
        b->xInside.assign(c.inside, c.inside + n);
        b->xStart.assign(c.start, c.start + n + 1);
        const size_t nv = size_t(b->xStart[size_t(n)]);
        b->xVerts.assign(c.verts, c.verts + 2 * nv);
        b->xLabels.assign(c.labels, c.labels + nv);
        b->xBound = std::move(c.bound);
        out->num_slots      = n;
        out->generators     = b->xGen.data();
        out->inside         = b->xInside.data();
        out->cell_start     = b->xStart.data();
        out->cell_verts     = reinterpret_cast<const vm_vec2 *>(b->xVerts.data());
        out->cell_labels    = b->xLabels.data();
        out->num_vertices   = int(b->xBound.vertices.size() / 2);
        out->vertices       = reinterpret_cast<const vm_vec2 *>(b->xBound.vertices.data());
        out->num_edges      = int(b->xBound.edges.size() / 2);
        out->edges          = b->xBound.edges.data();
        out->watertight     = b->xBound.watertight ? 1 : 0;
        out->loops          = b->xBound.loops;
        out->unsafe_cells   = c.unsafe;
        out->overflow_cells = c.overflow;
        out->gpu_ms         = c.gpuMs;
        out->host_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

int vm_batch_status(vm_batch *b, int id, int *iteration, int *done, double *lastMean)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(b && id >= 0 && id < int(b->eng->specs.size()), "invalid arguments");
        b->eng->readStates();
        const vmk::InstState &st = b->eng->specs[id].state;
        if (iteration) *iteration = st.iter;
        if (done) *done = st.done;
        if (lastMean) *lastMean = st.count > 0 ? st.loss / double(st.count) : 0.0;
    });
}

int vm_batch_saturation(vm_batch *b, int64_t *targets)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(b && targets, "invalid arguments");
        b->eng->readStates();
        int64_t n = 0;
        for (const auto &s : b->eng->specs) n += s.state.saturated;
        *targets = n;
    });
}

int vm_copy_probe(int64_t bytes, float *gpuMs)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(bytes > 0 && bytes % 4 == 0, "bytes must be a positive multiple of 4");
        DBuf<int> src, dst;
        src.reserve(size_t(bytes / 4)), dst.reserve(size_t(bytes / 4));
        hipStream_t s;
        VM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        hipEvent_t e0, e1;
        VM_HIP(hipEventCreate(&e0));
        VM_HIP(hipEventCreate(&e1));
        VM_HIP(vmk::launchFillInt(s, src.p, bytes / 4, 1));
        VM_HIP(hipEventRecord(e0, s));
        VM_HIP(hipMemcpyAsync(dst.p, src.p, size_t(bytes), hipMemcpyDeviceToDevice, s));
        VM_HIP(hipEventRecord(e1, s));
        VM_HIP(hipStreamSynchronize(s));
        float ms = 0;
        VM_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (gpuMs) *gpuMs = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
    });
}

int vm_batch_loss_pass(vm_batch *b, double *gradXY, double *lossSum, int64_t *count, int32_t *nearest,
                       int32_t *partner, double *lossTerms)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(b && b->committed, "batch not committed");
        const bool    wantTerms = nearest || partner || lossTerms;
        vmrt::DiagOut d;
        b->eng->diagAll(gradXY != nullptr, wantTerms, d);
        if (lossSum) *lossSum = d.lossSum;
        if (count) *count = d.count;
        if (gradXY) std::copy(d.grad.begin(), d.grad.end(), gradXY);
        for (size_t t = 0; t < d.terms.size(); ++t) {
            if (nearest) nearest[t] = d.terms[t].i0;
            if (partner) partner[t] = d.terms[t].j;
            if (lossTerms) lossTerms[t] = d.terms[t].loss;
        }
    });
}

int vm_batch_loss_stats(vm_batch *b, uint64_t *counters)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(b && b->committed && counters, "invalid arguments");
        b->eng->lossStats(counters);
    });
}

int vm_batch_choice(vm_batch *b, int id, int *k, float *cell, int *rings, float *chunkExtent, int *targetsPerLane)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(b && id >= 0 && id < int(b->eng->specs.size()), "invalid arguments");
        const vmrt::InstSpec &s = b->eng->specs[id];
        if (k) *k = s.K;
        if (cell) *cell = s.cell;
        if (rings) *rings = s.stageR;
        if (chunkExtent) *chunkExtent = s.chunkExt;
        if (targetsPerLane) *targetsPerLane = b->precision == 1 ? vmk::kTargetsPerLane : s.tpl;
    });
}

} // extern "C"
