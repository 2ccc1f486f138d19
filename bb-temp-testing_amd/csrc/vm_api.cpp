// vm_api.cpp -- the C ABI (include/voromesh_hip.h): the plain optimizer handle
// with the reference's slot semantics, the Poisson sampler, and batches.
#include "voromesh_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
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

inline bool liveX(float x) { return x >= 0.0f; } // NaN and negative x mark an empty slot

hipStream_t makeStream()
{
    hipStream_t s = nullptr;
    VM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

} // namespace

// ===========================================================================
// Plain optimizer (voromesh::VoroOptimizer)
// ===========================================================================
// Sites are kept in the reference's slot layout (voro_optimizer.cu:613-650):
// instance 0's site i is generator slot i, empty slots (x < 0) stay in place
// and are never binned or moved.  Targets are laid out sorted by hash cell.
// While the engine is committed, local edits patch the device state without a
// re-layout: UpdateGenerators copies the slot rectangle and bins again,
// BeginSchedule clears the rectangle's moments and rewrites the instance
// record, and UpdateTargets retires the dirty slots' sorted entries and lays
// the new ones out as chunks of their own in room kept past the targets.  The
// host work of an edit scales with the rectangle and the dirty slots; a full
// re-layout happens when that room runs out or the slot arrays are replaced.
struct vm_handle {
    float                          W = 1, H = 1, spacing = 1;
    vm_params                      prm{};
    hipStream_t                    stream = nullptr;
    std::unique_ptr<vmrt::Engine>  eng;
    int                            slotsW = 0, slotsH = 0;
    float                          slotCellSize = 0;
    // Host mirrors of the slot arrays; stale while the device is newer.
    std::vector<vm_vec2>           slotPos, slotM, slotV;
    std::vector<vm_vec2>           targetSlots; // always current
    int                            targetCapacity = 0;
    // Target slot -> pooled sorted position (-1: none), valid while committed.
    std::vector<int64_t>           tposOf;
    std::vector<uint32_t>          seen;
    uint32_t                       seenGen = 0;
    int                            liveTargets = 0;
    bool                           patched  = false; // targets edited since the last layout
    bool                           fullEdits = false; // every edit re-lays out (VM_EDIT_FULL=1)
    std::vector<uint64_t>          masks;
    DBuf<uint64_t>                 dMasks;
    HBuf<uint64_t>                 hMaskRows; // pinned staging of occupancy rows
    int                            tilesX = 0, tilesY = 0;
    float                          voxel = 0;
    int                            iteration = 0, scheduleLength = 0;
    float                          lastBatchLoss = 0;
    int                            rect[4] = {0, 0, 0, 0};
    bool                           haveGenerators = false;
    // Extract() storage
    std::vector<vm_vec2> mGen, mVerts, mVertices;
    std::vector<uint8_t> mInside;
    std::vector<int32_t> mStart, mLabels, mEdges;

    ~vm_handle()
    {
        eng.reset();
        dMasks.release();
        if (stream) (void)hipStreamDestroy(stream);
    }

    vmrt::InstSpec &spec() { return eng->specs[0]; }
    int             numSlots() const { return slotsW * slotsH; }

    // Device -> host mirrors (committed engine with newer device state).
    void pullSlots()
    {
        if (eng->dirty || !eng->deviceNewer) return;
        eng->pull();
        const vmrt::InstSpec &s = spec();
        for (int i = 0; i < numSlots(); ++i) {
            slotPos[size_t(i)] = vm_vec2{float(s.pos[2 * size_t(i)]), float(s.pos[2 * size_t(i) + 1])};
            slotM[size_t(i)]   = vm_vec2{float(s.m[2 * size_t(i)]), float(s.m[2 * size_t(i) + 1])};
            slotV[size_t(i)]   = vm_vec2{float(s.v[2 * size_t(i)]), float(s.v[2 * size_t(i) + 1])};
        }
    }

    // The next use re-lays out everything from the host mirrors.
    void markDirty()
    {
        if (!eng->dirty) pullSlots();
        eng->dirty = true;
    }

    // Host mirrors -> instance spec (slot layout), before a full commit.
    void buildSites()
    {
        vmrt::InstSpec &s = spec();
        const size_t    n = size_t(numSlots());
        s.pos.resize(2 * n), s.m.resize(2 * n), s.v.resize(2 * n);
        int live = 0;
        for (size_t i = 0; i < n; ++i) {
            s.pos[2 * i] = slotPos[i].x, s.pos[2 * i + 1] = slotPos[i].y;
            s.m[2 * i] = slotM[i].x, s.m[2 * i + 1] = slotM[i].y;
            s.v[2 * i] = slotV[i].x, s.v[2 * i + 1] = slotV[i].y;
            live += liveX(slotPos[i].x) ? 1 : 0;
        }
        s.liveS = live;
        float cell, ce;
        int   rs, tpl;
        vmrt::tunedChoice(std::max(live, 1), W, H, s.K, &cell, &rs, &ce, &tpl);
        s.cell     = prm.grid_cell > 0.0f ? prm.grid_cell : cell;
        s.stageR   = prm.stage_rings > 0 ? prm.stage_rings : rs;
        s.chunkExt = ce;
        s.tpl      = tpl;
        s.frozen.clear();
    }

    void compactTargets()
    {
        vmrt::InstSpec &s = spec();
        s.tgt.clear(), s.tslot.clear();
        for (int t = 0; t < int(targetSlots.size()); ++t) {
            if (!liveX(targetSlots[t].x)) continue;
            s.tgt.push_back(targetSlots[t].x), s.tgt.push_back(targetSlots[t].y);
            s.tslot.push_back(t);
        }
        liveTargets = int(s.tslot.size());
    }

    // Commits a pending full layout (with room for local target edits) and
    // maps target slots to their sorted positions.
    void ensureCommitted()
    {
        if (!eng->dirty) return;
        buildSites();
        compactTargets();
        const int64_t T = int64_t(spec().tslot.size());
        eng->tailTargets = std::max<int64_t>(16384, T / 4);
        eng->tailWork    = int(eng->tailTargets / 16 + 4096);
        eng->commit();
        tposOf.assign(size_t(std::max(targetCapacity, int(targetSlots.size()))), -1);
        const std::vector<int> &perm = eng->permOf(0);
        for (size_t t = 0; t < perm.size(); ++t) tposOf[size_t(spec().tslot[size_t(perm[t])])] = int64_t(t);
        patched = false;
    }

    void setScheduleFields(int iterations)
    {
        vmrt::InstSpec &s = spec();
        s.lr    = double(prm.lr_per_spacing * spacing);
        s.gamma = prm.gamma, s.beta1 = 0.9, s.beta2 = 0.999, s.eps = 1e-8;
        for (int k = 0; k < 4; ++k)
            s.milestones[k] = prm.iterations > 0
                                  ? int(std::lround(double(prm.milestones[k]) * iterations / prm.iterations))
                                  : prm.milestones[k];
        float margin = 2.0f * spacing;
        s.useBox     = 1;
        s.box[0] = float(rect[0]) * slotCellSize - margin, s.box[1] = float(rect[1]) * slotCellSize - margin;
        s.box[2] = float(rect[2]) * slotCellSize + margin, s.box[3] = float(rect[3]) * slotCellSize + margin;
        const bool all = rect[0] <= 0 && rect[1] <= 0 && rect[2] >= slotsW && rect[3] >= slotsH;
        s.rectW = all ? 0 : slotsW;
        for (int k = 0; k < 4; ++k) s.rect[k] = rect[k];
    }

    void beginSchedule(int iterations, int cx0, int cy0, int cx1, int cy1)
    {
        cx0 = std::max(cx0, 0), cy0 = std::max(cy0, 0);
        cx1 = std::min(cx1, slotsW), cy1 = std::min(cy1, slotsH);
        rect[0] = cx0, rect[1] = cy0, rect[2] = cx1, rect[3] = cy1;
        setScheduleFields(iterations);
        if (eng->dirty || fullEdits) {
            markDirty();
            if (cx0 < cx1 && cy0 < cy1)
                for (int y = cy0; y < cy1; ++y)
                    for (int x = cx0; x < cx1; ++x)
                        slotM[size_t(y) * slotsW + x] = vm_vec2{0, 0}, slotV[size_t(y) * slotsW + x] = vm_vec2{0, 0};
            spec().state = vmk::InstState{};
        } else {
            // Moments of the rectangle cleared on the device; the record and a
            // zero step state replace the instance's.
            eng->uploadSiteRect(slotsW, cx0, cy0, cx1, cy1, nullptr, true);
            eng->updateRecord(0, liveTargets, true);
        }
        scheduleLength = iterations;
        iteration      = 0;
        lastBatchLoss  = 0.0f;
    }
};

extern "C" {

void vm_default_params(vm_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->iterations       = 400;
    p->local_iterations = 100;
    p->lr_per_spacing   = 0.09f;
    p->sample_fraction  = 0.2f;
    p->gamma            = 0.5f;
    p->milestones[0] = 80, p->milestones[1] = 150, p->milestones[2] = 200, p->milestones[3] = 250;
    p->seed        = 1;
    p->loss_k      = 12;
    p->grid_cell   = 0.0f;
    p->stage_rings = 0;
    p->precision   = 0;
}

const char *vm_last_error(void) { return vmrt::lastError(); }
const char *vm_version(void) { return "voromesh-hip 0.1 (gfx950)"; }

int vm_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int vm_create(float width, float height, float spacing, const vm_params *params, vm_handle **out)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(out, "out is null");
        VM_REQUIRE(width > 0 && height > 0 && spacing > 0, "width, height and spacing must be positive");
        auto h = std::make_unique<vm_handle>();
        h->W = width, h->H = height, h->spacing = spacing;
        if (params) h->prm = *params;
        else vm_default_params(&h->prm);
        VM_REQUIRE(vmk::lossKSupported(h->prm.loss_k), "unsupported loss_k");
        const char *full = std::getenv("VM_EDIT_FULL");
        h->fullEdits     = full && full[0] == '1';
        h->stream = makeStream();
        h->eng    = vmrt::Engine::create(h->prm.precision, h->stream);
        vmrt::InstSpec s;
        s.W = width, s.H = height;
        s.K        = h->prm.loss_k;
        s.fraction = h->prm.sample_fraction;
        s.seed     = h->prm.seed;
        float cell, ce;
        int   rs, tpl;
        vmrt::tunedChoice(1, width, height, s.K, &cell, &rs, &ce, &tpl);
        s.cell     = h->prm.grid_cell > 0.0f ? h->prm.grid_cell : 2.0f * spacing;
        s.stageR   = h->prm.stage_rings > 0 ? h->prm.stage_rings : rs;
        s.chunkExt = ce;
        s.tpl      = tpl;
        h->eng->add(std::move(s));
        *out = h.release();
    });
}

int vm_destroy(vm_handle *h)
{
    return vmrt::guardedCall([&] { delete h; });
}

int vm_set_targets(vm_handle *h, const vm_vec2 *slots, int n)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h && n >= 0 && (slots || n == 0), "invalid arguments");
        h->markDirty();
        h->targetCapacity = std::max(n, 1);
        h->targetSlots.assign(slots, slots + n);
    });
}

int vm_update_targets(vm_handle *h, const int32_t *dirty, int numDirty, const vm_vec2 *slots, int numSlots)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h && numDirty >= 0 && numSlots >= 0, "invalid arguments");
        if (numDirty == 0) return;
        VM_REQUIRE(dirty && slots, "null arrays");
        if (numSlots > h->targetCapacity) {
            // Grow with slack and re-upload everything (voro_optimizer.cu:566-571).
            h->markDirty();
            size_t cap = std::max(size_t(numSlots) * 3 / 2, size_t(numSlots) + 64);
            h->targetSlots.assign(slots, slots + numSlots);
            h->targetSlots.resize(cap, vm_vec2{-1.0f, -1.0f});
            h->targetCapacity = int(cap);
            return;
        }
        if (int(h->targetSlots.size()) < h->targetCapacity)
            h->targetSlots.resize(size_t(h->targetCapacity), vm_vec2{NAN, NAN});
        for (int k = 0; k < numDirty; ++k) {
            const int d = dirty[k];
            VM_REQUIRE(d >= 0 && d < numSlots && d < h->targetCapacity, "dirty slot out of range");
        }
        if (h->eng->dirty || h->fullEdits || h->tposOf.size() < h->targetSlots.size()) {
            h->markDirty();
            for (int k = 0; k < numDirty; ++k) h->targetSlots[size_t(dirty[k])] = slots[dirty[k]];
            return;
        }
        // Local edit: the dirty slots' sorted entries are retired and their new
        // targets laid out as chunks of their own past the others.
        if (h->seen.size() < h->targetSlots.size()) h->seen.resize(h->targetSlots.size(), 0);
        ++h->seenGen;
        std::vector<int64_t> kill;
        std::vector<double>  addXY;
        std::vector<int>     addSlot;
        for (int k = 0; k < numDirty; ++k) {
            const int d = dirty[k];
            if (h->seen[size_t(d)] == h->seenGen) continue;
            h->seen[size_t(d)] = h->seenGen;
            if (h->tposOf[size_t(d)] >= 0) kill.push_back(h->tposOf[size_t(d)]);
            if (liveX(slots[d].x)) addXY.push_back(slots[d].x), addXY.push_back(slots[d].y), addSlot.push_back(d);
        }
        std::vector<int64_t> pos;
        if (!h->eng->appendTargets(addXY, addSlot, pos)) { // no room left: re-lay out at the next use
            h->markDirty();
            for (int k = 0; k < numDirty; ++k) h->targetSlots[size_t(dirty[k])] = slots[dirty[k]];
            return;
        }
        h->eng->killTargets(kill);
        h->liveTargets -= int(kill.size());
        for (int k = 0; k < numDirty; ++k) {
            const int d = dirty[k];
            h->targetSlots[size_t(d)] = slots[d];
            h->tposOf[size_t(d)]      = -1;
        }
        for (size_t k = 0; k < addSlot.size(); ++k) h->tposOf[size_t(addSlot[k])] = pos[k];
        h->liveTargets += int(addSlot.size());
        h->patched = true;
        h->eng->updateRecord(0, h->liveTargets, false); // the gradient shift follows the live target count
    });
}

int vm_set_occupancy(vm_handle *h, const uint64_t *masks, int tilesX, int tilesY, float voxelSize)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h && tilesX >= 0 && tilesY >= 0 && voxelSize > 0, "invalid arguments");
        size_t n = size_t(tilesX) * tilesY;
        VM_REQUIRE(masks || n == 0, "null masks");
        h->masks.assign(masks, masks + n);
        h->tilesX = tilesX, h->tilesY = tilesY, h->voxel = voxelSize;
        h->dMasks.reserve(n);
        // On the optimizer's own stream (ordered after its kernels), then waited for.
        if (n) VM_HIP(hipMemcpyAsync(h->dMasks.p, h->masks.data(), n * 8, hipMemcpyHostToDevice, h->stream));
        VM_HIP(hipStreamSynchronize(h->stream));
    });
}

int vm_update_occupancy(vm_handle *h, const uint64_t *masks, int tx0, int ty0, int tx1, int ty1)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h, "null handle");
        tx0 = std::max(tx0, 0), ty0 = std::max(ty0, 0);
        tx1 = std::min(tx1, h->tilesX), ty1 = std::min(ty1, h->tilesY);
        if (tx0 >= tx1 || ty0 >= ty1) return;
        VM_REQUIRE(masks, "null masks");
        for (int y = ty0; y < ty1; ++y)
            for (int x = tx0; x < tx1; ++x) h->masks[size_t(y) * h->tilesX + x] = masks[size_t(y) * h->tilesX + x];
        // Whole rows [ty0, ty1) go up in one DMA from pinned memory (the host copy
        // equals the device copy outside the rect).
        const size_t off = size_t(ty0) * h->tilesX, n = size_t(ty1 - ty0) * h->tilesX;
        h->hMaskRows.reserve(n);
        std::memcpy(h->hMaskRows.p, h->masks.data() + off, n * sizeof(uint64_t));
        VM_HIP(hipMemcpyAsync(h->dMasks.p + off, h->hMaskRows.p, n * sizeof(uint64_t), hipMemcpyHostToDevice, h->stream));
        VM_HIP(hipStreamSynchronize(h->stream));
    });
}

int vm_set_generators(vm_handle *h, const vm_vec2 *slots, int slotsW, int slotsH, float slotCellSize)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h && slotsW >= 0 && slotsH >= 0, "invalid arguments");
        size_t n = size_t(slotsW) * slotsH;
        VM_REQUIRE(slots || n == 0, "null slots");
        VM_REQUIRE(n < size_t(vmk::kMaxSitesPerInst), "too many generator slots");
        h->eng->dirty = true; // replaced wholesale: nothing to pull
        h->slotsW = slotsW, h->slotsH = slotsH, h->slotCellSize = slotCellSize;
        h->slotPos.assign(slots, slots + n);
        h->slotM.assign(n, vm_vec2{0, 0});
        h->slotV.assign(n, vm_vec2{0, 0});
        h->eng->deviceNewer = false;
        h->haveGenerators   = true;
        h->beginSchedule(h->prm.iterations, 0, 0, slotsW, slotsH);
    });
}

int vm_update_generators(vm_handle *h, const vm_vec2 *slots, int cx0, int cy0, int cx1, int cy1)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h, "null handle");
        cx0 = std::max(cx0, 0), cy0 = std::max(cy0, 0);
        cx1 = std::min(cx1, h->slotsW), cy1 = std::min(cy1, h->slotsH);
        if (cx0 >= cx1 || cy0 >= cy1) return;
        VM_REQUIRE(slots, "null slots");
        if (h->eng->dirty || h->fullEdits) {
            h->markDirty();
            for (int y = cy0; y < cy1; ++y)
                for (int x = cx0; x < cx1; ++x)
                    h->slotPos[size_t(y) * h->slotsW + x] = slots[size_t(y) * h->slotsW + x];
            return;
        }
        // Local edit: the rectangle's slots go up as one 2D copy; the sites are
        // binned again (empty slots stay out of the grid).
        h->eng->uploadSiteRect(h->slotsW, cx0, cy0, cx1, cy1, reinterpret_cast<const float *>(slots), false);
        h->eng->rebin();
    });
}

int vm_begin_schedule(vm_handle *h, int iterations, int cx0, int cy0, int cx1, int cy1)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h, "null handle");
        h->beginSchedule(iterations, cx0, cy0, cx1, cy1);
    });
}

int vm_step(vm_handle *h, int iterations, float *gpuMs)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h, "null handle");
        float ms = 0.0f;
        if (iterations > 0 && h->numSlots() > 0 && h->targetCapacity > 0) {
            h->ensureCommitted();
            ms = h->eng->step(iterations);
            h->eng->readStates();
            const vmk::InstState &st = h->spec().state;
            h->iteration     = st.iter;
            h->lastBatchLoss = st.count > 0 ? float(st.loss / double(st.count)) : 0.0f;
        }
        if (gpuMs) *gpuMs = ms;
    });
}

int vm_get_state(vm_handle *h, int *iteration, int *scheduleLength, float *lastBatchLoss, int *numSlots)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h, "null handle");
        if (iteration) *iteration = h->iteration;
        if (scheduleLength) *scheduleLength = h->scheduleLength;
        if (lastBatchLoss) *lastBatchLoss = h->lastBatchLoss;
        if (numSlots) *numSlots = h->numSlots();
    });
}

int vm_grad_saturation(vm_handle *h, int64_t *targets)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h && targets, "invalid arguments");
        *targets = 0;
        if (h->numSlots() == 0 || h->eng->dirty) return;
        h->eng->readStates();
        *targets = h->spec().state.saturated;
    });
}

int vm_graph_stats(vm_handle *h, int64_t *updates, int64_t *instantiations)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h, "null handle");
        if (updates) *updates = h->eng->graphUpdates;
        if (instantiations) *instantiations = h->eng->graphInstantiations;
    });
}

int vm_evaluate_loss(vm_handle *h, double *mean)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h && mean, "invalid arguments");
        *mean = 0.0;
        if (h->numSlots() == 0 || h->targetCapacity == 0) return;
        h->ensureCommitted();
        int64_t cnt = 0;
        double  sum = h->eng->evaluate(0, &cnt);
        *mean       = cnt > 0 ? sum / double(cnt) : 0.0;
    });
}

int vm_get_generators(vm_handle *h, vm_vec2 *slots, int n)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h && slots && n == h->numSlots(), "n must equal the number of slots");
        h->pullSlots();
        std::copy(h->slotPos.begin(), h->slotPos.end(), slots);
    });
}

int vm_extract(vm_handle *h, vm_mesh *out)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h && out, "invalid arguments");
        std::memset(out, 0, sizeof(*out));
        const int n = h->numSlots();
        h->mGen.clear(), h->mVerts.clear(), h->mVertices.clear(), h->mInside.clear();
        h->mStart.assign(1, 0), h->mLabels.clear(), h->mEdges.clear();
        if (n == 0) {
            out->cell_start = h->mStart.data();
            return;
        }
        h->ensureCommitted();
        vmrt::ClipOut c;
        h->eng->clip(0, h->masks.empty() ? nullptr : h->dMasks.p, h->tilesX, h->tilesY, h->voxel, c);
        auto t0 = std::chrono::steady_clock::now();
        h->pullSlots();
        h->mGen = h->slotPos;
        // Site i is slot i (empty slots have no cell).
        h->mInside.assign(c.inside, c.inside + n);
        h->mStart.assign(c.start, c.start + n + 1);
        const size_t nv = size_t(h->mStart[n]);
        h->mVerts.resize(nv);
        std::memcpy(h->mVerts.data(), c.verts, nv * sizeof(vm_vec2));
        h->mLabels.assign(c.labels, c.labels + nv);
        // The boundary was assembled on the device in site space, which is slot
        // space: its vertex numbering and edges are those of the slots.
        const vmrt::Boundary &bd = c.bound;
        const int V = int(bd.vertices.size() / 2), E = int(bd.edges.size() / 2);
        h->mVertices.resize(size_t(V));
        for (int v = 0; v < V; ++v) h->mVertices[size_t(v)] = vm_vec2{bd.vertices[2 * v], bd.vertices[2 * v + 1]};
        h->mEdges.assign(bd.edges.begin(), bd.edges.end());
        const bool wt    = bd.watertight;
        const int  loops = bd.loops;
        out->num_slots      = n;
        out->generators     = h->mGen.data();
        out->inside         = h->mInside.data();
        out->cell_start     = h->mStart.data();
        out->cell_verts     = h->mVerts.data();
        out->cell_labels    = h->mLabels.data();
        out->num_vertices   = V;
        out->vertices       = h->mVertices.data();
        out->num_edges      = E;
        out->edges          = h->mEdges.data();
        out->watertight     = wt ? 1 : 0;
        out->loops          = loops;
        out->unsafe_cells   = c.unsafe;
        out->overflow_cells = c.overflow;
        out->gpu_ms         = c.gpuMs;
        out->host_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

int vm_knn(vm_handle *h, const vm_vec2 *queries, int nq, int k, int32_t *idx, float *d2)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h && nq >= 0 && (queries || nq == 0) && idx && d2, "invalid arguments");
        VM_REQUIRE(k >= 1 && k <= 32, "k must be in [1, 32]");
        std::vector<double> q(2 * size_t(nq)), dd(size_t(nq) * k);
        for (int i = 0; i < nq; ++i) q[2 * i] = queries[i].x, q[2 * i + 1] = queries[i].y;
        h->ensureCommitted();
        h->eng->knnQuery(0, q.data(), nq, k, idx, dd.data()); // site indices are slots
        for (size_t i = 0; i < dd.size(); ++i) d2[i] = float(dd[i]);
    });
}

int vm_loss_pass(vm_handle *h, double *gradXY, double *lossSum, int64_t *count, int32_t *knn, int32_t *nearest,
                 int32_t *partner, double *lossTerms)
{
    return vmrt::guardedCall([&] {
        VM_REQUIRE(h, "null handle");
        bool wantDbg = knn || nearest || partner || lossTerms;
        // Per-target diagnostics follow the laid-out order: edited targets are
        // laid out again first.
        if (wantDbg && h->patched) h->markDirty();
        h->ensureCommitted();
        vmrt::DiagOut d;
        h->eng->diag(0, gradXY != nullptr, wantDbg, d);
        if (lossSum) *lossSum = d.lossSum;
        if (count) *count = d.count;
        if (gradXY) std::copy(d.grad.begin(), d.grad.begin() + 2 * size_t(h->numSlots()), gradXY); // sites are slots
        if (wantDbg) {
            const vmrt::InstSpec &s = h->spec();
            const int K = s.K, nT = int(h->targetSlots.size());
            if (knn) std::fill(knn, knn + size_t(nT) * K, -1);
            if (nearest) std::fill(nearest, nearest + nT, -1);
            if (partner) std::fill(partner, partner + nT, -1);
            if (lossTerms) std::fill(lossTerms, lossTerms + nT, 0.0);
            auto toSlot = [](int i) { return i >= 0 ? i : -1; };
            for (size_t t = 0; t < s.tslot.size(); ++t) {
                int ts = s.tslot[t];
                if (knn)
                    for (int k = 0; k < K; ++k) knn[size_t(ts) * K + k] = toSlot(d.knn[t * K + k]);
                if (nearest) nearest[ts] = toSlot(d.terms[t].i0);
                if (partner) partner[ts] = toSlot(d.terms[t].j);
                if (lossTerms) lossTerms[ts] = d.terms[t].loss;
            }
        }
    });
}

} // extern "C"
