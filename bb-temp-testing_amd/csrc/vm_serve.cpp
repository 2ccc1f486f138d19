// vm_serve.cpp -- continuous-batching request loop over resident fit state.
//
// Requests arrive on a schedule.  Worker threads prepare each arrived request
// (live sites/targets, target sort and wavefront chunking).  The GPU loop
// admits prepared requests in arrival order whenever an instance slot and
// room in the device pools (bounded by the HBM budget) are free, uploads and
// bins them, and keeps replaying one captured iteration over every resident
// instance.  Every `poll_every` iterations it reads the per-instance states:
// instances the scan kernel marked done (relative loss change below tol, or
// the iteration cap) are retired, their pool ranges freed, and new requests
// join the very next iteration.  Sites, Adam moments, hash grids and targets
// stay on the device for the whole life of a request.
#include "voromesh_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "vm_engine.h"

using vmrt::DBuf;
using vmrt::VmError;

namespace vmrt {
InstSpec specFromDesc(const vm_fit_desc &d);
}

namespace {

// First-fit allocator over [0, cap) in units of 64 elements.
class RangeAlloc {
public:
    void init(int64_t cap)
    {
        free_.clear();
        if (cap > 0) free_[0] = cap;
        cap_ = cap;
    }
    int64_t alloc(int64_t n)
    {
        n = round(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            int64_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > n) free_[off + n] = len - n;
            used_ += n;
            return off;
        }
        return -1;
    }
    void release(int64_t off, int64_t n)
    {
        n = round(n);
        used_ -= n;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + n == next->first) {
            n += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += n;
                return;
            }
        }
        free_[off] = n;
    }
    int64_t used() const { return used_; }
    static int64_t round(int64_t n) { return ((n > 0 ? n : 1) + 63) & ~int64_t(63); }

private:
    std::map<int64_t, int64_t> free_;
    int64_t                    cap_ = 0, used_ = 0;
};

using Clock = std::chrono::steady_clock;

struct Prepared {
    vmrt::InstSpec                spec;
    vmk::Inst                     in{};
    std::vector<vmk::SiteWork>    sw;
    std::vector<vmk::TargetWork>  tw;
    std::vector<int>              perm;
    int64_t                       cells = 0;
    // Upload images in the device element type, built by the preparation
    // worker: positions, targets in sorted order and their slot numbers.
    std::vector<unsigned char>    pos, tgt;
    std::vector<int>              tslot;
};

// Grow-only pinned host buffer.
struct PinnedBuf {
    unsigned char *p = nullptr;
    size_t         n = 0;
    PinnedBuf()                             = default;
    PinnedBuf(const PinnedBuf &)            = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf()
    {
        if (p) (void)hipHostFree(p);
    }
    void reserve(size_t bytes)
    {
        if (bytes <= n && p) return;
        if (p) (void)hipHostFree(p), p = nullptr;
        n = std::max(bytes + bytes / 2, size_t(1) << 20);
        VM_HIP(hipHostMalloc(reinterpret_cast<void **>(&p), n, hipHostMallocDefault));
    }
};

template <class R> class Server {
    using V = vmk::V2<R>;

public:
    Server(const vm_serve_config &cfg, int nReq, int64_t needSites, int64_t needTargets, int64_t needCells)
        : cfg_(cfg)
    {
        VM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        slots_ = cfg.max_resident > 0 ? cfg.max_resident : nReq;
        // Pool capacities: what the budget allows, but no more than the whole
        // request stream could ever hold at once.
        const double perSite = 4.0 * sizeof(V) + 2.0 * sizeof(vmk::GradT<R>) + 3 * sizeof(int) + 1;
        const double perTgt = sizeof(V) + sizeof(int) + sizeof(vmk::ExactItem), perCell = 2 * sizeof(int);
        const double need = perSite * needSites + perTgt * needTargets + perCell * needCells;
        const double scale = std::min(1.0, cfg.hbm_budget_bytes / std::max(need, 1.0));
        capSites_   = RangeAlloc::round(int64_t(needSites * scale) + 64 * slots_);
        capTargets_ = RangeAlloc::round(int64_t(needTargets * scale) + 64 * slots_);
        capCells_   = RangeAlloc::round(int64_t(needCells * scale) + 64 * slots_);
        budgetBytes_ = cfg.hbm_budget_bytes;
        sites_.init(capSites_), targets_.init(capTargets_), cells_.init(capCells_);
        dPos.reserve(capSites_), dM.reserve(capSites_), dV.reserve(capSites_), dSpos.reserve(capSites_);
        dGrad.reserve(2 * capSites_), dCellOf.reserve(capSites_), dRank.reserve(capSites_), dSidx.reserve(capSites_);
        dFrozen.reserve(1);
        dTpos.reserve(capTargets_), dTslot.reserve(capTargets_);
        dCounts.reserve(capCells_), dStart.reserve(capCells_);
        dInst.reserve(slots_), dState.reserve(slots_);
        maxTiles_ = int(capCells_ / vmk::kScanTileCells) + slots_ + 1;
        dTiles.reserve(maxTiles_), dScanCtl.reserve(1), dFlags.reserve(maxTiles_);
        dPartial.reserve(size_t(slots_) * vmk::kLossPartials);
        dSiteWork.reserve(capSites_ / vmk::kSiteChunk + slots_), dNewWork.reserve(capSites_ / vmk::kSiteChunk + slots_);
        dCnt.reserve(4), dNewCnt.reserve(1);
        dExactQ.reserve(capTargets_), dExactLen.reserve(1);
        twCap_ = 0;
        VM_HIP(hipEventCreate(&ev0_));
        VM_HIP(hipEventCreate(&ev1_));
        slotReq_.assign(slots_, -1);
        slotInfo_.resize(slots_);
        VM_HIP(hipMemsetAsync(dState.p, 0, slots_ * sizeof(vmk::InstState), stream_));
        VM_HIP(hipMemsetAsync(dScanCtl.p, 0, sizeof(vmk::ScanCtl), stream_));
        VM_HIP(hipMemsetAsync(dFlags.p, 0, dFlags.n * sizeof(unsigned long long), stream_));
        VM_HIP(hipMemsetAsync(dPartial.p, 0, size_t(slots_) * vmk::kLossPartials * sizeof(vmk::LossPartial), stream_));
        VM_HIP(hipStreamSynchronize(stream_));
    }
    ~Server()
    {
        // Graph replays and admission copies may still be queued (an exception
        // can leave the loop after launch()): let them finish before the
        // buffers and host images they use are released.
        (void)hipStreamSynchronize(stream_);
        if (graph_) (void)hipGraphExecDestroy(graph_);
        (void)hipEventDestroy(ev0_);
        (void)hipEventDestroy(ev1_);
        (void)hipStreamDestroy(stream_);
    }

    double peakBytes() const { return peakBytes_; }
    int    maxResident() const { return maxResident_; }

    // Upload images of a prepared request in this server's element type.
    static void buildImages(Prepared &p)
    {
        const int S = p.in.S, T = p.in.T;
        p.pos.resize(size_t(S) * sizeof(V)), p.tgt.resize(size_t(T) * sizeof(V)), p.tslot.resize(size_t(T));
        V *hp = reinterpret_cast<V *>(p.pos.data()), *ht = reinterpret_cast<V *>(p.tgt.data());
        for (int k = 0; k < S; ++k) hp[k] = V{R(p.spec.pos[2 * k]), R(p.spec.pos[2 * k + 1])};
        for (int t = 0; t < T; ++t) {
            const int o = p.perm[size_t(t)];
            ht[t]       = V{R(p.spec.tgt[2 * o]), R(p.spec.tgt[2 * o + 1])};
            p.tslot[size_t(t)] = p.spec.tslot.empty() ? o : p.spec.tslot[size_t(o)];
        }
    }

    // Tries to place a prepared request; returns false when no slot or pool room.
    // The uploads are only enqueued: the request's host images are kept (it is
    // moved into the pending list) until refresh() has synchronized the stream.
    bool admit(int req, std::unique_ptr<Prepared> &pp)
    {
        Prepared &p = *pp;
        int slot = -1;
        for (int s = 0; s < slots_; ++s)
            if (slotReq_[s] < 0) { slot = s; break; }
        if (slot < 0) return false;
        const int S = p.in.S, T = p.in.T;
        const double bytes = bytesFor(S, T, p.cells);
        if (residentBytes_ + bytes > budgetBytes_) return false;
        int64_t so = sites_.alloc(S), to = targets_.alloc(T), co = cells_.alloc(p.cells);
        if (so < 0 || to < 0 || co < 0) {
            if (so >= 0) sites_.release(so, S);
            if (to >= 0) targets_.release(to, T);
            if (co >= 0) cells_.release(co, p.cells);
            return false;
        }
        vmk::Inst in = p.in;
        in.siteOff = so, in.cellOff = co, in.targetOff = to;
        SlotInfo &si = slotInfo_[slot];
        si       = SlotInfo{so, to, co, S, T, p.cells, bytes, {}, {}};
        for (auto w : p.sw) si.sw.push_back({slot, w.begin, w.count, 0});
        for (auto w : p.tw) { // prepared at offset 0: move to the pool ranges
            w.inst = slot, w.tBegin += to, w.cellOff = co, w.siteOff = so;
            si.tw.push_back(w);
        }
        // Upload: positions, zero moments/gradient/counts, sorted targets, record, state.
        if (S) {
            VM_HIP(hipMemcpyAsync(dPos.p + so, p.pos.data(), S * sizeof(V), hipMemcpyHostToDevice, stream_));
            VM_HIP(hipMemsetAsync(dM.p + so, 0, S * sizeof(V), stream_));
            VM_HIP(hipMemsetAsync(dV.p + so, 0, S * sizeof(V), stream_));
            VM_HIP(hipMemsetAsync(dGrad.p + 2 * so, 0, 2 * S * sizeof(vmk::GradT<R>), stream_));
        }
        if (T) {
            VM_HIP(hipMemcpyAsync(dTpos.p + to, p.tgt.data(), T * sizeof(V), hipMemcpyHostToDevice, stream_));
            VM_HIP(hipMemcpyAsync(dTslot.p + to, p.tslot.data(), T * sizeof(int), hipMemcpyHostToDevice, stream_));
        }
        VM_HIP(hipMemsetAsync(dCounts.p + co, 0, p.cells * sizeof(int), stream_));
        p.in = in; // the record and a zero state are uploaded from the kept request
        VM_HIP(hipMemcpyAsync(dInst.p + slot, &p.in, sizeof(p.in), hipMemcpyHostToDevice, stream_));
        VM_HIP(hipMemcpyAsync(dState.p + slot, &zeroState_, sizeof(zeroState_), hipMemcpyHostToDevice, stream_));
        pending_.push_back(std::move(pp));
        newWork_.insert(newWork_.end(), si.sw.begin(), si.sw.end());
        slotReq_[slot] = req;
        residentBytes_ += bytes;
        ++resident_;
        maxResident_ = std::max(maxResident_, resident_);
        peakBytes_   = std::max(peakBytes_, residentBytes_);
        changed_     = true;
        return true;
    }

    // Uploads the work tables after admissions/retirements and bins new instances.
    void refresh()
    {
        if (!changed_) return;
        std::vector<vmk::SiteWork>   sw;
        std::vector<vmk::TargetWork> tw;
        std::vector<vmk::ScanTile>   tiles;
        int                          nDyn = 0;
        // Scan tiles: multi-tile instances first (look-back), then single-tile ones.
        for (int pass = 0; pass < 2; ++pass)
            for (int s = 0; s < slots_; ++s) {
                if (slotReq_[s] < 0) continue;
                const int nt = vmk::scanTilesFor(slotInfo_[s].C);
                if ((nt > 1) != (pass == 0)) continue;
                for (int k = 0; k < nt; ++k) {
                    const int b = k * vmk::kScanTileCells;
                    tiles.push_back({s, b, int(std::min<int64_t>(vmk::kScanTileCells, slotInfo_[s].C - b)), k == 0});
                }
                if (pass == 0) nDyn += nt;
            }
        for (int s = 0; s < slots_; ++s) {
            if (slotReq_[s] < 0) continue;
            sw.insert(sw.end(), slotInfo_[s].sw.begin(), slotInfo_[s].sw.end());
            tw.insert(tw.end(), slotInfo_[s].tw.begin(), slotInfo_[s].tw.end());
        }
        if (tw.size() > twCap_) {
            twCap_ = std::max(tw.size() * 3 / 2, size_t(4096));
            dTargetWork.release();
            dTargetWork.reserve(twCap_);
            if (graph_) (void)hipGraphExecDestroy(graph_), graph_ = nullptr;
        }
        int cnt[3] = {0, int(sw.size()), int(tw.size())};
        const int nTiles = int(tiles.size());
        if (nTiles > maxTiles_) throw VmError(VM_ESTATE, "scan tile table overflow");
        const int tileCounts[2] = {nTiles, nDyn};
        // The tables go up through one pinned staging buffer (DMA at full PCIe
        // rate; a pageable copy is staged by the runtime on the host thread).
        const size_t bT = tiles.size() * sizeof(tiles[0]), bS = sw.size() * sizeof(sw[0]);
        const size_t bW = tw.size() * sizeof(tw[0]), bytes = bT + bS + bW + sizeof(tileCounts) + sizeof(cnt);
        stage_.reserve(bytes);
        unsigned char *h = stage_.p;
        std::memcpy(h, tiles.data(), bT), std::memcpy(h + bT, sw.data(), bS), std::memcpy(h + bT + bS, tw.data(), bW);
        std::memcpy(h + bT + bS + bW, tileCounts, sizeof(tileCounts));
        std::memcpy(h + bT + bS + bW + sizeof(tileCounts), cnt, sizeof(cnt));
        if (bT) VM_HIP(hipMemcpyAsync(dTiles.p, h, bT, hipMemcpyHostToDevice, stream_));
        VM_HIP(hipMemcpyAsync(&dScanCtl.p->nTiles, h + bT + bS + bW, sizeof(tileCounts), hipMemcpyHostToDevice, stream_));
        if (bS) VM_HIP(hipMemcpyAsync(dSiteWork.p, h + bT, bS, hipMemcpyHostToDevice, stream_));
        if (bW) VM_HIP(hipMemcpyAsync(dTargetWork.p, h + bT + bS, bW, hipMemcpyHostToDevice, stream_));
        VM_HIP(hipMemcpyAsync(dCnt.p, h + bT + bS + bW + sizeof(tileCounts), sizeof(cnt), hipMemcpyHostToDevice, stream_));
        if (!newWork_.empty()) {
            int nn = int(newWork_.size());
            VM_HIP(hipMemcpyAsync(dNewWork.p, newWork_.data(), newWork_.size() * sizeof(newWork_[0]),
                                  hipMemcpyHostToDevice, stream_));
            VM_HIP(hipMemcpyAsync(dNewCnt.p, &nn, sizeof(nn), hipMemcpyHostToDevice, stream_));
            auto a = adamArgs(0);
            a.work = dNewWork.p, a.nWork = dNewCnt.p;
            VM_HIP(vmk::launchAdamBin<R>(std::min(nn, 8192), stream_, a));
        }
        VM_HIP(hipStreamSynchronize(stream_));
        pending_.clear(); // admitted requests' host images have been consumed
        newWork_.clear();
        changed_ = false;
    }

    // Enqueues `iters` iterations of every resident instance; wait() blocks
    // until they ran and returns their GPU ms.
    void launch(int iters, int K)
    {
        if (!graph_) buildGraph(K);
        VM_HIP(hipEventRecord(ev0_, stream_));
        for (int k = 0; k < iters; ++k) VM_HIP(hipGraphLaunch(graph_, stream_));
        VM_HIP(hipEventRecord(ev1_, stream_));
    }
    float wait()
    {
        VM_HIP(hipEventSynchronize(ev1_));
        float ms = 0;
        VM_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
        return ms;
    }

    // Reads states; retires finished instances.  Calls done(req, state, slot).
    template <class F> int retire(F &&done)
    {
        std::vector<vmk::InstState> st(slots_);
        VM_HIP(hipMemcpyAsync(st.data(), dState.p, slots_ * sizeof(vmk::InstState), hipMemcpyDeviceToHost, stream_));
        VM_HIP(hipStreamSynchronize(stream_));
        int n = 0;
        for (int s = 0; s < slots_; ++s) {
            if (slotReq_[s] < 0 || !st[s].done) continue;
            done(slotReq_[s], st[s], s);
            SlotInfo &si = slotInfo_[s];
            sites_.release(si.so, si.S), targets_.release(si.to, si.T), cells_.release(si.co, si.C);
            residentBytes_ -= si.bytes;
            slotReq_[s] = -1;
            --resident_;
            changed_ = true;
            ++n;
        }
        return n;
    }

    void downloadSites(int slot, vm_vec2 *out)
    {
        const SlotInfo &si = slotInfo_[slot];
        std::vector<V>  h(si.S);
        if (si.S) VM_HIP(hipMemcpyAsync(h.data(), dPos.p + si.so, si.S * sizeof(V), hipMemcpyDeviceToHost, stream_));
        VM_HIP(hipStreamSynchronize(stream_));
        for (int k = 0; k < si.S; ++k) out[k] = vm_vec2{float(h[k].x), float(h[k].y)};
    }

    int resident() const { return resident_; }

    double bytesFor(int S, int T, int64_t C) const
    {
        return double(RangeAlloc::round(S)) * (4.0 * sizeof(V) + 2.0 * sizeof(vmk::GradT<R>) + 12 + 1) +
               double(RangeAlloc::round(T)) * (sizeof(V) + 4 + sizeof(vmk::ExactItem)) + double(RangeAlloc::round(C)) * 8;
    }

private:
    struct SlotInfo {
        int64_t                      so = 0, to = 0, co = 0;
        int                          S = 0, T = 0;
        int64_t                      C = 0;
        double                       bytes = 0;
        std::vector<vmk::SiteWork>   sw;
        std::vector<vmk::TargetWork> tw;
    };

    vmk::AdamArgs<R> adamArgs(int doAdam)
    {
        vmk::AdamArgs<R> a{};
        a.insts = dInst.p, a.state = dState.p, a.work = dSiteWork.p, a.nWork = dCnt.p + 1;
        a.pos = dPos.p, a.m = dM.p, a.v = dV.p, a.grad = dGrad.p, a.frozen = dFrozen.p, a.partial = dPartial.p;
        a.cellOf = dCellOf.p, a.rank = dRank.p, a.counts = dCounts.p, a.doAdam = doAdam;
        return a;
    }

    void buildGraph(int K)
    {
        const int gScan = std::min(maxTiles_, 1024), gLoss = 65536;
        const int gSite = int(std::min<int64_t>(capSites_ / vmk::kSiteChunk + slots_, 65536));
        hipGraph_t g;
        VM_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            VM_HIP(vmk::launchScan(gScan, stream_, dInst.p, dState.p, dPartial.p, dTiles.p, dScanCtl.p, dFlags.p,
                                   dCounts.p, dStart.p, vmk::kModeIterate, dExactLen.p, 1, vmk::ScanScatter{}));
            VM_HIP(vmk::launchScatter<R>(gSite, stream_, dInst.p, dState.p, dSiteWork.p, dCnt.p + 1, dPos.p,
                                         dCellOf.p, dRank.p, dStart.p, dSpos.p, dSidx.p, vmk::kModeIterate,
                                         dScanCtl.p));
            vmk::LossArgs<R> la{};
            la.insts = dInst.p, la.state = dState.p, la.work = dTargetWork.p, la.nWork = dCnt.p + 2;
            la.pos = dPos.p, la.spos = dSpos.p, la.sidx = dSidx.p, la.start = dStart.p, la.tpos = dTpos.p;
            la.tslot = dTslot.p, la.grad = dGrad.p, la.partial = dPartial.p, la.mode = vmk::kModeIterate;
            la.exactQ = dExactQ.p, la.exactLen = dExactLen.p;
            const char *ex = std::getenv("VM_LOSS_EXACT");
            la.exactOnly   = ex && ex[0] == '1';
            la.exactGrid   = vmk::exactGridFor(capTargets_, sizeof(R) == 8 || la.exactOnly);
            VM_HIP(vmk::launchLoss<R>(K, vmk::kTargetsPerLane, gLoss, stream_, la));
            VM_HIP(vmk::launchExact<R>(K, stream_, la));
            VM_HIP(vmk::launchAdamBin<R>(gSite, stream_, adamArgs(1)));
        } catch (...) {
            hipGraph_t dummy;
            (void)hipStreamEndCapture(stream_, &dummy);
            throw;
        }
        VM_HIP(hipStreamEndCapture(stream_, &g));
        hipError_t e = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        VM_HIP(e);
    }

    vm_serve_config cfg_;
    hipStream_t     stream_ = nullptr;
    hipEvent_t      ev0_ = nullptr, ev1_ = nullptr;
    hipGraphExec_t  graph_ = nullptr;
    int             slots_ = 0, resident_ = 0, maxResident_ = 0;
    int64_t         capSites_ = 0, capTargets_ = 0, capCells_ = 0;
    double          budgetBytes_ = 0, residentBytes_ = 0, peakBytes_ = 0;
    size_t          twCap_ = 0;
    bool            changed_ = false;
    RangeAlloc      sites_, targets_, cells_;
    std::vector<int>      slotReq_;
    std::vector<SlotInfo> slotInfo_;
    std::vector<vmk::SiteWork> newWork_;
    std::vector<std::unique_ptr<Prepared>> pending_;
    PinnedBuf                              stage_;
    const vmk::InstState                   zeroState_{};

    DBuf<V>              dPos, dM, dV, dSpos, dTpos;
    DBuf<vmk::GradT<R>>  dGrad;
    DBuf<int>            dCellOf, dRank, dSidx, dCounts, dStart, dTslot, dCnt, dNewCnt, dExactLen;
    DBuf<vmk::ScanTile>  dTiles;
    DBuf<vmk::ScanCtl>   dScanCtl;
    DBuf<unsigned long long> dFlags;
    int                  maxTiles_ = 0;
    DBuf<vmk::ExactItem> dExactQ;
    DBuf<uint8_t>        dFrozen;
    DBuf<vmk::Inst>      dInst;
    DBuf<vmk::InstState> dState;
    DBuf<vmk::LossPartial> dPartial;
    DBuf<vmk::SiteWork>  dSiteWork, dNewWork;
    DBuf<vmk::TargetWork> dTargetWork;
};

template <class R>
void serve(const vm_fit_desc *reqs, const double *arrival, int n, const vm_serve_config &cfg, vm_serve_result *res,
           vm_serve_stats *stats)
{
    // Sizes of every request (for pool sizing and K).
    int     K = 0;
    int64_t needS = 0, needT = 0, needC = 0;
    for (int i = 0; i < n; ++i) {
        int k = reqs[i].loss_k > 0 ? reqs[i].loss_k : 16;
        if (K == 0) K = k;
        if (k != K) throw VmError(VM_EINVAL, "all requests of one serving run must share loss_k");
        if (i > 0 && arrival[i] < arrival[i - 1]) throw VmError(VM_EINVAL, "arrival times must be non-decreasing");
        needS += RangeAlloc::round(reqs[i].num_sites);
        needT += RangeAlloc::round(reqs[i].num_targets);
        float cell = reqs[i].grid_cell, ce;
        int   rs;
        if (!(cell > 0.0f))
            vmrt::tunedChoice(std::max(reqs[i].num_sites, 1), reqs[i].width, reqs[i].height, k, &cell, &rs, &ce, nullptr);
        needC += RangeAlloc::round(int64_t(std::ceil(reqs[i].width / cell)) * int64_t(std::ceil(reqs[i].height / cell)) + 1);
    }
    Server<R> srv(cfg, n, needS, needT, needC);

    // Preparation workers: request i is prepared once it has arrived.
    std::vector<std::unique_ptr<Prepared>> prep(n);
    std::vector<std::atomic<int>>          ready(n);
    for (auto &r : ready) r.store(0);
    std::atomic<int>  nextPrep{0};
    std::atomic<bool> stop{false};
    std::mutex        errMu;
    std::string       prepError;
    const auto        t0 = Clock::now();
    auto elapsed = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> workers;
    for (unsigned w = 0; w < hw; ++w)
        workers.emplace_back([&] {
            for (;;) {
                int i = nextPrep.fetch_add(1);
                if (i >= n || stop.load()) return;
                while (!stop.load() && elapsed() < arrival[i])
                    std::this_thread::sleep_for(std::chrono::microseconds(200));
                if (stop.load()) return;
                try {
                    auto p  = std::make_unique<Prepared>();
                    p->spec = vmrt::specFromDesc(reqs[i]);
                    p->spec.maxIter = cfg.max_iterations > 0 ? cfg.max_iterations : reqs[i].max_iterations;
                    p->spec.tol     = cfg.tol > 0 ? cfg.tol : reqs[i].tol;
                    // One loss launch serves every resident request: one k, one chunk shape.
                    p->spec.K   = K;
                    p->spec.tpl = vmk::kTargetsPerLane;
                    vmrt::layoutInstance(p->spec, -1, 0, 0, 0, p->in, p->sw, p->tw, p->perm);
                    p->cells = int64_t(p->in.gw) * p->in.gh + 1;
                    Server<R>::buildImages(*p);
                    prep[i]  = std::move(p);
                    ready[i].store(1);
                } catch (const std::exception &e) {
                    std::lock_guard<std::mutex> lk(errMu);
                    prepError = e.what();
                    ready[i].store(2);
                }
            }
        });
    struct Join {
        std::vector<std::thread> &w;
        std::atomic<bool>        &stop;
        ~Join()
        {
            stop.store(true);
            for (auto &t : w)
                if (t.joinable()) t.join();
        }
    } join{workers, stop};

    std::vector<double> admitT(n, -1.0);
    int    nextAdmit = 0, completed = 0;
    double siteIters = 0;
    const int poll = std::max(1, cfg.poll_every);
    while (completed < n) {
        // The next iterations of the resident instances run while this thread
        // admits arrived requests (their uploads queue behind the iterations).
        const bool running = srv.resident() > 0;
        if (running) srv.launch(poll, K);
        // Admit prepared requests in arrival order.
        while (nextAdmit < n && ready[nextAdmit].load() != 0) {
            if (ready[nextAdmit].load() == 2) {
                std::string msg;
                {
                    std::lock_guard<std::mutex> lk(errMu);
                    msg = prepError;
                }
                throw VmError(VM_EINVAL, "request preparation failed: " + msg);
            }
            if (!srv.admit(nextAdmit, prep[nextAdmit])) {
                if (srv.resident() == 0) throw VmError(VM_ENOMEM, "request does not fit the HBM budget");
                break;
            }
            admitT[nextAdmit] = elapsed();
            ++nextAdmit;
        }
        if (!running) {
            srv.refresh();
            if (srv.resident() == 0) std::this_thread::sleep_for(std::chrono::microseconds(100));
            continue;
        }
        srv.wait();
        srv.retire([&](int req, const vmk::InstState &st, int slot) {
            double now = elapsed();
            vm_serve_result &r = res[req];
            r.latency_s        = now - arrival[req];
            r.queue_s          = admitT[req] - arrival[req];
            r.iterations       = st.iter;
            r.final_mean_loss  = st.count > 0 ? st.loss / double(st.count) : 0.0;
            siteIters += double(reqs[req].num_sites) * st.iter;
            if (r.final_sites) srv.downloadSites(slot, r.final_sites);
            ++completed;
        });
        srv.refresh();
    }
    if (stats) {
        stats->wall_s            = elapsed();
        stats->site_iters        = siteIters;
        stats->completed         = completed;
        stats->max_resident_seen = srv.maxResident();
        stats->peak_bytes        = srv.peakBytes();
    }
}

} // namespace

extern "C" int vm_serve(const vm_fit_desc *requests, const double *arrival, int n, const vm_serve_config *cfg,
                        vm_serve_result *results, vm_serve_stats *stats)
{
    return vmrt::guardedCall([&] {
        if (n < 0 || (n > 0 && (!requests || !arrival || !results)) || !cfg)
            throw VmError(VM_EINVAL, "invalid arguments");
        if (cfg->hbm_budget_bytes <= 0) throw VmError(VM_EINVAL, "hbm budget must be positive");
        if (n == 0) return;
        if (cfg->precision == 1) serve<double>(requests, arrival, n, *cfg, results, stats);
        else if (cfg->precision == 0) serve<float>(requests, arrival, n, *cfg, results, stats);
        else throw VmError(VM_EINVAL, "precision must be 0 or 1");
    });
}
