// vm_engine.cpp -- pooled layout, work tables and launch sequencing for a batch
// of VoroMesh instances (see vm_engine.h).
#include "vm_engine.h"
#include "vm_tune_table.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>

namespace vmrt {

void hipCheck(hipError_t e, const char *what, const char *file, int line)
{
    if (e == hipSuccess) return;
    throw VmError(e == hipErrorOutOfMemory ? 3 : 2, std::string(what) + ": " + hipGetErrorString(e) + " (" + file +
                                                        ":" + std::to_string(line) + ")");
}

namespace {
thread_local std::string t_lastError;
}

int guardedCall(const std::function<void()> &f)
{
    try {
        f();
        return 0;
    } catch (const VmError &e) {
        t_lastError = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        t_lastError = "host allocation failed";
        return 3;
    } catch (const std::exception &e) {
        t_lastError = e.what();
        return 1;
    }
}

const char *lastError() { return t_lastError.c_str(); }

static const TuneRow &tuneRow(int numSites)
{
    const int lg = int(std::floor(std::log2(double(std::max(numSites, 1)))));
    return kTuneTable[std::min(std::max(lg, kTuneLog2Min), kTuneLog2Max) - kTuneLog2Min];
}

int tunedK(int numSites) { return tuneRow(numSites).k; }

void tunedChoice(int numSites, float W, float H, int K, float *cell, int *stageR, float *chunkExt, int *tpl)
{
    // Mean site spacing a = sqrt(area / S).  The loss kernel's fast path needs
    // two staged rings to cover the K-NN radius (~a sqrt(K/pi)) plus the chunk
    // extent: the base cell is max(1.5, (sqrt(K/pi) + 0.8) / 2) spacings, scaled
    // by the measured factor of the size class.
    const double a   = std::sqrt(double(W) * double(H) / double(std::max(numSites, 1)));
    const double f   = std::max(1.5, 0.5 * (std::sqrt(double(K) / 3.141592653589793) + 0.8));
    const auto  &row = tuneRow(numSites);
    *cell            = float(f * row.cellScale * a);
    *stageR          = 2;
    *chunkExt        = row.chunkExtent;
    if (tpl) *tpl = vmk::lossTplSupported(row.tpl, false) ? row.tpl : vmk::kTargetsPerLane;
}

namespace {

inline int cellCoordF(float v, float inv, int n)
{
    float f = v * inv;
    int   c = (f >= 0.0f) ? (f < 2147483520.0f ? int(f) : n - 1) : 0;
    return std::min(std::max(c, 0), n - 1);
}

inline int64_t align64(int64_t x) { return (x + 63) & ~int64_t(63); }

} // namespace

// ceil(log2(x)) for x > 0, exact (frexp gives x = m 2^e with m in [0.5, 1)).
static int ceilLog2(double x)
{
    int    e;
    double m = std::frexp(x, &e);
    return m == 0.5 ? e - 1 : e;
}

int gradShiftFor(int numTargets, float ext)
{
    // A site collects at most T terms of magnitude ~ few * ext, so sums stay
    // below 2^55 with this shift; clamped to [3, 47] (each scaled term is
    // clamped to 2^51, the range of the kernels' one-addition rounding).
    int g = 53 - ceilLog2(double(std::max(numTargets, 1))) - ceilLog2(double(std::max(ext, 1e-30f)));
    return std::min(std::max(g, 3), 47);
}

void layoutInstance(const InstSpec &s, int id, int64_t siteOff, int64_t cellOff, int64_t targetOff, vmk::Inst &in,
                    std::vector<vmk::SiteWork> &siteWork, std::vector<vmk::TargetWork> &targetWork,
                    std::vector<int> &perm, float extentCap)
{
    std::memset(&in, 0, sizeof(in));
    const int S = int(s.pos.size() / 2), T = int(s.tgt.size() / 2);
    if (S >= vmk::kMaxSitesPerInst) throw VmError(1, "instance has too many sites");
    if (!vmk::lossKSupported(s.K)) throw VmError(1, "unsupported loss_k " + std::to_string(s.K));
    if (s.tpl < 1 || s.tpl > vmk::kMaxTargetsPerLane) throw VmError(1, "unsupported targets per lane");
    const int chunkMax = 64 * s.tpl;
    if (!(s.cell > 0.0f)) throw VmError(1, "grid cell must be positive");
    in.W = s.W, in.H = s.H;
    in.cell = s.cell, in.inv = 1.0f / s.cell, in.ext = std::max(s.W, s.H);
    double gwd = std::ceil(double(s.W) / s.cell), ghd = std::ceil(double(s.H) / s.cell);
    if (gwd * ghd > double(1 << 28)) throw VmError(1, "hash grid too large");
    in.gw = std::max(1, int(gwd)), in.gh = std::max(1, int(ghd));
    in.S = S, in.T = T, in.K = s.K;
    in.stageR     = std::min(std::max(s.stageR, 0), 2);
    in.allMovable = s.frozen.empty() && s.rectW <= 0 ? 1 : 0;
    in.rectW      = std::max(s.rectW, 0);
    in.rx0 = s.rect[0], in.ry0 = s.rect[1], in.rx1 = s.rect[2], in.ry1 = s.rect[3];
    in.siteOff = siteOff, in.cellOff = cellOff, in.targetOff = targetOff;
    in.lr = s.lr, in.gamma = s.gamma, in.beta1 = s.beta1, in.beta2 = s.beta2, in.eps = s.eps;
    for (int k = 0; k < 4; ++k) in.milestones[k] = s.milestones[k];
    in.fraction = s.fraction, in.seed = s.seed;
    in.useBox = s.useBox, in.bx0 = s.box[0], in.by0 = s.box[1], in.bx1 = s.box[2], in.by1 = s.box[3];
    in.maxIter = s.maxIter, in.tol = s.tol;
    in.gradShift = gradShiftFor(T, in.ext);

    for (int b = 0; b < S; b += vmk::kSiteChunk)
        siteWork.push_back({id, b, std::min(vmk::kSiteChunk, S - b), 0, siteOff, cellOff});

    // Targets sorted by a Morton key over the grid at 1/16 of a cell (the bits
    // of the fine x and y coordinates interleaved), input order breaking ties.
    // A chunk (one wavefront) takes consecutive targets of that order, at most
    // 64 * tpl, while its bounding-box diagonal stays within the chunk extent
    // (in mean spacings).  Chunks cross cell boundaries: the loss kernel stages
    // the cells around the one that holds the chunk's centre and checks that
    // they cover every target's candidate disc.
    auto spread32 = [](uint64_t v) { // bits 0..31 of v to the even bits
        v &= 0xFFFFFFFFull;
        v = (v | (v << 16)) & 0x0000FFFF0000FFFFull;
        v = (v | (v << 8)) & 0x00FF00FF00FF00FFull;
        v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0Full;
        v = (v | (v << 2)) & 0x3333333333333333ull;
        v = (v | (v << 1)) & 0x5555555555555555ull;
        return v;
    };
    const size_t          nT = static_cast<size_t>(T);
    std::vector<uint64_t> key(nT), key2(nT);
    for (int t = 0; t < T; ++t) {
        float fx = float(s.tgt[2 * t]) * in.inv, fy = float(s.tgt[2 * t + 1]) * in.inv;
        int   ccx = cellCoordF(float(s.tgt[2 * t]), in.inv, in.gw);
        int   ccy = cellCoordF(float(s.tgt[2 * t + 1]), in.inv, in.gh);
        int   sx  = std::min(std::max(int((fx - float(ccx)) * 16.0f), 0), 15);
        int   sy  = std::min(std::max(int((fy - float(ccy)) * 16.0f), 0), 15);
        key[size_t(t)] = spread32(uint64_t(ccx) * 16 + uint64_t(sx)) | (spread32(uint64_t(ccy) * 16 + uint64_t(sy)) << 1);
    }
    // LSD radix sort on the key, 12-bit digits: the order of a stable
    // comparison sort.
    perm.resize(nT);
    {
        const uint64_t maxKey = spread32(uint64_t(in.gw) * 16 - 1) | (spread32(uint64_t(in.gh) * 16 - 1) << 1);
        int            bits   = 1;
        while (bits < 64 && (maxKey >> bits)) ++bits;
        std::vector<int> idx2(nT);
        std::iota(perm.begin(), perm.end(), 0);
        constexpr int kDigit = 12, kBuckets = 1 << kDigit;
        std::vector<int> count(kBuckets);
        for (int shift = 0; shift < bits; shift += kDigit) {
            std::fill(count.begin(), count.end(), 0);
            for (int t = 0; t < T; ++t) ++count[size_t((key[size_t(t)] >> shift) & (kBuckets - 1))];
            int sum = 0;
            for (int d = 0; d < kBuckets; ++d) {
                const int c = count[size_t(d)];
                count[size_t(d)] = sum;
                sum += c;
            }
            for (int t = 0; t < T; ++t) {
                const int d = int((key[size_t(t)] >> shift) & (kBuckets - 1)), o = count[size_t(d)]++;
                key2[size_t(o)] = key[size_t(t)], idx2[size_t(o)] = perm[size_t(t)];
            }
            key.swap(key2), perm.swap(idx2);
        }
    }
    const double spacing = std::sqrt(double(s.W) * double(s.H) / double(std::max(s.liveS >= 0 ? s.liveS : S, 1)));
    const double maxDiag = double(std::min(s.chunkExt > 0.0f ? s.chunkExt : 1.6f, extentCap)) * spacing;
    // Target positions in sorted order, read sequentially by the chunk scan.
    std::vector<double> sxT(nT), syT(nT);
    for (size_t t = 0; t < nT; ++t) {
        const int o = perm[t];
        sxT[t] = s.tgt[2 * size_t(o)], syT[t] = s.tgt[2 * size_t(o) + 1];
    }
    // The loss kernel stages the block of cells within `rs` rings of the cell
    // holding the chunk centre q; it covers the candidate disc (radius D + 2 rho,
    // D = q's K-th neighbour distance, rho = the chunk's radius) when the block
    // reaches farther than that from q on every side that is not the grid's
    // edge.  `room` is that reach; D is estimated as sqrt(K / pi) spacings, the
    // radius of the disc that holds K sites at the mean density.
    const double dEst = std::sqrt(double(s.K) / 3.141592653589793) * spacing;
    // Three rings (7 x 7 cells) only where they hold at most 116 sites at the
    // mean density, about 10 % inside the 128 the kernel stages.
    const double cellSp   = double(in.cell) / spacing;
    const int    maxRings = 49.0 * cellSp * cellSp <= 116.0 ? vmk::kStageRetry : 2;
    auto         room = [&](double qx, double qy, int rs) {
        const double c  = double(in.cell);
        const int    hx = cellCoordF(float(qx), in.inv, in.gw), hy = cellCoordF(float(qy), in.inv, in.gh);
        double       r  = 1e300;
        if (hx - rs > 0) r = std::min(r, qx - double(hx - rs) * c);
        if (hx + rs < in.gw - 1) r = std::min(r, double(hx + rs + 1) * c - qx);
        if (hy - rs > 0) r = std::min(r, qy - double(hy - rs) * c);
        if (hy + rs < in.gh - 1) r = std::min(r, double(hy + rs + 1) * c - qy);
        return r;
    };
    for (int t = 0; t < T;) {
        int    e  = t;
        double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
        while (e < T && e - t < chunkMax) {
            double px = sxT[size_t(e)], py = syT[size_t(e)];
            double nx0 = std::min(x0, px), ny0 = std::min(y0, py), nx1 = std::max(x1, px), ny1 = std::max(y1, py);
            if (e > t) {
                const double dx = nx1 - nx0, dy = ny1 - ny0;
                if (dx * dx + dy * dy > maxDiag * maxDiag) break;
                // A chunk stops growing before its largest block would miss its
                // disc (with a fifth of D to spare).
                if (room(0.5 * (nx0 + nx1), 0.5 * (ny0 + ny1), maxRings) <
                    1.2 * dEst + std::sqrt(dx * dx + dy * dy))
                    break;
            }
            x0 = nx0, y0 = ny0, x1 = nx1, y1 = ny1;
            ++e;
        }
        // Chunk centre (FP32) and an upper bound on the distance of its targets to it.
        float  qx = float(0.5 * (x0 + x1)), qy = float(0.5 * (y0 + y1));
        double rho = 0.0;
        for (int k = t; k < e; ++k) {
            const double dx = sxT[size_t(k)] - double(qx), dy = syT[size_t(k)] - double(qy);
            rho = std::max(rho, dx * dx + dy * dy);
        }
        rho = std::sqrt(rho); // double: within an ulp, and rounded up twice in float below
        float rhof = std::nextafter(std::nextafter(float(rho), 1e30f), 1e30f);
        vmk::TargetWork tw{};
        tw.inst = id, tw.count = e - t;
        tw.cx = cellCoordF(qx, in.inv, in.gw), tw.cy = cellCoordF(qy, in.inv, in.gh);
        tw.qx = qx, tw.qy = qy, tw.rho = rhof, tw.gw = in.gw, tw.gh = in.gh;
        // Two rings, or three where two would likely miss the disc (the kernel
        // retries with the other count when its choice fails).
        const int rings = in.stageR == 2 && maxRings > 2 && room(qx, qy, 2) < dEst + 2.0 * rho ? maxRings : in.stageR;
        tw.flags   = (rings & vmk::kWorkRingsMask) | (in.fraction < 1.0f ? vmk::kWorkSubsample : 0);
        tw.tBegin  = targetOff + t, tw.cellOff = cellOff, tw.siteOff = siteOff;
        targetWork.push_back(tw);
        // Lane l takes the chunk's targets l, l + 64, l + 128, ...  Interleave the
        // Morton order with stride H = ceil(n / 64), so the 64 targets of one
        // pass are spread over the chunk and fewer lanes of a pass add their
        // gradient to the same site (LDS atomics on one address serialize).
        const int n = e - t, H = (n + 63) / 64;
        if (H > 1) {
            std::vector<int> run(perm.begin() + t, perm.begin() + e);
            int              j = t;
            for (int r0 = 0; r0 < H; ++r0)
                for (int r = r0; r < n; r += H) perm[j++] = run[r];
        }
        t = e;
    }
    // The chunks follow the Morton order of their targets: runs of consecutive
    // chunks, which the loss kernel hands to one XCD, cover compact patches of
    // the grid.
}

namespace {

constexpr int kMaxLossGrid = 131072; // wavefronts (measured on the batch: 0.4 % faster than 65,536, 2.7 % faster than one per chunk)
constexpr int kMaxSiteGrid = 65536; // one site chunk per workgroup up to 16.7 M sites
constexpr int kMaxScanGrid = 4096;

template <class R> class EngineT final : public Engine {
    using V = vmk::V2<R>;

public:
    explicit EngineT(hipStream_t s) { stream = s; }
    ~EngineT() override
    {
        destroyGraph();
        if (!side.empty()) (void)hipStreamSynchronize(stream);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e), e = nullptr;
        for (auto &e : evJoin) (void)hipEventDestroy(e);
        if (evFork) (void)hipEventDestroy(evFork);
        for (auto &sd : side) (void)hipStreamDestroy(sd);
    }

    // ---- layout ---------------------------------------------------------------
    struct Layout {
        std::vector<vmk::Inst>       inst;
        std::vector<vmk::SiteWork>   siteWork;
        std::vector<vmk::TargetWork> targetWork; // grouped by K
        struct Group { int tpl, begin, count; int64_t qOff, qLen; }; // qOff/qLen: the group's exact-selection queue
        std::vector<Group>            groups;
        std::vector<vmk::ScanTile>    tiles; // scan tiles, instance by instance
        std::vector<std::vector<int>> perm; // per instance: sorted position -> input target index
        int64_t sites = 0, cells = 0, targets = 0, queue = 0;
        int     dynTiles = 0;
    } L;

    DBuf<V>              dPos, dM, dV, dSpos, dTpos;
    DBuf<vmk::GradT<R>>  dGrad; // two per site
    DBuf<int>            dCellOf, dRank, dSidx, dCounts, dStart, dTslot, dCnt;
    DBuf<vmk::ScanTile>  dTiles;
    DBuf<vmk::ScanCtl>   dScanCtl;
    DBuf<unsigned long long> dFlags; // one look-back flag per scan tile
    DBuf<uint8_t>        dFrozen;
    DBuf<vmk::Inst>      dInst;
    DBuf<vmk::InstState> dState;
    DBuf<vmk::LossPartial> dPartial;
    DBuf<vmk::SiteWork>  dSiteWork;
    DBuf<vmk::TargetWork> dTargetWork;
    DBuf<vmk::ExactItem> dExactQ;
    DBuf<int>            dExactLen; // one per group
    hipGraphExec_t       graph = nullptr, graphN = nullptr; // one iteration, kGraphIters iterations
    bool                 stale1 = false, staleN = false;   // a commit changed their launches
    hipEvent_t           ev[8] = {};
    // Loss groups after the first run on side streams forked from the main
    // stream (and joined back before Adam), so their launches overlap.
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t>  evJoin;
    hipEvent_t               evFork = nullptr;
    int                  gridScan = 1, gridSite = 1;
    std::vector<int>     gridLoss;

    void destroyGraph()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (graphN) (void)hipGraphExecDestroy(graphN);
        graph = graphN = nullptr;
    }

    void ensureEvents()
    {
        for (auto &e : ev)
            if (!e) VM_HIP(hipEventCreate(&e));
    }

    size_t deviceBytes() const override
    {
        size_t b = 0;
        b += (dPos.n + dM.n + dV.n + dSpos.n + dTpos.n) * sizeof(V) + dGrad.n * sizeof(vmk::GradT<R>);
        b += (dCellOf.n + dRank.n + dSidx.n + dCounts.n + dStart.n + dTslot.n + dCnt.n) * sizeof(int);
        b += dTiles.n * sizeof(vmk::ScanTile) + dScanCtl.n * sizeof(vmk::ScanCtl) + dFlags.n * sizeof(unsigned long long);
        b += dFrozen.n + dInst.n * sizeof(vmk::Inst) + dState.n * sizeof(vmk::InstState) +
             dPartial.n * sizeof(vmk::LossPartial);
        b += dSiteWork.n * sizeof(vmk::SiteWork) + dTargetWork.n * sizeof(vmk::TargetWork);
        b += dExactQ.n * sizeof(vmk::ExactItem) + dExactLen.n * sizeof(int);
        return b;
    }

    void buildLayout()
    {
        L = Layout();
        const int n = int(specs.size());
        // Chunk shapes compiled for this precision (FP64 sends every target to
        // the exact selection and has one).
        for (auto &s : specs)
            if (!vmk::lossTplSupported(s.tpl, sizeof(R) == 8)) s.tpl = vmk::kTargetsPerLane;
        L.inst.resize(n);
        L.perm.resize(n);
        std::vector<std::vector<vmk::TargetWork>> twByInst(n);
        for (float cap : {1e30f, kSmallLayoutExtent}) {
            extentCap_ = cap;
            L.sites = L.cells = L.targets = 0;
            L.siteWork.clear();
            size_t chunks = 0;
            for (int i = 0; i < n; ++i) {
                InstSpec  &s  = specs[i];
                vmk::Inst &in = L.inst[i];
                twByInst[i].clear();
                layoutInstance(s, i, L.sites, L.cells, L.targets, in, L.siteWork, twByInst[i], L.perm[i], cap);
                L.sites += align64(in.S);
                L.cells += align64(int64_t(in.gw) * in.gh + 1);
                L.targets += align64(in.T);
                chunks += twByInst[i].size();
            }
            if (chunks >= size_t(kMinLayoutChunks)) break;
        }
        // A single instance may keep room past its targets for local edits.
        tailBase_ = L.targets, tailUsed_ = 0;
        if (n == 1 && tailTargets > 0) L.targets += align64(tailTargets);
        // Scan tiles: the tiles of multi-tile instances first (look-back), then
        // the single-tile instances.
        for (int pass = 0; pass < 2; ++pass)
            for (int i = 0; i < n; ++i) {
                const int64_t cells = int64_t(L.inst[i].gw) * L.inst[i].gh + 1;
                const int     nt    = vmk::scanTilesFor(cells);
                if ((nt > 1) != (pass == 0)) continue;
                for (int k = 0; k < nt; ++k) {
                    const int b = k * vmk::kScanTileCells;
                    L.tiles.push_back({i, b, int(std::min<int64_t>(vmk::kScanTileCells, cells - b)), k == 0});
                }
                if (pass == 0) L.dynTiles += nt;
            }
        // Group target work by targets per lane: one loss launch per group (the
        // kernel reads each fit's k from its record, so fits of every k share one).
        std::vector<int> tpls;
        for (auto &s : specs) tpls.push_back(s.tpl);
        std::sort(tpls.begin(), tpls.end());
        tpls.erase(std::unique(tpls.begin(), tpls.end()), tpls.end());
        for (int tpl : tpls) {
            typename Layout::Group g{tpl, int(L.targetWork.size()), 0, L.queue, 0};
            for (int i = 0; i < n; ++i)
                if (specs[i].tpl == tpl) {
                    L.targetWork.insert(L.targetWork.end(), twByInst[i].begin(), twByInst[i].end());
                    L.queue += L.inst[i].T; // every target of the group may be queued once per launch
                }
            g.count = int(L.targetWork.size()) - g.begin;
            if (n == 1 && tailTargets > 0) L.queue += align64(tailTargets);
            g.qLen = L.queue - g.qOff;
            L.groups.push_back(g);
        }
    }

    void commit() override
    {
        stale1 = staleN = true; // the graphs take the new launches before their next replay
        buildLayout();
        const int n = int(specs.size());
        dPos.reserve(L.sites), dM.reserve(L.sites), dV.reserve(L.sites), dGrad.reserve(2 * L.sites);
        dSpos.reserve(L.sites), dCellOf.reserve(L.sites), dRank.reserve(L.sites), dSidx.reserve(L.sites);
        dFrozen.reserve(L.sites);
        dCounts.reserve(L.cells), dStart.reserve(L.cells);
        dTpos.reserve(L.targets), dTslot.reserve(L.targets);
        dInst.reserve(n), dState.reserve(n), dPartial.reserve(size_t(n) * vmk::kLossPartials);
        dTiles.reserve(L.tiles.size()), dScanCtl.reserve(1), dFlags.reserve(L.tiles.size());
        const bool tail = n == 1 && tailTargets > 0;
        dSiteWork.reserve(L.siteWork.size()), dTargetWork.reserve(L.targetWork.size() + (tail ? size_t(tailWork) : 0));
        dCnt.reserve(2 + L.groups.size());
        dExactQ.reserve(std::max<int64_t>(L.queue, 1)), dExactLen.reserve(L.groups.size());

        std::vector<V>       hp(L.sites), hm(L.sites), hv(L.sites), ht(L.targets, V{R(-1), R(-1)});
        std::vector<int>     hts(L.targets, -1);
        std::vector<uint8_t> hf(L.sites, 0);
        std::vector<vmk::InstState> hs(n);
        for (int i = 0; i < n; ++i) {
            const InstSpec  &s  = specs[i];
            const vmk::Inst &in = L.inst[i];
            for (int k = 0; k < in.S; ++k) {
                hp[in.siteOff + k] = V{R(s.pos[2 * k]), R(s.pos[2 * k + 1])};
                hm[in.siteOff + k] = s.m.empty() ? V{0, 0} : V{R(s.m[2 * k]), R(s.m[2 * k + 1])};
                hv[in.siteOff + k] = s.v.empty() ? V{0, 0} : V{R(s.v[2 * k]), R(s.v[2 * k + 1])};
                if (!s.frozen.empty()) hf[in.siteOff + k] = s.frozen[k];
            }
            for (int t = 0; t < in.T; ++t) {
                int o = L.perm[i][t];
                ht[in.targetOff + t]  = V{R(s.tgt[2 * o]), R(s.tgt[2 * o + 1])};
                hts[in.targetOff + t] = s.tslot.empty() ? o : s.tslot[o];
            }
            hs[i] = s.state;
        }
        std::vector<int> cnt(2 + L.groups.size());
        cnt[0] = n, cnt[1] = int(L.siteWork.size());
        for (size_t g = 0; g < L.groups.size(); ++g) cnt[2 + g] = L.groups[g].count;

        // The host tables go up through one pinned staging buffer when they are
        // small (each copy a DMA; a pageable copy is staged by the runtime on the
        // host thread, which dominates a local edit's commit), else directly.
        struct Up {
            void       *dst;
            const void *src;
            size_t      bytes, off;
        };
        std::vector<Up> ups;
        size_t          total = 0;
        auto up = [&](auto *dst, const auto &src) {
            if (src.empty()) return;
            const size_t b = src.size() * sizeof(src[0]);
            ups.push_back({dst, src.data(), b, total});
            total += (b + 255) & ~size_t(255);
        };
        up(dPos.p, hp), up(dM.p, hm), up(dV.p, hv), up(dTpos.p, ht), up(dTslot.p, hts), up(dFrozen.p, hf);
        up(dInst.p, L.inst), up(dState.p, hs), up(dTiles.p, L.tiles), up(dSiteWork.p, L.siteWork);
        const std::vector<vmk::ScanCtl> ctl{{0, 0u, int(L.tiles.size()), L.dynTiles}};
        up(dScanCtl.p, ctl);
        up(dTargetWork.p, L.targetWork), up(dCnt.p, cnt);
        const bool staged = total <= kPinnedUploadMax;
        if (staged) hUpload.reserve(total);
        for (const Up &u : ups) {
            const void *src = u.src;
            if (staged) {
                std::memcpy(hUpload.p + u.off, u.src, u.bytes);
                src = hUpload.p + u.off;
            }
            VM_HIP(hipMemcpyAsync(u.dst, src, u.bytes, hipMemcpyHostToDevice, stream));
        }
        VM_HIP(hipMemsetAsync(dFlags.p, 0, dFlags.n * sizeof(unsigned long long), stream));
        VM_HIP(hipMemsetAsync(dGrad.p, 0, dGrad.n * sizeof(vmk::GradT<R>), stream));
        VM_HIP(hipMemsetAsync(dPartial.p, 0, dPartial.n * sizeof(vmk::LossPartial), stream));
        VM_HIP(vmk::launchFillInt(stream, dCounts.p, int64_t(dCounts.n), 0));
        gridScan = std::max(1, std::min(int(L.tiles.size()), kMaxScanGrid));
        gridSite = std::max(1, std::min(int(L.siteWork.size()), kMaxSiteGrid));
        gridLoss.clear();
        for (auto &g : L.groups) gridLoss.push_back(std::max(1, std::min(g.count + (tail ? tailWork : 0), kMaxLossGrid)));
        bin();
        VM_HIP(hipStreamSynchronize(stream));
        dirty       = false;
        deviceNewer = false;
    }

    // ---- local edits of instance 0 ---------------------------------------------
    const std::vector<int> &permOf(int id) const override { return L.perm[size_t(id)]; }

    void killTargets(const std::vector<int64_t> &pos) override
    {
        if (pos.empty()) return;
        const size_t bytes = pos.size() * sizeof(int64_t);
        hPatch.reserve(bytes);
        std::memcpy(hPatch.p, pos.data(), bytes);
        dPatchIdx.reserve(pos.size());
        VM_HIP(hipMemcpyAsync(dPatchIdx.p, hPatch.p, bytes, hipMemcpyHostToDevice, stream));
        VM_HIP(vmk::launchKillTargets<R>(stream, dTpos.p, dTslot.p, dPatchIdx.p, int(pos.size())));
        VM_HIP(hipStreamSynchronize(stream));
    }

    bool appendTargets(const std::vector<double> &tgt, const std::vector<int> &tslot,
                       std::vector<int64_t> &pos) override
    {
        const int64_t n = int64_t(tgt.size() / 2);
        pos.assign(size_t(n), -1);
        if (n == 0) return true;
        if (dirty || specs.size() != 1 || L.groups.size() != 1 || tailUsed_ + n > tailTargets) return false;
        // The new targets laid out alone with instance 0's geometry and chunk
        // shape, at the next free tail position.
        const InstSpec &s = specs[0];
        InstSpec        sub;
        sub.W = s.W, sub.H = s.H, sub.K = s.K, sub.cell = s.cell, sub.stageR = s.stageR;
        sub.chunkExt = s.chunkExt, sub.tpl = s.tpl, sub.fraction = s.fraction;
        sub.liveS = s.liveS >= 0 ? s.liveS : int(s.pos.size() / 2);
        sub.tgt = tgt, sub.tslot = tslot;
        const vmk::Inst              &in = L.inst[0];
        vmk::Inst                     tin;
        std::vector<vmk::SiteWork>    sw;
        std::vector<vmk::TargetWork>  tw;
        std::vector<int>              perm;
        const int64_t                 at = tailBase_ + tailUsed_;
        layoutInstance(sub, 0, in.siteOff, in.cellOff, at, tin, sw, tw, perm, extentCap_);
        const size_t w0 = L.targetWork.size();
        if (w0 + tw.size() > dTargetWork.n) return false;
        std::vector<V>   ht(static_cast<size_t>(n));
        std::vector<int> hs(static_cast<size_t>(n));
        for (int64_t t = 0; t < n; ++t) {
            const int o = perm[size_t(t)];
            ht[size_t(t)] = V{R(tgt[2 * size_t(o)]), R(tgt[2 * size_t(o) + 1])};
            hs[size_t(t)] = tslot[size_t(o)];
            pos[size_t(o)] = at + t;
        }
        const int    cnt = int(w0 + tw.size());
        const size_t bT = ht.size() * sizeof(V), bS = hs.size() * sizeof(int), bW = tw.size() * sizeof(tw[0]);
        hPatch.reserve(bT + bS + bW + sizeof(int));
        unsigned char *h = hPatch.p;
        std::memcpy(h, ht.data(), bT), std::memcpy(h + bT, hs.data(), bS), std::memcpy(h + bT + bS, tw.data(), bW);
        std::memcpy(h + bT + bS + bW, &cnt, sizeof(int));
        VM_HIP(hipMemcpyAsync(dTpos.p + at, h, bT, hipMemcpyHostToDevice, stream));
        VM_HIP(hipMemcpyAsync(dTslot.p + at, h + bT, bS, hipMemcpyHostToDevice, stream));
        if (bW) VM_HIP(hipMemcpyAsync(dTargetWork.p + w0, h + bT + bS, bW, hipMemcpyHostToDevice, stream));
        VM_HIP(hipMemcpyAsync(dCnt.p + 2, h + bT + bS + bW, sizeof(int), hipMemcpyHostToDevice, stream));
        VM_HIP(hipStreamSynchronize(stream));
        L.targetWork.insert(L.targetWork.end(), tw.begin(), tw.end());
        L.groups[0].count = cnt;
        tailUsed_ += n;
        return true;
    }

    void uploadSiteRect(int slotsW, int x0, int y0, int x1, int y1, const float *xy, bool zeroMoments) override
    {
        if (dirty || x1 <= x0 || y1 <= y0) return;
        const vmk::Inst &in = L.inst[0];
        const int        w = x1 - x0, h = y1 - y0;
        const size_t     pitch = size_t(slotsW) * sizeof(V), row = size_t(w) * sizeof(V);
        const size_t     off   = size_t(in.siteOff) + size_t(y0) * slotsW + x0;
        if (xy) {
            hPatch.reserve(size_t(w) * h * sizeof(V));
            V *st = reinterpret_cast<V *>(hPatch.p);
            for (int y = 0; y < h; ++y)
                for (int x = 0; x < w; ++x) {
                    const size_t k = size_t(y0 + y) * slotsW + x0 + x;
                    st[size_t(y) * w + x] = V{R(xy[2 * k]), R(xy[2 * k + 1])};
                }
            VM_HIP(hipMemcpy2DAsync(dPos.p + off, pitch, st, row, row, size_t(h), hipMemcpyHostToDevice, stream));
        }
        if (zeroMoments) {
            VM_HIP(hipMemset2DAsync(dM.p + off, pitch, 0, row, size_t(h), stream));
            VM_HIP(hipMemset2DAsync(dV.p + off, pitch, 0, row, size_t(h), stream));
        }
        VM_HIP(hipStreamSynchronize(stream));
        deviceNewer = true;
    }

    void updateRecord(int id, int liveTargets, bool resetState) override
    {
        if (dirty) return;
        const InstSpec &s  = specs[size_t(id)];
        vmk::Inst      &in = L.inst[size_t(id)];
        in.lr = s.lr, in.gamma = s.gamma, in.beta1 = s.beta1, in.beta2 = s.beta2, in.eps = s.eps;
        for (int k = 0; k < 4; ++k) in.milestones[k] = s.milestones[k];
        in.useBox = s.useBox, in.bx0 = s.box[0], in.by0 = s.box[1], in.bx1 = s.box[2], in.by1 = s.box[3];
        in.allMovable = s.frozen.empty() && s.rectW <= 0 ? 1 : 0;
        in.rectW      = std::max(s.rectW, 0);
        in.rx0 = s.rect[0], in.ry0 = s.rect[1], in.rx1 = s.rect[2], in.ry1 = s.rect[3];
        in.maxIter = s.maxIter, in.tol = s.tol;
        in.gradShift = gradShiftFor(liveTargets, in.ext);
        VM_HIP(hipMemcpyAsync(dInst.p + id, &in, sizeof(in), hipMemcpyHostToDevice, stream));
        if (resetState) {
            specs[size_t(id)].state = vmk::InstState{};
            VM_HIP(hipMemcpyAsync(dState.p + id, &specs[size_t(id)].state, sizeof(vmk::InstState),
                                  hipMemcpyHostToDevice, stream));
        }
        VM_HIP(hipStreamSynchronize(stream));
    }

    void rebin() override
    {
        if (dirty) return;
        VM_HIP(vmk::launchFillInt(stream, dCounts.p, int64_t(dCounts.n), 0));
        bin();
        VM_HIP(hipStreamSynchronize(stream));
    }

    // ---- launch helpers -------------------------------------------------------
    vmk::LossArgs<R> lossArgs(int g, int mode, bool wantGrad)
    {
        vmk::LossArgs<R> a{};
        a.insts = dInst.p, a.state = dState.p, a.work = dTargetWork.p + L.groups[g].begin, a.nWork = dCnt.p + 2 + g;
        a.pos = dPos.p, a.spos = dSpos.p, a.sidx = dSidx.p, a.start = dStart.p, a.tpos = dTpos.p, a.tslot = dTslot.p;
        a.grad = wantGrad ? dGrad.p : nullptr, a.partial = dPartial.p, a.mode = mode, a.dbgInst = -1;
        // VM_LOSS_EXACT=1 sends every lane through the exact k-NN selection (parity checks of that path).
        const char *ex = std::getenv("VM_LOSS_EXACT");
        a.exactOnly    = ex && ex[0] == '1';
        a.exactQ = dExactQ.p + L.groups[g].qOff, a.exactLen = dExactLen.p + g;
        a.exactGrid = vmk::exactGridFor(L.groups[g].qLen, sizeof(R) == 8 || a.exactOnly);
        return a;
    }

    vmk::AdamArgs<R> adamArgs(int doAdam)
    {
        vmk::AdamArgs<R> a{};
        a.insts = dInst.p, a.state = dState.p, a.work = dSiteWork.p, a.nWork = dCnt.p + 1;
        a.pos = dPos.p, a.m = dM.p, a.v = dV.p, a.grad = dGrad.p, a.frozen = dFrozen.p, a.partial = dPartial.p;
        a.cellOf = dCellOf.p, a.rank = dRank.p, a.counts = dCounts.p, a.doAdam = doAdam;
        return a;
    }

    void bin() { VM_HIP(vmk::launchAdamBin<R>(gridSite, stream, adamArgs(0))); }

    // Hash-grid build: scan, then scatter (folded into the scan when no instance
    // has more than one scan tile: one launch less per iteration).
    void gridBuild(int mode)
    {
        vmk::ScanScatter sc{};
        if (L.dynTiles == 0)
            sc = vmk::ScanScatter{dPos.p, dCellOf.p, dRank.p, dSpos.p, dSidx.p, int(sizeof(V))};
        VM_HIP(vmk::launchScan(gridScan, stream, dInst.p, dState.p, dPartial.p, dTiles.p, dScanCtl.p, dFlags.p,
                               dCounts.p, dStart.p, mode, dExactLen.p, int(L.groups.size()), sc));
        if (L.dynTiles != 0)
            VM_HIP(vmk::launchScatter<R>(gridSite, stream, dInst.p, dState.p, dSiteWork.p, dCnt.p + 1, dPos.p,
                                         dCellOf.p, dRank.p, dStart.p, dSpos.p, dSidx.p, mode, dScanCtl.p));
    }

    void ensureSide(size_t groups)
    {
        while (side.size() + 1 < groups) {
            hipStream_t sd;
            VM_HIP(hipStreamCreateWithFlags(&sd, hipStreamNonBlocking));
            side.push_back(sd);
        }
        while (evJoin.size() + 1 < groups) {
            hipEvent_t e;
            VM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            evJoin.push_back(e);
        }
        if (!evFork) VM_HIP(hipEventCreateWithFlags(&evFork, hipEventDisableTiming));
    }

    void enqueueIteration(hipEvent_t *marks)
    {
        if (marks) VM_HIP(hipEventRecord(marks[0], stream));
        vmk::ScanScatter sc{};
        if (L.dynTiles == 0)
            sc = vmk::ScanScatter{dPos.p, dCellOf.p, dRank.p, dSpos.p, dSidx.p, int(sizeof(V))};
        VM_HIP(vmk::launchScan(gridScan, stream, dInst.p, dState.p, dPartial.p, dTiles.p, dScanCtl.p, dFlags.p,
                               dCounts.p, dStart.p, vmk::kModeIterate, dExactLen.p, int(L.groups.size()), sc));
        if (marks) VM_HIP(hipEventRecord(marks[1], stream));
        if (L.dynTiles != 0)
            VM_HIP(vmk::launchScatter<R>(gridSite, stream, dInst.p, dState.p, dSiteWork.p, dCnt.p + 1, dPos.p,
                                         dCellOf.p, dRank.p, dStart.p, dSpos.p, dSidx.p, vmk::kModeIterate,
                                         dScanCtl.p));
        if (marks) VM_HIP(hipEventRecord(marks[2], stream));
        if (L.groups.size() == 1) {
            VM_HIP(vmk::launchLoss<R>(L.groups[0].tpl, gridLoss[0], stream,
                                      lossArgs(0, vmk::kModeIterate, true)));
            if (marks) VM_HIP(hipEventRecord(marks[3], stream));
            VM_HIP(vmk::launchExact<R>(stream, lossArgs(0, vmk::kModeIterate, true)));
        } else if (!L.groups.empty()) {
            // Each group's loss and exact selection on its own stream (the groups
            // touch disjoint instances); the main stream waits for all of them.
            // With several groups the loss phase below covers both kernels.
            ensureSide(L.groups.size());
            VM_HIP(hipEventRecord(evFork, stream));
            for (size_t g = 0; g < L.groups.size(); ++g) {
                hipStream_t sg = g == 0 ? stream : side[g - 1];
                if (g > 0) VM_HIP(hipStreamWaitEvent(sg, evFork, 0));
                VM_HIP(vmk::launchLoss<R>(L.groups[g].tpl, gridLoss[g], sg,
                                          lossArgs(int(g), vmk::kModeIterate, true)));
                VM_HIP(vmk::launchExact<R>(sg, lossArgs(int(g), vmk::kModeIterate, true)));
                if (g > 0) VM_HIP(hipEventRecord(evJoin[g - 1], sg));
            }
            for (size_t g = 1; g < L.groups.size(); ++g) VM_HIP(hipStreamWaitEvent(stream, evJoin[g - 1], 0));
            if (marks) VM_HIP(hipEventRecord(marks[3], stream));
        }
        if (marks) VM_HIP(hipEventRecord(marks[4], stream));
        VM_HIP(vmk::launchAdamBin<R>(gridSite, stream, adamArgs(1)));
        if (marks) VM_HIP(hipEventRecord(marks[5], stream));
    }

    // Iterations per replay of the long graph: each graph launch costs about
    // 5 us on the queue, which dominates small fits.
    static constexpr int kGraphIters = 8;

    // Captures `iters` iterations.  An existing exec of the same shape (the same
    // launches, as after a local edit) takes the new parameters in place, which
    // costs less than instantiating again; otherwise it is replaced.
    void buildGraph(hipGraphExec_t &out, int iters)
    {
        hipGraph_t g;
        VM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            for (int k = 0; k < iters; ++k) enqueueIteration(nullptr);
        } catch (...) {
            hipGraph_t dummy;
            (void)hipStreamEndCapture(stream, &dummy);
            throw;
        }
        VM_HIP(hipStreamEndCapture(stream, &g));
        if (out) {
            hipGraphNode_t           errNode = nullptr;
            hipGraphExecUpdateResult res     = hipGraphExecUpdateError;
            if (hipGraphExecUpdate(out, g, &errNode, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(out);
                out = nullptr;
            } else {
                ++graphUpdates;
            }
        }
        hipError_t e = hipSuccess;
        if (!out) {
            e = hipGraphInstantiate(&out, g, nullptr, nullptr, 0);
            if (e == hipSuccess) ++graphInstantiations;
            else out = nullptr;
        }
        (void)hipGraphDestroy(g);
        VM_HIP(e);
    }

    float step(int iterations) override
    {
        if (dirty) commit();
        if (iterations <= 0) return 0.0f;
        if (!graph || stale1) buildGraph(graph, 1), stale1 = false;
        if (iterations >= kGraphIters && (!graphN || staleN)) buildGraph(graphN, kGraphIters), staleN = false;
        ensureEvents();
        VM_HIP(hipEventRecord(ev[5], stream));
        int k = 0;
        for (; k + kGraphIters <= iterations; k += kGraphIters) VM_HIP(hipGraphLaunch(graphN, stream));
        for (; k < iterations; ++k) VM_HIP(hipGraphLaunch(graph, stream));
        VM_HIP(hipEventRecord(ev[6], stream));
        VM_HIP(hipEventSynchronize(ev[6]));
        float ms = 0;
        VM_HIP(hipEventElapsedTime(&ms, ev[5], ev[6]));
        deviceNewer = true;
        return ms;
    }

    void profile(int iterations, float *phaseMs) override
    {
        if (dirty) commit();
        ensureEvents();
        double acc[kPhases] = {};
        for (int k = 0; k < iterations; ++k) {
            enqueueIteration(ev);
            VM_HIP(hipEventSynchronize(ev[kPhases]));
            for (int p = 0; p < kPhases; ++p) {
                float ms = 0;
                VM_HIP(hipEventElapsedTime(&ms, ev[p], ev[p + 1]));
                acc[p] += ms;
            }
        }
        for (int p = 0; p < kPhases; ++p) phaseMs[p] = float(acc[p] / std::max(iterations, 1));
        deviceNewer = true;
    }

    void readStates() override
    {
        if (dirty) return;
        std::vector<vmk::InstState> hs(specs.size());
        if (!hs.empty())
            VM_HIP(hipMemcpyAsync(hs.data(), dState.p, hs.size() * sizeof(vmk::InstState), hipMemcpyDeviceToHost,
                                  stream));
        VM_HIP(hipStreamSynchronize(stream));
        for (size_t i = 0; i < hs.size(); ++i) specs[i].state = hs[i];
    }

    void pull() override
    {
        if (dirty || !deviceNewer) return;
        std::vector<V> hp(L.sites), hm(L.sites), hv(L.sites);
        if (L.sites) {
            VM_HIP(hipMemcpyAsync(hp.data(), dPos.p, L.sites * sizeof(V), hipMemcpyDeviceToHost, stream));
            VM_HIP(hipMemcpyAsync(hm.data(), dM.p, L.sites * sizeof(V), hipMemcpyDeviceToHost, stream));
            VM_HIP(hipMemcpyAsync(hv.data(), dV.p, L.sites * sizeof(V), hipMemcpyDeviceToHost, stream));
        }
        readStates();
        for (size_t i = 0; i < specs.size(); ++i) {
            InstSpec        &s  = specs[i];
            const vmk::Inst &in = L.inst[i];
            s.m.resize(2 * size_t(in.S)), s.v.resize(2 * size_t(in.S));
            for (int k = 0; k < in.S; ++k) {
                s.pos[2 * k] = double(hp[in.siteOff + k].x), s.pos[2 * k + 1] = double(hp[in.siteOff + k].y);
                s.m[2 * k] = double(hm[in.siteOff + k].x), s.m[2 * k + 1] = double(hm[in.siteOff + k].y);
                s.v[2 * k] = double(hv[in.siteOff + k].x), s.v[2 * k + 1] = double(hv[in.siteOff + k].y);
            }
        }
        deviceNewer = false;
    }

    double evaluate(int id, int64_t *count) override
    {
        if (dirty) commit();
        gridBuild(vmk::kModeEvaluate);
        for (size_t g = 0; g < L.groups.size(); ++g)
        {
            VM_HIP(vmk::launchLoss<R>(L.groups[g].tpl, gridLoss[g], stream, lossArgs(int(g), vmk::kModeEvaluate, false)));
            VM_HIP(vmk::launchExact<R>(stream, lossArgs(int(g), vmk::kModeEvaluate, false)));
        }
        bin();
        vmk::InstState st;
        VM_HIP(hipMemcpyAsync(&st, dState.p + id, sizeof(st), hipMemcpyDeviceToHost, stream));
        VM_HIP(hipStreamSynchronize(stream));
        if (count) *count = st.evalCount;
        return st.evalLoss;
    }

    void diag(int id, bool wantGrad, bool wantDbg, DiagOut &out) override
    {
        if (dirty) commit();
        const vmk::Inst &in = L.inst[id];
        DBuf<int>          dK;
        DBuf<vmk::TermDbg> dT;
        int                K = in.K;
        if (wantDbg) {
            dK.reserve(std::max<size_t>(size_t(in.T) * K, 1));
            dT.reserve(size_t(L.targets));
        }
        gridBuild(vmk::kModeEvaluate);
        for (size_t g = 0; g < L.groups.size(); ++g) {
            auto a = lossArgs(int(g), vmk::kModeEvaluate, wantGrad);
            if (wantDbg) a.dbgKnn = dK.p, a.dbgInst = id, a.dbgTerm = dT.p, a.exactGrid = vmk::exactGridFor(L.groups[g].qLen, true);
            VM_HIP(vmk::launchLoss<R>(L.groups[g].tpl, gridLoss[g], stream, a));
            VM_HIP(vmk::launchExact<R>(stream, a));
        }
        std::vector<vmk::GradT<R>> hg;
        if (wantGrad) {
            hg.resize(2 * size_t(in.S));
            if (in.S)
                VM_HIP(hipMemcpyAsync(hg.data(), dGrad.p + 2 * in.siteOff, hg.size() * sizeof(vmk::GradT<R>),
                                      hipMemcpyDeviceToHost, stream));
        }
        std::vector<int>          hk;
        std::vector<vmk::TermDbg> ht;
        if (wantDbg) {
            hk.resize(size_t(in.T) * K), ht.resize(in.T);
            if (in.T) {
                VM_HIP(hipMemcpyAsync(hk.data(), dK.p, hk.size() * sizeof(int),
                                      hipMemcpyDeviceToHost, stream));
                VM_HIP(hipMemcpyAsync(ht.data(), dT.p + in.targetOff, ht.size() * sizeof(vmk::TermDbg),
                                      hipMemcpyDeviceToHost, stream));
            }
        }
        vmk::InstState st;
        VM_HIP(hipMemcpyAsync(&st, dState.p + id, sizeof(st), hipMemcpyDeviceToHost, stream));
        if (wantGrad) VM_HIP(hipMemsetAsync(dGrad.p, 0, dGrad.n * sizeof(vmk::GradT<R>), stream));
        bin();
        VM_HIP(hipStreamSynchronize(stream));
        out.lossSum = st.evalLoss, out.count = st.evalCount;
        out.grad.assign(2 * size_t(in.S), 0.0);
        // FP32 runs: exact value of the fixed-point sums, sum * 2^-gradShift.
        const double inv = sizeof(R) == 4 ? std::ldexp(1.0, -in.gradShift) : 1.0;
        for (size_t k = 0; k < hg.size(); ++k) out.grad[k] = double(hg[k]) * inv;
        if (wantDbg) {
            out.knn.assign(size_t(in.T) * K, -1);
            out.terms.assign(in.T, vmk::TermDbg{-1, -1, 0.0});
            for (int t = 0; t < in.T; ++t) {
                int o = L.perm[id][t];
                std::copy_n(hk.begin() + size_t(t) * K, K, out.knn.begin() + size_t(o) * K);
                out.terms[o] = ht[t];
            }
        }
    }

    void diagAll(bool wantGrad, bool wantTerms, DiagOut &out) override
    {
        if (dirty) commit();
        DBuf<vmk::TermDbg> dT;
        if (wantTerms) dT.reserve(size_t(std::max<int64_t>(L.targets, 1)));
        gridBuild(vmk::kModeEvaluate);
        for (size_t g = 0; g < L.groups.size(); ++g) {
            auto a = lossArgs(int(g), vmk::kModeEvaluate, wantGrad);
            if (wantTerms) a.dbgTerm = dT.p;
            VM_HIP(vmk::launchLoss<R>(L.groups[g].tpl, gridLoss[g], stream, a));
            VM_HIP(vmk::launchExact<R>(stream, a));
        }
        std::vector<vmk::GradT<R>> hg;
        std::vector<vmk::TermDbg>  ht;
        if (wantGrad && L.sites) {
            hg.resize(2 * size_t(L.sites));
            VM_HIP(hipMemcpyAsync(hg.data(), dGrad.p, hg.size() * sizeof(hg[0]), hipMemcpyDeviceToHost, stream));
        }
        if (wantTerms && L.targets) {
            ht.resize(size_t(L.targets));
            VM_HIP(hipMemcpyAsync(ht.data(), dT.p, ht.size() * sizeof(ht[0]), hipMemcpyDeviceToHost, stream));
        }
        std::vector<vmk::InstState> hs(specs.size());
        if (!hs.empty())
            VM_HIP(hipMemcpyAsync(hs.data(), dState.p, hs.size() * sizeof(hs[0]), hipMemcpyDeviceToHost, stream));
        if (wantGrad) VM_HIP(hipMemsetAsync(dGrad.p, 0, dGrad.n * sizeof(vmk::GradT<R>), stream));
        bin();
        VM_HIP(hipStreamSynchronize(stream));
        out = DiagOut{};
        size_t nS = 0, nT = 0;
        for (size_t i = 0; i < specs.size(); ++i) {
            nS += size_t(L.inst[i].S), nT += size_t(L.inst[i].T);
            out.lossSum += hs[i].evalLoss, out.count += hs[i].evalCount;
        }
        if (wantGrad) out.grad.reserve(2 * nS);
        if (wantTerms) out.terms.assign(nT, vmk::TermDbg{-1, -1, 0.0});
        size_t tBase = 0;
        for (size_t i = 0; i < specs.size(); ++i) {
            const vmk::Inst &in = L.inst[i];
            if (wantGrad) {
                const double inv = sizeof(R) == 4 ? std::ldexp(1.0, -in.gradShift) : 1.0;
                for (int k = 0; k < 2 * in.S; ++k) out.grad.push_back(double(hg[2 * size_t(in.siteOff) + k]) * inv);
            }
            if (wantTerms)
                for (int t = 0; t < in.T; ++t) out.terms[tBase + size_t(L.perm[i][t])] = ht[size_t(in.targetOff) + t];
            tBase += size_t(in.T);
        }
    }

    void lossStats(uint64_t out[8]) override
    {
        if (dirty) commit();
        DBuf<unsigned long long> dS;
        dS.reserve(64 * 8);
        VM_HIP(hipMemsetAsync(dS.p, 0, 64 * 8 * sizeof(unsigned long long), stream));
        gridBuild(vmk::kModeEvaluate);
        for (size_t g = 0; g < L.groups.size(); ++g) {
            auto a  = lossArgs(int(g), vmk::kModeEvaluate, false);
            a.stats = dS.p;
            VM_HIP(vmk::launchLoss<R>(L.groups[g].tpl, gridLoss[g], stream, a));
            VM_HIP(vmk::launchExact<R>(stream, a));
        }
        bin();
        std::vector<unsigned long long> h(64 * 8);
        VM_HIP(hipMemcpyAsync(h.data(), dS.p, h.size() * sizeof(h[0]), hipMemcpyDeviceToHost, stream));
        VM_HIP(hipStreamSynchronize(stream));
        for (int k = 0; k < 8; ++k) {
            out[k] = 0;
            for (int l = 0; l < 64; ++l) out[k] += h[8 * l + k];
        }
    }

    void knnQuery(int id, const double *q, int nq, int K, int *idx, double *d2) override
    {
        if (dirty) commit();
        if (K < 1 || K > 32) throw VmError(1, "k must be in [1, 32]");
        gridBuild(vmk::kModeEvaluate);
        DBuf<V>   dq;
        DBuf<int> di;
        DBuf<R>   dd;
        dq.reserve(nq), di.reserve(size_t(nq) * K), dd.reserve(size_t(nq) * K);
        std::vector<V> hq(nq);
        for (int i = 0; i < nq; ++i) hq[i] = V{R(q[2 * i]), R(q[2 * i + 1])};
        if (nq) VM_HIP(hipMemcpyAsync(dq.p, hq.data(), nq * sizeof(V), hipMemcpyHostToDevice, stream));
        VM_HIP(vmk::launchKnnQuery<R>(K, stream, dInst.p, id, dSpos.p, dSidx.p, dStart.p, dq.p, nq, di.p, dd.p));
        std::vector<R> hd(size_t(nq) * K);
        if (nq) {
            VM_HIP(hipMemcpyAsync(idx, di.p, size_t(nq) * K * sizeof(int), hipMemcpyDeviceToHost, stream));
            VM_HIP(hipMemcpyAsync(hd.data(), dd.p, hd.size() * sizeof(R), hipMemcpyDeviceToHost, stream));
        }
        bin();
        VM_HIP(hipStreamSynchronize(stream));
        for (size_t i = 0; i < hd.size(); ++i) d2[i] = double(hd[i]);
    }

    // Pinned staging of commit's uploads (layouts up to kPinnedUploadMax bytes).
    static constexpr size_t kPinnedUploadMax = size_t(64) << 20;
    HBuf<unsigned char>     hUpload;
    // Local edits: pinned staging, pooled target indices to retire, the tail
    // (first position, targets used) and the chunk extent cap of the last layout.
    HBuf<unsigned char>     hPatch;
    DBuf<int64_t>           dPatchIdx;
    int64_t                 tailBase_ = 0, tailUsed_ = 0;
    float                   extentCap_ = 1e30f;

    DBuf<float>   clipVerts, clipCVerts;
    DBuf<int>     clipLabels, clipCount, clipFlags, clipStart, clipTiles, clipCLabels;
    DBuf<uint8_t> clipInside;
    // Pinned copies of the clip outputs (ClipOut views them: CSR) and of the
    // boundary's: V, watertight, loops, overflow, E, then vertices and edges.
    HBuf<float>   hClipVerts, hBndVerts;
    HBuf<int>     hClipLabels, hClipStart, hClipFlags, hBndEdges, hBndRes;
    HBuf<uint8_t> hClipInside;
    // Boundary assembly buffers, kept between calls like the clip's; their end
    // point capacity grows when a boundary does not fit.
    DBuf<int>         bndEcount, bndEoff, bndSlot, bndRep, bndVid, bndEdges, bndOut, bndIn, bndNxt, bndLab0, bndLab1,
                      bndNxt1, bndResult;
    DBuf<vmk::BndKey> bndKey;
    DBuf<float>       bndPos, bndVerts;
    DBuf<uint32_t>    bndTable;
    int               bndCap = 8192;

    // Boundary of the inside cells of the last clip, assembled on the device
    // (site space: labels are instance site indices, which keeps slot order).
    // The outputs are copied at capacity, so the clip's synchronization serves
    // both in the usual case.
    void boundaryEnqueue(int S)
    {
        const size_t cap = size_t(bndCap);
        size_t       tsz = 64;
        while (tsz < 2 * cap) tsz *= 2;
        bndEcount.reserve(size_t(S)), bndEoff.reserve(size_t(S));
        bndKey.reserve(cap), bndPos.reserve(2 * cap), bndTable.reserve(tsz), bndSlot.reserve(cap);
        bndRep.reserve(cap), bndVid.reserve(cap), bndVerts.reserve(2 * cap), bndEdges.reserve(cap);
        bndOut.reserve(cap), bndIn.reserve(cap), bndNxt.reserve(cap), bndLab0.reserve(cap), bndLab1.reserve(cap);
        bndNxt1.reserve(cap), bndResult.reserve(5);
        vmk::BoundaryArgs b{};
        b.S = S, b.count = clipCount.p, b.labels = clipLabels.p, b.verts = clipVerts.p, b.inside = clipInside.p;
        b.cap = int(cap), b.ecount = bndEcount.p, b.eoff = bndEoff.p, b.key = bndKey.p, b.pos = bndPos.p;
        b.table = bndTable.p, b.tableMask = uint32_t(tsz - 1), b.slot = bndSlot.p, b.rep = bndRep.p;
        b.vid = bndVid.p, b.vertices = bndVerts.p, b.edges = bndEdges.p, b.outDeg = bndOut.p, b.inDeg = bndIn.p;
        b.nxt = bndNxt.p, b.lab0 = bndLab0.p, b.lab1 = bndLab1.p, b.nxt1 = bndNxt1.p, b.result = bndResult.p;
        VM_HIP(vmk::launchBoundary(stream, b));
        hBndRes.reserve(5), hBndVerts.reserve(2 * cap), hBndEdges.reserve(cap);
        VM_HIP(hipMemcpyAsync(hBndRes.p, bndResult.p, 5 * sizeof(int), hipMemcpyDeviceToHost, stream));
        VM_HIP(hipMemcpyAsync(hBndVerts.p, bndVerts.p, 2 * cap * sizeof(float), hipMemcpyDeviceToHost, stream));
        VM_HIP(hipMemcpyAsync(hBndEdges.p, bndEdges.p, cap * sizeof(int), hipMemcpyDeviceToHost, stream));
    }

    // After the copies completed: false (and a larger capacity) when the
    // boundary had more end points than the capacity.
    bool boundaryFinish(Boundary &out)
    {
        const int *res = hBndRes.p;
        if (res[3]) {
            while (bndCap < 2 * res[4]) bndCap *= 2;
            return false;
        }
        out.vertices.assign(hBndVerts.p, hBndVerts.p + 2 * size_t(res[0]));
        out.edges.assign(hBndEdges.p, hBndEdges.p + 2 * size_t(res[4]));
        out.watertight = res[1] != 0, out.loops = res[2];
        return true;
    }

    void clip(int id, const uint64_t *dMasks, int tilesX, int tilesY, float voxel, ClipOut &out) override
    {
        if (dirty) commit();
        ensureEvents();
        const vmk::Inst &in = L.inst[id];
        const int        S  = in.S;
        // Output buffers are kept between calls (an extract per local edit).
        DBuf<float>   &dv = clipVerts;
        DBuf<int>     &dl = clipLabels, &dc = clipCount, &df = clipFlags;
        DBuf<uint8_t> &di = clipInside;
        dv.reserve(size_t(S) * vmk::kMaxCellVerts * 2), dl.reserve(size_t(S) * vmk::kMaxCellVerts);
        dc.reserve(S), di.reserve(S), df.reserve(2);
        VM_HIP(hipEventRecord(ev[5], stream));
        gridBuild(vmk::kModeEvaluate);
        VM_HIP(vmk::launchFillInt(stream, df.p, 2, 0));
        vmk::ClipArgs<R> a{};
        a.insts = dInst.p, a.inst = id, a.pos = dPos.p, a.spos = dSpos.p, a.sidx = dSidx.p, a.start = dStart.p;
        a.masks = dMasks, a.tilesX = tilesX, a.tilesY = tilesY, a.invVoxel = voxel > 0 ? 1.0f / voxel : 0.0f;
        a.verts = dv.p, a.labels = dl.p, a.count = dc.p, a.inside = di.p, a.flags = df.p;
        VM_HIP(vmk::launchClip<R>(S, stream, a));
        VM_HIP(hipEventRecord(ev[6], stream));
        // Without occupancy no cell is inside and the boundary is empty.
        const bool haveBoundary = dMasks != nullptr && S > 0;
        if (haveBoundary) boundaryEnqueue(S);
        bin();
        // The cells go down as CSR (only the vertices that exist).
        const size_t nv = size_t(S) * vmk::kMaxCellVerts;
        clipStart.reserve(size_t(S) + 1), clipTiles.reserve(size_t(std::max(vmk::csrTiles(S), 1)));
        clipCVerts.reserve(2 * nv), clipCLabels.reserve(nv);
        VM_HIP(vmk::launchClipCsr(stream, dv.p, dl.p, dc.p, S, clipStart.p, clipTiles.p, clipCVerts.p, clipCLabels.p));
        hClipStart.reserve(size_t(S) + 1), hClipInside.reserve(S), hClipFlags.reserve(2);
        if (S) {
            VM_HIP(hipMemcpyAsync(hClipStart.p, clipStart.p, (size_t(S) + 1) * sizeof(int), hipMemcpyDeviceToHost,
                                  stream));
            VM_HIP(hipMemcpyAsync(hClipInside.p, di.p, size_t(S), hipMemcpyDeviceToHost, stream));
        }
        VM_HIP(hipMemcpyAsync(hClipFlags.p, df.p, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
        VM_HIP(hipStreamSynchronize(stream));
        const size_t nnz = S ? size_t(hClipStart.p[S]) : 0;
        hClipVerts.reserve(std::max<size_t>(2 * nnz, 1)), hClipLabels.reserve(std::max<size_t>(nnz, 1));
        if (nnz) {
            VM_HIP(hipMemcpyAsync(hClipVerts.p, clipCVerts.p, 2 * nnz * sizeof(float), hipMemcpyDeviceToHost, stream));
            VM_HIP(hipMemcpyAsync(hClipLabels.p, clipCLabels.p, nnz * sizeof(int), hipMemcpyDeviceToHost, stream));
            VM_HIP(hipStreamSynchronize(stream));
        }
        VM_HIP(hipEventElapsedTime(&out.gpuMs, ev[5], ev[6]));
        out.S = S, out.verts = hClipVerts.p, out.labels = hClipLabels.p, out.start = hClipStart.p;
        out.inside = hClipInside.p;
        out.unsafe = hClipFlags.p[0], out.overflow = hClipFlags.p[1];
        out.bound  = Boundary{};
        out.bound.watertight = true;
        if (haveBoundary && !boundaryFinish(out.bound)) {
            boundaryEnqueue(S);
            VM_HIP(hipStreamSynchronize(stream));
            if (!boundaryFinish(out.bound)) throw VmError(2, "boundary assembly: capacity did not grow");
        }
    }
};

} // namespace

std::unique_ptr<Engine> Engine::create(int precision, hipStream_t stream)
{
    if (precision == 1) return std::unique_ptr<Engine>(new EngineT<double>(stream));
    if (precision == 0) return std::unique_ptr<Engine>(new EngineT<float>(stream));
    throw VmError(1, "precision must be 0 (fp32) or 1 (fp64)");
}

} // namespace vmrt
