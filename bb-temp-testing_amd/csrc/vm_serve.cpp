// vm_serve.cpp -- continuous-batching request loop over resident fit state.
//
// Requests arrive on a schedule: a caller's list (vm_serve) or a synthetic
// stream that the library generates lazily, request by request, from a seed
// (vm_serve_synthetic).  Worker threads prepare requests that have arrived,
// at most a fixed number ahead of the admission point, so a backlog costs
// only its arrival times: live sites and targets, the target sort and
// wavefront chunking, and upload images in the device element type.
// Synthetic requests first get their sites from the batched Poisson-disk dart
// thrower: a sampler thread takes the arrived requests in groups and samples
// each group in one call of the no-allocation entry point (a hipFree inside
// the loop would wait for the whole device).
//
// The GPU loop replays one captured iteration over every resident instance,
// `poll_every` iterations per launch, followed by an asynchronous readback of
// the per-instance states.  While a launch runs, the host retires the
// instances that the previous readback showed done (relative loss change
// below tol, or the iteration cap), admits prepared requests in arrival order
// while an instance slot, room within the HBM budget and room under the
// resident-site cap are free, and enqueues their uploads and the new work
// tables (through pinned staging) behind the running launch; the next launch
// picks them up.  The host waits only for the readback of the launch it just
// issued.  Device pools (sites, targets, hash cells, instance slots) start
// small and grow when an admission needs room: the stream is drained, the
// contents are copied into the larger buffers and the iteration graph is
// captured again.  Pools never grow past what the budget can fill or past
// the device's free memory.
#include "voromesh_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "vm_engine.h"

using vmrt::DBuf;
using vmrt::HBuf;
using vmrt::VmError;

namespace vmrt {
InstSpec specFromDesc(const vm_fit_desc &d);
}

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over [0, cap) in units of `granule` elements; grow()
// appends [cap, newCap) to the free list.
class RangeAlloc {
public:
    explicit RangeAlloc(int64_t granule = 64) : granule_(granule) {}
    void init(int64_t cap)
    {
        free_.clear();
        cap_ = 0, used_ = 0;
        grow(cap);
    }
    void grow(int64_t cap)
    {
        if (cap <= cap_) return;
        insertFree(cap_, cap - cap_);
        cap_ = cap;
    }
    int64_t alloc(int64_t n)
    {
        n = up(n);
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
        n = up(n);
        used_ -= n;
        insertFree(off, n);
    }
    int64_t cap() const { return cap_; }
    // End of the last allocated range (0 when nothing is allocated).
    int64_t highWater() const
    {
        if (free_.empty()) return cap_;
        const auto last = std::prev(free_.end());
        return last->first + last->second == cap_ ? last->first : cap_;
    }
    static int64_t round(int64_t n) { return ((n > 0 ? n : 1) + 63) & ~int64_t(63); }

private:
    int64_t up(int64_t n) const { return ((n > 0 ? n : 1) + granule_ - 1) / granule_ * granule_; }
    void    insertFree(int64_t off, int64_t n)
    {
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
    std::map<int64_t, int64_t> free_;
    int64_t                    granule_, cap_ = 0, used_ = 0;
};

// Grow-only pinned host buffer (contents lost on growth).
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

// Reallocates `b` to hold n elements, keeping its first `keep` elements and
// zero-filling the rest when asked; the stream is drained before the old
// buffer is freed.
template <class T> void regrow(DBuf<T> &b, size_t n, size_t keep, bool zero, hipStream_t s)
{
    if (b.p && n <= b.n) return;
    DBuf<T> nb;
    nb.reserve(n);
    if (zero) VM_HIP(hipMemsetAsync(nb.p, 0, nb.n * sizeof(T), s));
    keep = b.p ? std::min(keep, b.n) : 0;
    if (keep) VM_HIP(hipMemcpyAsync(nb.p, b.p, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
    VM_HIP(hipStreamSynchronize(s));
    std::swap(b.p, nb.p), std::swap(b.n, nb.n);
}

// ---------------------------------------------------------------------------
// Synthetic request stream
// ---------------------------------------------------------------------------
// SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): a 64-bit counter advanced by the golden-ratio
// increment, each value passed through a 64-bit finalizing mixer.
uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

class Rng {
public:
    // Stream `tag` of request `index` of the stream seeded `seed`.
    Rng(uint64_t seed, uint64_t index, uint64_t tag)
        : s_(mix64(mix64(seed + 0x9E3779B97F4A7C15ull) ^ (index * 8 + tag)))
    {
    }
    uint64_t next() { return mix64(s_ += 0x9E3779B97F4A7C15ull); }
    double   uniform() { return double(next() >> 11) * 0x1.0p-53; } // [0, 1)
    float    uniformf() { return float(next() >> 40) * 0x1.0p-24f; }  // [0, 1)
    // A pair of independent standard normals (single precision) by Marsaglia's
    // polar method (Marsaglia & Bray, SIAM Review 6, 1964): a point uniform in
    // the unit disc, scaled by sqrt(-2 ln q / q), q its squared radius.
    void normal2(float &a, float &b)
    {
        float u, v, q;
        do {
            u = 2.0f * uniformf() - 1.0f, v = 2.0f * uniformf() - 1.0f;
            q = u * u + v * v;
        } while (q >= 1.0f || q == 0.0f);
        const float f = std::sqrt(-2.0f * std::log(q) / q);
        a = u * f, b = v * f;
    }

private:
    uint64_t s_;
};

enum : uint64_t { kTagGap = 0, kTagSize = 1, kTagSeed = 2, kTagShape = 3 };

// Live sites per r^2 of area that the dart thrower leaves (30 rounds): the
// radius that gives a request about its drawn site count in the unit square.
constexpr double kPoissonYield = 0.668;

struct SynthPlan {
    double   gap;    // seconds since the previous arrival (the first: since the start)
    int      sites;  // drawn site count
    float    radius; // Poisson-disk radius for it
    unsigned seed;   // sampler seed
};

SynthPlan planOf(const vm_serve_stream &st, int64_t i)
{
    SynthPlan p;
    p.gap           = -std::log1p(-Rng(st.seed, uint64_t(i), kTagGap).uniform()) / st.rate;
    const double lo = std::log(double(st.min_sites)), hi = std::log(double(st.max_sites));
    p.sites         = int(std::lround(std::exp(lo + (hi - lo) * Rng(st.seed, uint64_t(i), kTagSize).uniform())));
    p.radius        = float(std::sqrt(kPoissonYield / double(p.sites)));
    p.seed          = unsigned(Rng(st.seed, uint64_t(i), kTagSeed).next() >> 32);
    return p;
}

// Request i's targets: targets_per_site * numSites points uniform by arc
// length on a random simple polygon (5-40 vertices at sorted uniform angles
// around (0.5, 0.5), radii U[0.15, 0.45]) with N(0, sigma^2) jitter per axis.
void synthTargets(const vm_serve_stream &st, int64_t i, int numSites, vm_vec2 *out)
{
    Rng                 g(st.seed, uint64_t(i), kTagShape);
    const int           m = 5 + int(g.next() % 36);
    std::vector<double> ang(m), vx(m), vy(m), sx(m), sy(m), len(m), cum(m + 1, 0.0);
    for (auto &a : ang) a = 6.283185307179586 * g.uniform();
    std::sort(ang.begin(), ang.end());
    for (int k = 0; k < m; ++k) {
        const double r = 0.15 + 0.30 * g.uniform();
        vx[k] = 0.5 + r * std::cos(ang[k]), vy[k] = 0.5 + r * std::sin(ang[k]);
    }
    for (int k = 0; k < m; ++k) {
        sx[k] = vx[(k + 1) % m] - vx[k], sy[k] = vy[(k + 1) % m] - vy[k];
        len[k]     = std::hypot(sx[k], sy[k]);
        cum[k + 1] = cum[k] + len[k];
    }
    // Segment lookup: the arc length is split into 256 equal bins; bin b
    // starts its search at the last edge that begins at or before the bin.
    constexpr int kBins = 256;
    int           first[kBins];
    for (int b = 0, k = 0; b < kBins; ++b) {
        const double at = cum[m] * b / kBins;
        while (k + 1 < m && cum[k + 1] <= at) ++k;
        first[b] = k;
    }
    const int64_t T = int64_t(st.targets_per_site) * numSites;
    for (int64_t t = 0; t < T; ++t) {
        const double u = g.uniform() * cum[m];
        int          k = first[std::min(int(u * (kBins / cum[m])), kBins - 1)];
        while (k > 0 && cum[k] > u) --k;           // (the bin index may round up past u)
        while (k + 1 < m && cum[k + 1] <= u) ++k; // the edge with cum[k] <= u < cum[k + 1]
        const double f = len[k] > 0 ? (u - cum[k]) / len[k] : 0.0;
        float        zx, zy;
        g.normal2(zx, zy);
        out[t] = vm_vec2{float(vx[k] + f * sx[k] + double(st.sigma * zx)), float(vy[k] + f * sy[k] + double(st.sigma * zy))};
    }
}

void checkStream(const vm_serve_stream &st)
{
    if (!(st.rate > 0) || !(st.duration_s >= 0) || st.min_sites < 1 || st.max_sites < st.min_sites ||
        st.targets_per_site < 1 || !(st.sigma >= 0) || !(st.lr > 0) || st.loss_k < 0)
        throw VmError(VM_EINVAL, "invalid serving stream");
}

// ---------------------------------------------------------------------------
// Requests between arrival and admission
// ---------------------------------------------------------------------------
struct Prepared {
    vmk::Inst                    in{};
    std::vector<vmk::SiteWork>   sw;
    std::vector<vmk::TargetWork> tw;
    int64_t                      cells = 0;
    // Upload images in the device element type: positions, targets in sorted
    // order and their slot numbers.
    std::vector<unsigned char> pos, tgt;
    std::vector<int>           tslot;
};

// Lays out one request for a server of element type R: every resident
// request shares one k and one chunk shape (one loss launch serves them all).
template <class R>
std::unique_ptr<Prepared> prepare(vmrt::InstSpec spec, const vm_serve_config &cfg, int K)
{
    using V      = vmk::V2<R>;
    spec.maxIter = cfg.max_iterations > 0 ? cfg.max_iterations : spec.maxIter;
    spec.tol     = cfg.tol > 0 ? cfg.tol : spec.tol;
    spec.K       = K;
    spec.tpl     = vmk::kTargetsPerLane;
    auto             p = std::make_unique<Prepared>();
    std::vector<int> perm;
    vmrt::layoutInstance(spec, -1, 0, 0, 0, p->in, p->sw, p->tw, perm);
    p->cells    = int64_t(p->in.gw) * p->in.gh + 1;
    const int S = p->in.S, T = p->in.T;
    p->pos.resize(size_t(S) * sizeof(V)), p->tgt.resize(size_t(T) * sizeof(V)), p->tslot.resize(size_t(T));
    V *hp = reinterpret_cast<V *>(p->pos.data()), *ht = reinterpret_cast<V *>(p->tgt.data());
    for (int k = 0; k < S; ++k) hp[k] = V{R(spec.pos[2 * k]), R(spec.pos[2 * k + 1])};
    for (int t = 0; t < T; ++t) {
        const int o         = perm[size_t(t)];
        ht[t]               = V{R(spec.tgt[2 * o]), R(spec.tgt[2 * o + 1])};
        p->tslot[size_t(t)] = spec.tslot.empty() ? o : spec.tslot[size_t(o)];
    }
    return p;
}

// Shared between the server thread, the preparation workers and a source's
// helper threads.
struct Pipeline {
    std::atomic<int64_t> arrived{0}, admitted{0};
    std::atomic<bool>    stop{false};
    int64_t              ahead = 160; // requests prepared beyond the admission point
    int64_t limit() const
    {
        return std::min(arrived.load(std::memory_order_acquire), admitted.load(std::memory_order_acquire) + ahead);
    }
};

// A schedule of requests.  arrivedBy() and closed() run on the server thread,
// spec() on the preparation workers (for arrived requests only).
class Source {
public:
    virtual ~Source() = default;
    virtual int64_t        capacity() const                                  = 0;
    virtual int64_t        arrivedBy(double now)                             = 0;
    virtual bool           closed() const                                    = 0; // no arrivals beyond those counted
    virtual double         arrival(int64_t i) const                          = 0;
    virtual vmrt::InstSpec spec(int64_t i, const std::atomic<bool> &stop)    = 0;
    virtual void           start(Pipeline &) {}
    virtual void           finish() {}
    virtual double         sampleSeconds() const { return 0.0; }
    virtual double         waitSeconds() const { return 0.0; }
};

class ListSource final : public Source {
public:
    ListSource(const vm_fit_desc *reqs, const double *arrival, int n) : reqs_(reqs), arr_(arrival), n_(n) {}
    int64_t capacity() const override { return n_; }
    int64_t arrivedBy(double now) override
    {
        while (count_ < n_ && arr_[count_] <= now) ++count_;
        return count_;
    }
    bool           closed() const override { return count_ == n_; }
    double         arrival(int64_t i) const override { return arr_[i]; }
    vmrt::InstSpec spec(int64_t i, const std::atomic<bool> &) override { return vmrt::specFromDesc(reqs_[i]); }

private:
    const vm_fit_desc *reqs_;
    const double      *arr_;
    int64_t            n_, count_ = 0;
};

class SynthSource final : public Source {
public:
    SynthSource(const vm_serve_stream &st, int64_t cap, vm_serve_record *rec)
        : st_(st), cap_(cap), rec_(rec), arr_(size_t(cap)), sites_(size_t(cap)), sampled_(new std::atomic<int>[cap])
    {
        for (int64_t i = 0; i < cap; ++i) sampled_[i].store(0);
        nextT_ = planOf(st_, 0).gap;
        VM_HIP(hipGetDevice(&device_));
    }
    ~SynthSource() override { finish(); }

    int64_t capacity() const override { return cap_; }
    int64_t arrivedBy(double now) override
    {
        while (nextT_ < st_.duration_s && nextT_ <= now) {
            if (count_ == cap_)
                throw VmError(VM_EINVAL, "the record buffer (" + std::to_string(cap_) +
                                             " requests) is too small for the stream's arrivals");
            arr_[size_t(count_)] = nextT_;
            rec_[count_]         = vm_serve_record{nextT_, -1.0, -1.0, 0, 0, 0, 0.0};
            ++count_;
            nextT_ += planOf(st_, count_).gap;
        }
        return count_;
    }
    bool   closed() const override { return nextT_ >= st_.duration_s; }
    double arrival(int64_t i) const override { return arr_[size_t(i)]; }

    vmrt::InstSpec spec(int64_t i, const std::atomic<bool> &stop) override
    {
        int        s;
        const auto tw = Clock::now();
        while ((s = sampled_[i].load(std::memory_order_acquire)) == 0) {
            if (stop.load()) throw VmError(VM_ESTATE, "serving loop stopped");
            std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
        waitNs_.fetch_add(std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - tw).count());
        if (s == 2) {
            std::lock_guard<std::mutex> lk(errMu_);
            throw VmError(VM_EINVAL, "site sampling failed: " + samplerError_);
        }
        std::vector<vm_vec2> sites = std::move(sites_[size_t(i)]);
        std::vector<vm_vec2> tg(size_t(st_.targets_per_site) * sites.size());
        synthTargets(st_, i, int(sites.size()), tg.data());
        vm_fit_desc d{};
        d.width = 1.0f, d.height = 1.0f;
        d.sites = sites.data(), d.num_sites = int(sites.size());
        d.targets = tg.data(), d.num_targets = int(tg.size());
        d.lr = st_.lr, d.gamma = 0.5f;
        const int ms[4] = {80, 150, 200, 250};
        std::memcpy(d.milestones, ms, sizeof(ms));
        d.sample_fraction = 1.0f, d.seed = 1, d.loss_k = st_.loss_k > 0 ? st_.loss_k : 16;
        return vmrt::specFromDesc(d);
    }

    void start(Pipeline &pl) override
    {
        sampler_ = std::thread([this, &pl] { samplerLoop(pl); });
    }
    void finish() override
    {
        if (sampler_.joinable()) sampler_.join();
    }
    double sampleSeconds() const override { return sampleS_; }
    double waitSeconds() const override { return double(waitNs_.load()) * 1e-9; }

private:
    // Grid cells per sampler call: the no-allocation sampler's fixed device
    // array holds about 2 M cells with its tables.  A single request beyond
    // this is sampled alone through the allocating entry point.
    static constexpr int64_t kGroupCells = 1800000;
    static constexpr int     kGroupMax   = 64;
    static constexpr int     kTrials     = 30;

    void samplerLoop(Pipeline &pl)
    {
        (void)hipSetDevice(device_);
        std::vector<float>    w, h, r;
        std::vector<unsigned> sd;
        std::vector<int64_t>  off, idx;
        std::vector<vm_vec2>  cells;
        int64_t               next = 0;
        while (!pl.stop.load()) {
            const int64_t lim = std::min(pl.limit(), cap_);
            if (next >= lim) {
                std::this_thread::sleep_for(std::chrono::microseconds(200));
                continue;
            }
            w.clear(), h.clear(), r.clear(), sd.clear(), off.clear(), idx.clear();
            int64_t total = 0;
            for (int64_t j = next; j < lim && int(idx.size()) < kGroupMax; ++j) {
                const SynthPlan p = planOf(st_, j);
                int             gw = 0, gh = 0;
                float           cs = 0;
                if (vm_poisson_grid_dims(1.0f, 1.0f, p.radius, &gw, &gh, &cs) != 0) break;
                const int64_t c = int64_t(gw) * gh;
                if (!idx.empty() && total + c > kGroupCells) break;
                w.push_back(1.0f), h.push_back(1.0f), r.push_back(p.radius), sd.push_back(p.seed);
                off.push_back(total), idx.push_back(j);
                total += c;
            }
            const int n = int(idx.size());
            if (n == 0) { // grid dims refused the radius
                fail(next, next + 1, vm_last_error());
                ++next;
                continue;
            }
            cells.resize(size_t(total));
            const auto ts = Clock::now();
            const int  rc = total <= kGroupCells
                               ? vm_poisson_batch_noalloc(n, w.data(), h.data(), r.data(), sd.data(), kTrials,
                                                          off.data(), cells.data(), nullptr)
                               : vm_poisson_batch(n, w.data(), h.data(), r.data(), sd.data(), kTrials, off.data(),
                                                  cells.data(), nullptr);
            sampleS_ += std::chrono::duration<double>(Clock::now() - ts).count();
            if (rc != 0) {
                fail(idx.front(), idx.back() + 1, vm_last_error());
            } else {
                for (int k = 0; k < n; ++k) {
                    const int64_t        end = k + 1 < n ? off[size_t(k) + 1] : total;
                    std::vector<vm_vec2> live;
                    for (int64_t c = off[size_t(k)]; c < end; ++c)
                        if (cells[size_t(c)].x >= 0.0f) live.push_back(cells[size_t(c)]);
                    sites_[size_t(idx[size_t(k)])] = std::move(live);
                    sampled_[idx[size_t(k)]].store(1, std::memory_order_release);
                }
            }
            next = idx.back() + 1;
        }
    }
    void fail(int64_t a, int64_t b, const char *msg)
    {
        {
            std::lock_guard<std::mutex> lk(errMu_);
            samplerError_ = msg ? msg : "unknown error";
        }
        for (int64_t j = a; j < b; ++j) sampled_[j].store(2, std::memory_order_release);
    }

    vm_serve_stream                   st_;
    int64_t                           cap_;
    vm_serve_record                  *rec_;
    std::vector<double>               arr_;
    std::vector<std::vector<vm_vec2>> sites_;
    std::unique_ptr<std::atomic<int>[]> sampled_;
    int64_t                           count_ = 0;
    double                            nextT_ = 0;
    int                               device_ = 0;
    double                            sampleS_ = 0; // written by the sampler thread, read after it joined
    std::atomic<int64_t>              waitNs_{0};
    std::thread                       sampler_;
    std::mutex                        errMu_;
    std::string                       samplerError_;
};

// ---------------------------------------------------------------------------
// Resident state on the device
// ---------------------------------------------------------------------------
enum class Admit { Ok, Wait, Never };

template <class R> class Server {
    using V = vmk::V2<R>;

public:
    Server(const vm_serve_config &cfg, int K) : cfg_(cfg), K_(K)
    {
        VM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        VM_HIP(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking));
        VM_HIP(hipEventCreateWithFlags(&evCopy_, hipEventDisableTiming));
        VM_HIP(hipEventCreate(&ev0_));
        VM_HIP(hipEventCreate(&ev1_));
        VM_HIP(hipEventCreate(&evRead_));
        budget_ = cfg.hbm_budget_bytes;
        // No pool needs more than the whole budget spent on it alone.
        maxS_ = RangeAlloc::round(int64_t(budget_ / kPerSite));
        maxT_ = RangeAlloc::round(int64_t(budget_ / kPerTarget));
        maxC_ = RangeAlloc::round(int64_t(budget_ / kPerCell));
        // Initial pools: `initial_pool_bytes` (default 256 MiB) split in the
        // serving law's proportions (8 targets and about half a hash cell per site).
        const double ib   = cfg.initial_pool_bytes > 0 ? cfg.initial_pool_bytes : double(256 << 20);
        const double unit = kPerSite + 8 * kPerTarget + 0.5 * kPerCell;
        const int64_t s0  = std::max<int64_t>(int64_t(ib / unit), 64);
        // Slot 0 is a permanent placeholder (an empty instance marked done):
        // work-table entries of retired instances are pointed at it, and every
        // kernel skips them.  cfg.max_resident counts the request slots.
        fixedSlots_ = cfg.max_resident > 0;
        growSlots(1 + (fixedSlots_ ? cfg.max_resident : 64));
        sites_.init(0), targets_.init(0), cells_.init(0);
        growSites(std::min(maxS_, RangeAlloc::round(s0)));
        growTargets(std::min(maxT_, RangeAlloc::round(8 * s0)));
        growCells(std::min(maxC_, RangeAlloc::round(s0 / 2)));
        tileAlloc_.init(0), swAlloc_.init(0), twAlloc_.init(0);
        growTable(0, 1024), growTable(1, 4096), growTable(2, 16384);
        poolGrows_ = 0;
        dFrozen.reserve(1), dScanCtl.reserve(1), dCnt.reserve(4), dNewCnt.reserve(1), dExactLen.reserve(1);
        VM_HIP(hipMemsetAsync(dScanCtl.p, 0, sizeof(vmk::ScanCtl), stream_));
        vmk::Inst      none{};
        vmk::InstState doneState{};
        none.W = none.H = 1.0f, none.cell = none.inv = 1.0f, none.ext = 1.0f, none.gw = none.gh = 1;
        none.K = K, none.allMovable = 1;
        doneState.done = 1;
        VM_HIP(hipMemcpy(dInst.p, &none, sizeof(none), hipMemcpyHostToDevice));
        VM_HIP(hipMemcpy(dState.p, &doneState, sizeof(doneState), hipMemcpyHostToDevice));
        VM_HIP(hipStreamSynchronize(stream_));
    }
    ~Server()
    {
        // Graph replays, readbacks and uploads may still be queued (an
        // exception can leave the loop after launch()): let them finish before
        // the buffers and the staging memory they use are released.
        (void)hipStreamSynchronize(copy_);
        (void)hipStreamSynchronize(stream_);
        if (graph_) (void)hipGraphExecDestroy(graph_);
        (void)hipEventDestroy(evCopy_);
        (void)hipStreamDestroy(copy_);
        (void)hipEventDestroy(ev0_);
        (void)hipEventDestroy(ev1_);
        (void)hipEventDestroy(evRead_);
        (void)hipStreamDestroy(stream_);
    }

    // Places a prepared request (its uploads are enqueued by the next refresh()).
    Admit admit(int64_t req, std::unique_ptr<Prepared> &pp)
    {
        Prepared    &p = *pp;
        const int    S = p.in.S, T = p.in.T;
        const double bytes = bytesFor(S, T, p.cells);
        if (bytes > budget_) return Admit::Never;
        if (residentBytes_ + bytes > budget_) return Admit::Wait;
        if (cfg_.max_resident_sites > 0 && resident_ > 0 && residentSites_ + S > cfg_.max_resident_sites)
            return Admit::Wait;
        int slot = freeSlot();
        if (slot < 0) {
            if (fixedSlots_) return Admit::Wait;
            growSlots(2 * slots_);
            slot = freeSlot();
        }
        int64_t so = -1, to = -1, co = -1;
        if ((so = allocGrow(sites_, S, maxS_, &Server::growSites)) < 0 ||
            (to = allocGrow(targets_, T, maxT_, &Server::growTargets)) < 0 ||
            (co = allocGrow(cells_, p.cells, maxC_, &Server::growCells)) < 0) {
            if (so >= 0) sites_.release(so, S);
            if (to >= 0) targets_.release(to, T);
            return resident_ == 0 ? Admit::Never : Admit::Wait;
        }
        const int nt = vmk::scanTilesFor(p.cells);
        const int64_t tOff = allocTable(0, nt), sOff = allocTable(1, int64_t(p.sw.size())),
                      wOff = allocTable(2, int64_t(p.tw.size()));
        p.in.siteOff = so, p.in.cellOff = co, p.in.targetOff = to;
        SlotInfo &si = slotInfo_[size_t(slot)];
        si           = SlotInfo{so, to, co, S, T, p.cells, bytes, launchSeq_, tOff, sOff, wOff, {}, {}, {}};
        for (int k = 0; k < nt; ++k) {
            const int b = k * vmk::kScanTileCells;
            si.tiles.push_back({slot, b, int(std::min<int64_t>(vmk::kScanTileCells, p.cells - b)), k == 0});
        }
        for (auto w : p.sw) si.sw.push_back({slot, w.begin, w.count, 0, so, co});
        for (auto w : p.tw) { // laid out at offset 0: move to the pool ranges
            w.inst = slot, w.tBegin += to, w.cellOff = co, w.siteOff = so;
            si.tw.push_back(w);
        }
        newWork_.insert(newWork_.end(), si.sw.begin(), si.sw.end());
        uploads_.push_back({slot, std::move(pp)});
        slotReq_[size_t(slot)] = req;
        residentBytes_ += bytes, residentSites_ += S;
        ++resident_;
        maxResident_ = std::max(maxResident_, resident_);
        peakBytes_   = std::max(peakBytes_, residentBytes_);
        changed_     = true;
        return Admit::Ok;
    }

    // Enqueues, behind whatever the stream already holds: placeholders over
    // the work-table ranges of retired instances, the admitted requests'
    // uploads and work-table entries, the table lengths and the binning of
    // the new instances.  Host work and traffic scale with the admissions
    // and retirements, not with what stays resident.
    void refresh()
    {
        if (!changed_) return;
        const int nTiles = int(tileAlloc_.highWater()), nSW = int(swAlloc_.highWater());
        const int nTW = int(twAlloc_.highWater()), nn = int(newWork_.size());
        const int tileCounts[2] = {nTiles, nTiles}; // every tile takes the look-back path
        const int cnt[3]        = {0, nSW, nTW};
        if (size_t(nn) > dNewWork.n) {
            quiesce();
            regrow(dNewWork, std::max(size_t(nn) * 3 / 2, size_t(4096)), 0, false, stream_);
        }
        // Placeholders over freed ranges go first: an admission may reuse them.
        for (const Freed &f : freed_) {
            if (f.kind == 0) fillDummy(dTiles.p + f.off, f.n);
            else if (f.kind == 1) fillDummy(dSiteWork.p + f.off, f.n);
            else fillDummy(dTargetWork.p + f.off, f.n);
        }
        freed_.clear();
        // Everything else goes up from one pinned staging buffer (DMA at full
        // PCIe rate, and nothing is read from pageable memory after this returns).
        std::vector<size_t> at;
        size_t              bytes = 0;
        auto                put   = [&](size_t b) {
            at.push_back(bytes);
            bytes += (b + 255) & ~size_t(255);
        };
        put(sizeof(tileCounts)), put(sizeof(cnt)), put(newWork_.size() * sizeof(vmk::SiteWork)), put(sizeof(nn));
        for (auto &u : uploads_) {
            const SlotInfo &si = slotInfo_[size_t(u.slot)];
            put(u.p->pos.size()), put(u.p->tgt.size()), put(u.p->tslot.size() * sizeof(int));
            put(sizeof(vmk::Inst)), put(sizeof(vmk::InstState));
            put(si.tiles.size() * sizeof(vmk::ScanTile)), put(si.sw.size() * sizeof(vmk::SiteWork));
            put(si.tw.size() * sizeof(vmk::TargetWork));
        }
        const int k = stageNext_;
        if (stageTag_[k] >= doneSeq_) drain(); // its last upload may still be in flight
        uploadBytes_ += double(bytes);
        stage_[size_t(k)].reserve(bytes);
        unsigned char *h      = stage_[size_t(k)].p;
        auto           copyUp = [&](void *dst, size_t slot, const void *src, size_t b, hipStream_t on) {
            if (!b) return;
            std::memcpy(h + at[slot], src, b);
            VM_HIP(hipMemcpyAsync(dst, h + at[slot], b, hipMemcpyHostToDevice, on));
        };
        // Admitted requests' pool ranges (positions, zero moments / gradient /
        // cell counts, sorted targets and slots) go up on the copy stream while
        // the running launch computes: a range belongs either to no instance
        // or to one retired as done, which every kernel skips (they may read
        // it before testing the state, and discard what they read).
        for (auto &u : uploads_) {
            const SlotInfo &si = slotInfo_[size_t(u.slot)];
            const Prepared &p  = *u.p;
            const size_t    a  = 4 + 8 * size_t(&u - uploads_.data());
            copyUp(dPos.p + si.so, a, p.pos.data(), p.pos.size(), copy_);
            if (si.S) {
                VM_HIP(hipMemsetAsync(dM.p + si.so, 0, size_t(si.S) * sizeof(V), copy_));
                VM_HIP(hipMemsetAsync(dV.p + si.so, 0, size_t(si.S) * sizeof(V), copy_));
                VM_HIP(hipMemsetAsync(dGrad.p + 2 * si.so, 0, 2 * size_t(si.S) * sizeof(vmk::GradT<R>), copy_));
            }
            VM_HIP(hipMemsetAsync(dCounts.p + si.co, 0, size_t(si.C) * sizeof(int), copy_));
            copyUp(dTpos.p + si.to, a + 1, p.tgt.data(), p.tgt.size(), copy_);
            copyUp(dTslot.p + si.to, a + 2, p.tslot.data(), p.tslot.size() * sizeof(int), copy_);
        }
        if (!uploads_.empty()) {
            VM_HIP(hipEventRecord(evCopy_, copy_));
            VM_HIP(hipStreamWaitEvent(stream_, evCopy_, 0));
        }
        // Records, zero states and work entries follow the running launch on
        // the compute stream (it still reads the slots' previous occupants).
        size_t a = 4;
        for (auto &u : uploads_) {
            const SlotInfo &si = slotInfo_[size_t(u.slot)];
            const Prepared &p  = *u.p;
            copyUp(dInst.p + u.slot, a + 3, &p.in, sizeof(p.in), stream_);
            copyUp(dState.p + u.slot, a + 4, &zeroState_, sizeof(zeroState_), stream_);
            copyUp(dTiles.p + si.tileOff, a + 5, si.tiles.data(), si.tiles.size() * sizeof(vmk::ScanTile), stream_);
            copyUp(dSiteWork.p + si.swOff, a + 6, si.sw.data(), si.sw.size() * sizeof(vmk::SiteWork), stream_);
            copyUp(dTargetWork.p + si.twOff, a + 7, si.tw.data(), si.tw.size() * sizeof(vmk::TargetWork), stream_);
            a += 8;
        }
        copyUp(&dScanCtl.p->nTiles, 0, tileCounts, sizeof(tileCounts), stream_);
        copyUp(dCnt.p, 1, cnt, sizeof(cnt), stream_);
        if (nn) {
            copyUp(dNewWork.p, 2, newWork_.data(), newWork_.size() * sizeof(vmk::SiteWork), stream_);
            copyUp(dNewCnt.p, 3, &nn, sizeof(nn), stream_);
            auto ab = adamArgs(0);
            ab.work = dNewWork.p, ab.nWork = dNewCnt.p;
            VM_HIP(vmk::launchAdamBin<R>(std::min(nn, 8192), stream_, ab));
        }
        stageTag_[k] = launchSeq_;
        stageNext_ ^= 1;
        uploads_.clear(); // the images now sit in pinned memory
        newWork_.clear();
        changed_ = false;
    }

    // Enqueues `iters` iterations of every resident instance and a readback
    // of the instance states; waitReadback() blocks until it has landed.
    void launch(int iters)
    {
        if (!graph_) buildGraph();
        VM_HIP(hipEventRecord(ev0_, stream_));
        for (int k = 0; k < iters; ++k) VM_HIP(hipGraphLaunch(graph_, stream_));
        VM_HIP(hipEventRecord(ev1_, stream_));
        readback_.reserve(size_t(slots_));
        VM_HIP(hipMemcpyAsync(readback_.p, dState.p, size_t(slots_) * sizeof(vmk::InstState), hipMemcpyDeviceToHost,
                              stream_));
        VM_HIP(hipEventRecord(evRead_, stream_));
        ++launchSeq_;
        readSlots_   = slots_;
        readPending_ = true;
    }
    void waitReadback()
    {
        if (!readPending_) return;
        VM_HIP(hipEventSynchronize(evRead_));
        float ms = 0;
        VM_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
        gpuMs_ += ms;
        observed_.assign(readback_.p, readback_.p + readSlots_);
        observedSeq_ = launchSeq_;
        // Uploads enqueued before this launch's readback have completed.
        doneSeq_      = std::max(doneSeq_, launchSeq_);
        readPending_  = false;
        haveObserved_ = true;
    }

    // Retires the instances the last readback showed done (each readback is
    // used once).  Calls done(req, state, slot) before the slot is freed.
    template <class F> int retire(F &&done)
    {
        if (!haveObserved_) return 0;
        haveObserved_ = false;
        int n = 0;
        for (int s = 1; s < int(observed_.size()) && s < slots_; ++s) {
            SlotInfo &si = slotInfo_[size_t(s)];
            // A slot refilled after that launch was issued shows its previous occupant.
            if (slotReq_[size_t(s)] < 0 || si.admitSeq >= observedSeq_ || !observed_[size_t(s)].done) continue;
            done(slotReq_[size_t(s)], observed_[size_t(s)], s);
            sites_.release(si.so, si.S), targets_.release(si.to, si.T), cells_.release(si.co, si.C);
            releaseTable(0, si.tileOff, int64_t(si.tiles.size()));
            releaseTable(1, si.swOff, int64_t(si.sw.size()));
            releaseTable(2, si.twOff, int64_t(si.tw.size()));
            si.tiles.clear(), si.sw.clear(), si.tw.clear();
            residentBytes_ -= si.bytes, residentSites_ -= si.S;
            slotReq_[size_t(s)] = -1;
            --resident_;
            changed_ = true;
            ++n;
        }
        return n;
    }

    // Copies a retired (done, so frozen) instance's sites out; waits for the stream.
    void downloadSites(int slot, vm_vec2 *out)
    {
        const SlotInfo &si = slotInfo_[size_t(slot)];
        std::vector<V>  h(size_t(si.S));
        if (si.S)
            VM_HIP(hipMemcpyAsync(h.data(), dPos.p + si.so, size_t(si.S) * sizeof(V), hipMemcpyDeviceToHost, stream_));
        VM_HIP(hipStreamSynchronize(stream_));
        for (int k = 0; k < si.S; ++k) out[k] = vm_vec2{float(h[size_t(k)].x), float(h[size_t(k)].y)};
    }

    int    resident() const { return resident_; }
    int    maxResident() const { return maxResident_; }
    double peakBytes() const { return peakBytes_; }
    int    poolGrows() const { return poolGrows_; }
    double gpuSeconds() const { return gpuMs_ * 1e-3; }
    double uploadBytes() const { return uploadBytes_; }
    // Site-iterations of the resident instances as of the last readback.
    double residentSiteIters() const
    {
        double n = 0;
        for (int s = 1; s < int(observed_.size()) && s < slots_; ++s)
            if (slotReq_[size_t(s)] >= 0 && slotInfo_[size_t(s)].admitSeq < observedSeq_)
                n += double(slotInfo_[size_t(s)].S) * observed_[size_t(s)].iter;
        return n;
    }
    double poolBytes() const
    {
        return double(sites_.cap()) * kPerSite + double(targets_.cap()) * kPerTarget + double(cells_.cap()) * kPerCell;
    }

    // Resident bytes charged for a request: per site four position-sized
    // arrays, two gradient words, three ints and a flag; per target its
    // position, slot and exact-queue entry; per hash cell two ints.
    static double bytesFor(int S, int T, int64_t C)
    {
        return double(RangeAlloc::round(S)) * kPerSite + double(RangeAlloc::round(T)) * kPerTarget +
               double(RangeAlloc::round(C)) * kPerCell;
    }

private:
    static constexpr double kPerSite   = 4.0 * sizeof(V) + 2.0 * sizeof(vmk::GradT<R>) + 12 + 1;
    static constexpr double kPerTarget = sizeof(V) + 4 + sizeof(vmk::ExactItem);
    static constexpr double kPerCell   = 2 * sizeof(int);

    struct SlotInfo {
        int64_t                      so = 0, to = 0, co = 0;
        int                          S = 0, T = 0;
        int64_t                      C = 0;
        double                       bytes = 0;
        int64_t                      admitSeq = 0;                  // launches issued before its uploads
        int64_t                      tileOff = 0, swOff = 0, twOff = 0; // its ranges of the work tables
        std::vector<vmk::ScanTile>   tiles;
        std::vector<vmk::SiteWork>   sw;
        std::vector<vmk::TargetWork> tw;
    };
    struct Upload {
        int                       slot;
        std::unique_ptr<Prepared> p;
    };
    struct Freed {
        int     kind; // 0 scan tiles, 1 site work, 2 target work
        int64_t off, n;
    };

    int freeSlot() const
    {
        for (int s = 1; s < slots_; ++s)
            if (slotReq_[size_t(s)] < 0) return s;
        return -1;
    }

    // Every queued command has finished.
    void drain()
    {
        VM_HIP(hipStreamSynchronize(copy_));
        VM_HIP(hipStreamSynchronize(stream_));
        doneSeq_ = launchSeq_ + 1;
    }
    // Drained, and no captured graph holds the current buffers.
    void quiesce()
    {
        drain();
        if (graph_) (void)hipGraphExecDestroy(graph_), graph_ = nullptr;
        changed_ = true;
    }

    // Room for `add` more bytes of device memory (with 1 GiB to spare)?
    static bool deviceRoom(double add)
    {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return false;
        return add + double(size_t(1) << 30) <= double(fr);
    }

    // Allocates n elements from a pool, growing it (to at least twice its
    // capacity, at most maxCap) when it has no room.
    int64_t allocGrow(RangeAlloc &ra, int64_t n, int64_t maxCap, void (Server::*grow)(int64_t))
    {
        int64_t off = ra.alloc(n);
        if (off >= 0) return off;
        const int64_t cap = std::min(maxCap, std::max(2 * ra.cap(), ra.cap() + 2 * RangeAlloc::round(n)));
        if (cap <= ra.cap()) return -1;
        (this->*grow)(cap);
        return ra.alloc(n);
    }
    void growSites(int64_t cap)
    {
        if (!deviceRoom(double(cap) * kPerSite)) return;
        quiesce();
        const size_t keep = size_t(sites_.cap());
        regrow(dPos, cap, keep, false, stream_), regrow(dM, cap, keep, false, stream_);
        regrow(dV, cap, keep, false, stream_), regrow(dSpos, cap, keep, false, stream_);
        regrow(dGrad, 2 * cap, 2 * keep, false, stream_);
        regrow(dCellOf, cap, keep, false, stream_), regrow(dRank, cap, keep, false, stream_);
        regrow(dSidx, cap, keep, false, stream_);
        sites_.grow(cap);
        ++poolGrows_;
    }
    void growTargets(int64_t cap)
    {
        if (!deviceRoom(double(cap) * kPerTarget)) return;
        quiesce();
        const size_t keep = size_t(targets_.cap());
        regrow(dTpos, cap, keep, false, stream_), regrow(dTslot, cap, keep, false, stream_);
        regrow(dExactQ, cap, 0, false, stream_);
        targets_.grow(cap);
        ++poolGrows_;
    }
    void growCells(int64_t cap)
    {
        if (!deviceRoom(double(cap) * kPerCell)) return;
        quiesce();
        const size_t keep = size_t(cells_.cap());
        regrow(dCounts, cap, keep, false, stream_), regrow(dStart, cap, keep, false, stream_);
        cells_.grow(cap);
        ++poolGrows_;
    }
    // Instance slots: records, states (zero for new slots) and loss partials.
    void growSlots(int n)
    {
        if (slots_) quiesce();
        regrow(dInst, size_t(n), size_t(slots_), false, stream_);
        regrow(dState, size_t(n), size_t(slots_), true, stream_);
        regrow(dPartial, size_t(n) * vmk::kLossPartials, size_t(slots_) * vmk::kLossPartials, true, stream_);
        slots_ = n;
        slotReq_.resize(size_t(n), -1);
        slotInfo_.resize(size_t(n));
    }

    // ---- work tables: one range per resident instance ----------------------
    RangeAlloc &tableAlloc(int kind) { return kind == 0 ? tileAlloc_ : kind == 1 ? swAlloc_ : twAlloc_; }
    int64_t allocTable(int kind, int64_t n)
    {
        if (n == 0) return 0;
        RangeAlloc &ra  = tableAlloc(kind);
        int64_t     off = ra.alloc(n);
        if (off < 0) {
            growTable(kind, std::max(2 * ra.cap(), ra.cap() + 2 * n));
            off = ra.alloc(n);
        }
        return off;
    }
    void releaseTable(int kind, int64_t off, int64_t n)
    {
        if (n == 0) return;
        tableAlloc(kind).release(off, n);
        freed_.push_back({kind, off, n});
    }
    // Enlarges a table (contents kept, the new part filled with placeholders).
    void growTable(int kind, int64_t cap)
    {
        RangeAlloc   &ra  = tableAlloc(kind);
        const int64_t old = ra.cap();
        quiesce();
        if (kind == 0) {
            regrow(dTiles, size_t(cap), size_t(old), false, stream_);
            regrow(dFlags, size_t(cap), size_t(old), true, stream_);
            fillDummy(dTiles.p + old, cap - old);
        } else if (kind == 1) {
            regrow(dSiteWork, size_t(cap), size_t(old), false, stream_);
            fillDummy(dSiteWork.p + old, cap - old);
        } else {
            regrow(dTargetWork, size_t(cap), size_t(old), false, stream_);
            fillDummy(dTargetWork.p + old, cap - old);
        }
        ra.grow(cap);
    }
    // Placeholder entries (instance 0, nothing to do) from pinned arrays of them.
    template <class E> void fillDummy(E *dst, int64_t n)
    {
        PinnedBuf &buf = dummy_[std::is_same<E, vmk::ScanTile>::value ? 0 : std::is_same<E, vmk::SiteWork>::value ? 1 : 2];
        const size_t bytes = size_t(n) * sizeof(E);
        if (bytes > buf.n || !buf.p) {
            drain(); // copies from the old array may be in flight
            buf.reserve(bytes);
            E dummy = placeholder<E>();
            for (size_t i = 0; i < buf.n / sizeof(E); ++i) std::memcpy(buf.p + i * sizeof(E), &dummy, sizeof(E));
        }
        VM_HIP(hipMemcpyAsync(dst, buf.p, bytes, hipMemcpyHostToDevice, stream_));
        uploadBytes_ += double(bytes);
    }
    template <class E> static E placeholder()
    {
        E e{};
        if constexpr (std::is_same<E, vmk::ScanTile>::value) e.first = 1; // begins (and skips) instance 0's step
        if constexpr (std::is_same<E, vmk::TargetWork>::value) e.gw = e.gh = 1;
        return e;
    }

    vmk::AdamArgs<R> adamArgs(int doAdam)
    {
        vmk::AdamArgs<R> a{};
        a.insts = dInst.p, a.state = dState.p, a.work = dSiteWork.p, a.nWork = dCnt.p + 1;
        a.pos = dPos.p, a.m = dM.p, a.v = dV.p, a.grad = dGrad.p, a.frozen = dFrozen.p, a.partial = dPartial.p;
        a.cellOf = dCellOf.p, a.rank = dRank.p, a.counts = dCounts.p, a.doAdam = doAdam;
        return a;
    }

    void buildGraph()
    {
        const int gScan = int(std::min<size_t>(dTiles.n, 1024)), gLoss = 65536;
        const int gSite = int(std::min<size_t>(dSiteWork.n, 65536));
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
            la.exactGrid   = vmk::exactGridFor(targets_.cap(), sizeof(R) == 8 || la.exactOnly);
            VM_HIP(vmk::launchLoss<R>(vmk::kTargetsPerLane, gLoss, stream_, la));
            VM_HIP(vmk::launchExact<R>(stream_, la));
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
    int             K_;
    hipStream_t     stream_ = nullptr, copy_ = nullptr; // compute; admissions' pool uploads
    hipEvent_t      ev0_ = nullptr, ev1_ = nullptr, evRead_ = nullptr, evCopy_ = nullptr;
    hipGraphExec_t  graph_ = nullptr;
    int             slots_ = 0, resident_ = 0, maxResident_ = 0, readSlots_ = 0, poolGrows_ = 0;
    bool            fixedSlots_ = false, changed_ = false, readPending_ = false, haveObserved_ = false;
    int64_t         maxS_ = 0, maxT_ = 0, maxC_ = 0;
    double budget_ = 0, residentBytes_ = 0, residentSites_ = 0, peakBytes_ = 0, gpuMs_ = 0, uploadBytes_ = 0;
    // Launch bookkeeping: uploads enqueued while launchSeq_ == u complete once
    // the readback of launch u + 1 has landed (doneSeq_ > u).
    int64_t launchSeq_ = 0, observedSeq_ = 0, doneSeq_ = 0;
    int64_t stageTag_[2] = {-1, -1};
    int     stageNext_   = 0;
    PinnedBuf                   stage_[2], dummy_[3];
    HBuf<vmk::InstState>        readback_;
    std::vector<vmk::InstState> observed_;
    RangeAlloc                  sites_, targets_, cells_;
    RangeAlloc                  tileAlloc_{1}, swAlloc_{1}, twAlloc_{1};
    std::vector<int64_t>        slotReq_;
    std::vector<SlotInfo>       slotInfo_;
    std::vector<vmk::SiteWork>  newWork_;
    std::vector<Upload>         uploads_;
    std::vector<Freed>          freed_;
    const vmk::InstState        zeroState_{};

    DBuf<V>                  dPos, dM, dV, dSpos, dTpos;
    DBuf<vmk::GradT<R>>      dGrad;
    DBuf<int>                dCellOf, dRank, dSidx, dCounts, dStart, dTslot, dCnt, dNewCnt, dExactLen;
    DBuf<vmk::ScanTile>      dTiles;
    DBuf<vmk::ScanCtl>       dScanCtl;
    DBuf<unsigned long long> dFlags;
    DBuf<vmk::ExactItem>     dExactQ;
    DBuf<uint8_t>            dFrozen;
    DBuf<vmk::Inst>          dInst;
    DBuf<vmk::InstState>     dState;
    DBuf<vmk::LossPartial>   dPartial;
    DBuf<vmk::SiteWork>      dSiteWork, dNewWork;
    DBuf<vmk::TargetWork>    dTargetWork;
};

// ---------------------------------------------------------------------------
// The loop
// ---------------------------------------------------------------------------
struct LoopHooks {
    std::function<void(int64_t req, double now, const vmk::Inst &in)> admitted;
    // Called before the slot is freed (the instance's sites are still resident).
    std::function<void(int64_t req, double now, const vmk::InstState &st, int slot, const vmk::Inst &in)> done;
};

template <class R>
void runLoop(Source &src, const vm_serve_config &cfg, int K, bool drainAll, double stopAt, LoopHooks &hooks,
             vm_serve_stats *stats, const std::function<void(Server<R> &, int, vm_vec2 *)> *download = nullptr,
             const std::function<vm_vec2 *(int64_t)> *sitesOut = nullptr)
{
    Server<R>     srv(cfg, K);
    const int64_t cap = src.capacity();
    std::vector<std::unique_ptr<Prepared>> prep(static_cast<size_t>(cap));
    std::vector<vmk::Inst>                 inst(static_cast<size_t>(cap));
    std::unique_ptr<std::atomic<int>[]>    ready(new std::atomic<int>[size_t(std::max<int64_t>(cap, 1))]);
    for (int64_t i = 0; i < cap; ++i) ready[i].store(0);
    Pipeline             pl;
    std::atomic<int64_t> nextPrep{0}, prepNs{0};
    std::mutex           errMu;
    std::string          prepError;
    const auto           t0      = Clock::now();
    auto                 elapsed = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    int device = 0;
    VM_HIP(hipGetDevice(&device));
    const unsigned           hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> workers;
    struct Join {
        std::vector<std::thread> &w;
        Pipeline                 &pl;
        Source                   &src;
        ~Join()
        {
            pl.stop.store(true);
            for (auto &t : w)
                if (t.joinable()) t.join();
            src.finish();
        }
    } join{workers, pl, src};
    for (unsigned w = 0; w < hw; ++w)
        workers.emplace_back([&] {
            (void)hipSetDevice(device);
            for (;;) {
                const int64_t i = nextPrep.fetch_add(1);
                if (i >= cap) return;
                while (!pl.stop.load() && i >= pl.limit())
                    std::this_thread::sleep_for(std::chrono::microseconds(200));
                if (pl.stop.load()) return;
                try {
                    const auto tp   = Clock::now();
                    prep[size_t(i)] = prepare<R>(src.spec(i, pl.stop), cfg, K);
                    prepNs.fetch_add(std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - tp).count());
                    ready[i].store(1, std::memory_order_release);
                } catch (const std::exception &e) {
                    std::lock_guard<std::mutex> lk(errMu);
                    prepError = e.what();
                    ready[i].store(2, std::memory_order_release);
                }
            }
        });
    src.start(pl);

    int64_t   nextAdmit = 0, completed = 0, arrived = 0;
    double    siteIters = 0, now = 0, loopS = 0;
    const int poll = std::max(1, cfg.poll_every);
    for (;;) {
        now     = elapsed();
        arrived = src.arrivedBy(now);
        pl.arrived.store(arrived, std::memory_order_release);
        if (!drainAll && now >= stopAt) break;
        if (src.closed() && completed == arrived) break;
        // The next iterations run while this thread retires, admits and
        // enqueues uploads (they queue behind the iterations).
        const bool running = srv.resident() > 0;
        if (running) srv.launch(poll);
        const auto tl = Clock::now();
        srv.retire([&](int64_t req, const vmk::InstState &st, int slot) {
            const vmk::Inst &in = inst[size_t(req)];
            if (sitesOut && download) {
                vm_vec2 *out = (*sitesOut)(req);
                if (out) (*download)(srv, slot, out);
            }
            hooks.done(req, elapsed(), st, slot, in);
            siteIters += double(in.S) * st.iter;
            ++completed;
        });
        // Admit prepared requests in arrival order.
        while (nextAdmit < arrived && ready[nextAdmit].load(std::memory_order_acquire) != 0) {
            if (ready[nextAdmit].load() == 2) {
                std::string msg;
                {
                    std::lock_guard<std::mutex> lk(errMu);
                    msg = prepError;
                }
                throw VmError(VM_EINVAL, "request preparation failed: " + msg);
            }
            const vmk::Inst in0 = prep[size_t(nextAdmit)]->in;
            const Admit     a   = srv.admit(nextAdmit, prep[size_t(nextAdmit)]);
            if (a == Admit::Never) throw VmError(VM_ENOMEM, "request does not fit the HBM budget or device memory");
            if (a == Admit::Wait) break;
            inst[size_t(nextAdmit)] = in0;
            hooks.admitted(nextAdmit, elapsed(), in0);
            ++nextAdmit;
        }
        pl.admitted.store(nextAdmit, std::memory_order_release);
        srv.refresh();
        loopS += std::chrono::duration<double>(Clock::now() - tl).count();
        if (running) srv.waitReadback();
        else if (srv.resident() == 0) std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    if (stats) {
        *stats                   = vm_serve_stats{};
        stats->wall_s            = now;
        stats->site_iters        = siteIters;
        stats->site_iters_run    = siteIters + srv.residentSiteIters();
        stats->completed         = int(completed);
        stats->max_resident_seen = srv.maxResident();
        stats->peak_bytes        = srv.peakBytes();
        stats->arrived           = int(arrived);
        stats->admitted          = int(nextAdmit);
        stats->pool_grows        = srv.poolGrows();
        stats->pool_bytes        = srv.poolBytes();
        stats->gpu_busy_s        = srv.gpuSeconds();
        stats->oldest_wait_s     = nextAdmit < arrived ? now - src.arrival(nextAdmit) : 0.0;
        stats->prep_busy_s       = double(prepNs.load()) * 1e-9;
        stats->loop_busy_s       = loopS;
        stats->upload_bytes      = srv.uploadBytes();
        stats->prep_workers      = int(hw);
    }
}

template <class R>
void serveList(const vm_fit_desc *reqs, const double *arrival, int n, const vm_serve_config &cfg, vm_serve_result *res,
               vm_serve_stats *stats)
{
    int K = 0;
    for (int i = 0; i < n; ++i) {
        const int k = reqs[i].loss_k > 0 ? reqs[i].loss_k : 16;
        if (K == 0) K = k;
        if (k != K) throw VmError(VM_EINVAL, "all requests of one serving run must share loss_k");
        if (i > 0 && arrival[i] < arrival[i - 1]) throw VmError(VM_EINVAL, "arrival times must be non-decreasing");
    }
    ListSource src(reqs, arrival, n);
    std::vector<double> admitT(static_cast<size_t>(n), -1.0);
    LoopHooks hooks;
    hooks.admitted = [&](int64_t req, double now, const vmk::Inst &) { admitT[size_t(req)] = now; };
    hooks.done     = [&](int64_t req, double now, const vmk::InstState &st, int, const vmk::Inst &) {
        vm_serve_result &r = res[req];
        r.latency_s        = now - arrival[req];
        r.queue_s          = admitT[size_t(req)] - arrival[req];
        r.iterations       = st.iter;
        r.final_mean_loss  = st.count > 0 ? st.loss / double(st.count) : 0.0;
    };
    const std::function<void(Server<R> &, int, vm_vec2 *)> dl = [](Server<R> &s, int slot, vm_vec2 *out) {
        s.downloadSites(slot, out);
    };
    const std::function<vm_vec2 *(int64_t)> where = [&](int64_t req) { return res[req].final_sites; };
    runLoop<R>(src, cfg, K, true, 0.0, hooks, stats, &dl, &where);
}

template <class R>
void serveSynthetic(const vm_serve_stream &st, const vm_serve_config &cfg, vm_serve_record *rec, int maxRec,
                    int *numRec, vm_serve_stats *stats)
{
    SynthSource src(st, maxRec, rec);
    LoopHooks   hooks;
    hooks.admitted = [&](int64_t req, double now, const vmk::Inst &in) {
        rec[req].admit_s = now, rec[req].num_sites = in.S, rec[req].num_targets = in.T;
    };
    hooks.done = [&](int64_t req, double now, const vmk::InstState &s, int, const vmk::Inst &) {
        rec[req].done_s          = now;
        rec[req].iterations      = s.iter;
        rec[req].final_mean_loss = s.count > 0 ? s.loss / double(s.count) : 0.0;
    };
    vm_serve_stats local{};
    runLoop<R>(src, cfg, st.loss_k > 0 ? st.loss_k : 16, st.drain != 0, st.duration_s, hooks, &local);
    local.sample_busy_s = src.sampleSeconds(); // the sampler thread has been joined
    local.prep_wait_s   = src.waitSeconds();
    *numRec             = local.arrived;
    if (stats) *stats = local;
}

void checkConfig(const vm_serve_config *cfg)
{
    if (!cfg) throw VmError(VM_EINVAL, "invalid arguments");
    if (!(cfg->hbm_budget_bytes > 0)) throw VmError(VM_EINVAL, "hbm budget must be positive");
    if (cfg->precision != 0 && cfg->precision != 1) throw VmError(VM_EINVAL, "precision must be 0 or 1");
    if (cfg->max_resident < 0 || cfg->max_resident_sites < 0 || cfg->initial_pool_bytes < 0)
        throw VmError(VM_EINVAL, "negative serving limit");
}

} // namespace

extern "C" int vm_serve(const vm_fit_desc *requests, const double *arrival, int n, const vm_serve_config *cfg,
                        vm_serve_result *results, vm_serve_stats *stats)
{
    return vmrt::guardedCall([&] {
        if (n < 0 || (n > 0 && (!requests || !arrival || !results))) throw VmError(VM_EINVAL, "invalid arguments");
        checkConfig(cfg);
        if (n == 0) return;
        if (cfg->precision == 1) serveList<double>(requests, arrival, n, *cfg, results, stats);
        else serveList<float>(requests, arrival, n, *cfg, results, stats);
    });
}

extern "C" int vm_serve_synthetic(const vm_serve_stream *stream, const vm_serve_config *cfg, vm_serve_record *records,
                                  int max_records, int *num_records, vm_serve_stats *stats)
{
    return vmrt::guardedCall([&] {
        if (!stream || !records || max_records < 1 || !num_records) throw VmError(VM_EINVAL, "invalid arguments");
        checkStream(*stream);
        checkConfig(cfg);
        *num_records = 0;
        if (cfg->precision == 1) serveSynthetic<double>(*stream, *cfg, records, max_records, num_records, stats);
        else serveSynthetic<float>(*stream, *cfg, records, max_records, num_records, stats);
    });
}

extern "C" int vm_serve_synth_plan(const vm_serve_stream *stream, int64_t index, double *arrival_s, int *num_sites,
                                   float *radius, unsigned *sampler_seed)
{
    return vmrt::guardedCall([&] {
        if (!stream || index < 0) throw VmError(VM_EINVAL, "invalid arguments");
        checkStream(*stream);
        double t = 0;
        for (int64_t j = 0; j <= index; ++j) t += planOf(*stream, j).gap;
        const SynthPlan p = planOf(*stream, index);
        if (arrival_s) *arrival_s = t;
        if (num_sites) *num_sites = p.sites;
        if (radius) *radius = p.radius;
        if (sampler_seed) *sampler_seed = p.seed;
    });
}

extern "C" int vm_serve_synth_targets(const vm_serve_stream *stream, int64_t index, int num_sites, vm_vec2 *targets,
                                      int capacity)
{
    return vmrt::guardedCall([&] {
        if (!stream || index < 0 || num_sites < 0 || (!targets && capacity > 0))
            throw VmError(VM_EINVAL, "invalid arguments");
        checkStream(*stream);
        if (int64_t(stream->targets_per_site) * num_sites > capacity)
            throw VmError(VM_EINVAL, "target buffer holds " + std::to_string(capacity) + " points; the request has " +
                                         std::to_string(int64_t(stream->targets_per_site) * num_sites));
        synthTargets(*stream, index, num_sites, targets);
    });
}
