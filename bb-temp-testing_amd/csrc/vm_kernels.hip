// vm_kernels.hip -- HIP kernels of the VoroMesh optimizer step for gfx950.
//
// One optimizer iteration over a packed table of instances is four launches:
//   scan     one workgroup per instance: exclusive scan of the per-cell site
//            counts into cell starts (and zeroes the counts), begins the step
//            (iter += 1, clears the loss accumulators, convergence test);
//   scatter  site chunks: counting-sort scatter into cell order;
//   loss     target chunks (one wavefront = up to 64 targets of one hash
//            cell): stage the sites of the surrounding cell block into LDS,
//            k-NN by ring search with a register top-K, VoroLoss and its
//            gradient accumulated in LDS and flushed with one global atomic
//            per touched site;
//   adam     site chunks: Adam + MultiStepLR + clamp, zero the gradient, and
//            bin the new position (count + rank) for the next scan.
// Reference: src/voromesh/voro_optimizer.cu (binKernel :168, scatterKernel
// :179, vorolossKernel :206, adamKernel :271, voronoiCellKernel :361) and
// src/poisson/poisson_disk.cu (dartPhaseKernel :138, clearRegionKernel :173).
//
// Arithmetic contract (shared with the CPU oracle): squared distances are
// fma(dy, dy, dx*dx), dot products fma(a.y, b.y, a.x*b.x), k-NN lists are
// ordered by (d2, site index); the file is built with -ffp-contract=off and
// correctly rounded division and square root, so every per-target term is
// bit-identical to the oracle's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <float.h>
#include <limits.h>
#include <stdint.h>

#include "vm_kernels.h"

namespace vmk {

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pcg_hash(uint32_t v)
{
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word  = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

__device__ __forceinline__ float rand01_loss(uint32_t index, uint32_t iter, uint32_t seed)
{
    uint32_t h = pcg_hash(index ^ pcg_hash(iter * 2654435761u ^ pcg_hash(seed)));
    return float(h & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float rand01_poisson(uint32_t cell, uint32_t round, uint32_t seed, uint32_t stream)
{
    uint32_t h = pcg_hash(cell ^ pcg_hash(round * 2654435761u ^ pcg_hash(seed + stream * 0x9E3779B9u)));
    return float(h & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}

// Wave-uniform records that no kernel writes while it runs (instances and work
// tables, written by the host; queue entries and instance states written by an
// earlier launch) are read through the constant address space: the loads are
// scalar, their fields stay in scalar registers, and stores elsewhere in the
// kernel do not force re-reads.  The address must be wave-uniform.
template <class T> __device__ __forceinline__ T loadUniform(const T *p)
{
    static_assert(sizeof(T) % 4 == 0 && alignof(T) >= 4, "records of whole dwords");
#if defined(__HIP_DEVICE_COMPILE__)
    // Dword by dword: a whole-record copy from this address space is partly
    // lowered to vector loads.
    using CU32 = const __attribute__((address_space(4))) uint32_t;
    CU32    *q = (CU32 *)p;
    uint32_t w[sizeof(T) / 4];
#pragma unroll
    for (unsigned k = 0; k < sizeof(T) / 4; ++k) w[k] = q[k];
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
#else
    return *p; // host pass only parses device code
#endif
}

__device__ __forceinline__ float  fmaR(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaR(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float  sqrtR(float a) { return sqrtf(a); }
__device__ __forceinline__ double sqrtR(double a) { return sqrt(a); }

// Correctly rounded quotients a / b for 2^-60 <= |b| <= 2^60 and a == 0 or
// 2^-70 <= |a| <= 2^60 (so the quotient and both residuals below stay normal):
// the core of the division sequence (reciprocal, one refinement, quotient and
// two residual corrections), whose operand scaling and special-case fixup
// leave such operands unchanged.  recipRefined(b) is the refined reciprocal
// the quotients by b share.
__device__ __forceinline__ float recipRefined(float b)
{
    const float r = __builtin_amdgcn_rcpf(b);
    return __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r);
}
__device__ __forceinline__ float divCore(float a, float b, float r1)
{
    const float q0 = a * r1;
    const float q1 = __builtin_fmaf(__builtin_fmaf(-b, q0, a), r1, q0);
    return __builtin_fmaf(__builtin_fmaf(-b, q1, a), r1, q1);
}

// Correctly rounded sqrt for finite x >= 2^-96: v_sqrt_f32 is within 1 ulp,
// and the residuals x - s' s of its neighbours s' pick the rounded root (the
// same steps sqrtf takes there, without its rescaling of tiny inputs and its
// zero / infinity cases).
__device__ __forceinline__ float sqrtNormal(float x)
{
    const float s  = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    float       r  = __builtin_fmaf(-sd, s, x) <= 0.0f ? sd : s;
    return __builtin_fmaf(-su, s, x) > 0.0f ? su : r;
}

template <class R> __device__ __forceinline__ R dist2(V2<R> p, V2<R> q)
{
    R dx = p.x - q.x, dy = p.y - q.y;
    return fmaR(dy, dy, dx * dx);
}

template <class R> __device__ __forceinline__ R dotR(V2<R> a, V2<R> b) { return fmaR(a.y, b.y, a.x * b.x); }

template <class R> __device__ __forceinline__ int cellCoord(R v, R inv, int n)
{
    R   f = v * inv;
    int c = (f >= R(0)) ? (f < R(2147483520.0) ? int(f) : n - 1) : 0;
    return min(max(c, 0), n - 1);
}

template <class R> __device__ __forceinline__ R ringRel();
template <> __device__ __forceinline__ float  ringRel<float>() { return 1e-6f; }
template <> __device__ __forceinline__ double ringRel<double>() { return 1e-13; }

__device__ __forceinline__ double powi(double b, int e)
{
    double r = 1.0;
    while (e > 0) {
        if (e & 1) r = r * b;
        b = b * b;
        e >>= 1;
    }
    return r;
}

// Gradient accumulation.  FP32: a term c becomes round(c * 2^shift) (double
// scaling is exact, rounding to nearest even), clamped to +-2^51, and is summed
// with 64-bit integer atomics; the sum is order-independent.  FP64: double atomics.
// scale = 2^shift, shift in [3, 47].  The (finite) term is clamped to
// +-2^(51 - shift) in FP32 before the exact scaling, which is the same as
// clamping c * scale to +-2^51 after it.
__device__ __forceinline__ float quantLimit(double scale) // 2^(51 - shift)
{
    const int g = int(uint64_t(__double_as_longlong(scale)) >> 52) - 1023;
    return __uint_as_float(uint32_t(127 + 51 - g) << 23);
}
// The rounding to an integer is one FP64 addition: for |x| <= 2^51 the sum
// x + 1.5 * 2^52 lies in [2^52, 2^53], where doubles are spaced 1 apart, so it
// is 1.5 * 2^52 + round(x) (ties to even: 1.5 * 2^52 is even), and its bit
// pattern minus that of 1.5 * 2^52 is round(x) in two's complement.  The
// subtrahend's low word is zero, so only the high word changes.  x = c * scale
// is exact, so one fma forms the same sum.
__device__ __forceinline__ long long quantize(float c, double scale)
{
    constexpr double   kMagic   = 0x1.8p52;
    constexpr uint32_t kMagicHi = 0x43380000u; // high word of kMagic's bit pattern
    const float        lim      = quantLimit(scale);
    const double       t        = __builtin_fma(double(__builtin_amdgcn_fmed3f(c, -lim, lim)), scale, kMagic);
    const uint64_t     b        = uint64_t(__double_as_longlong(t));
    return (long long)((uint64_t(uint32_t(b >> 32) - kMagicHi) << 32) | (b & 0xFFFFFFFFull));
}

template <class R> struct GradOps;
template <> struct GradOps<float> {
    using T = long long;
    static __device__ __forceinline__ void add(T *p, float c, double scale)
    {
        atomicAdd(reinterpret_cast<unsigned long long *>(p), (unsigned long long)quantize(c, scale));
    }
    static __device__ __forceinline__ void addRaw(T *p, T v)
    {
        atomicAdd(reinterpret_cast<unsigned long long *>(p), (unsigned long long)v);
    }
    static __device__ __forceinline__ float toReal(T v, double inv) { return float(double(v) * inv); }
};
template <> struct GradOps<double> {
    using T = double;
    static __device__ __forceinline__ void add(T *p, double c, double) { unsafeAtomicAdd(p, c); }
    static __device__ __forceinline__ void addRaw(T *p, T v) { unsafeAtomicAdd(p, v); }
    static __device__ __forceinline__ double toReal(T v, double) { return v; }
};

// ---------------------------------------------------------------------------
// k-NN keys: total order (d2, site index).  The low word of a key is
// (site index << kLdsSlotBits) | LDS slot, so ties on d2 are broken by the
// site index and the slot rides along for free.
// ---------------------------------------------------------------------------
template <class R> struct Key;

template <> struct Key<float> {
    uint64_t k;
    static __device__ __forceinline__ Key make(float d2, uint32_t low)
    {
        return Key{(uint64_t(__float_as_uint(d2)) << 32) | low};
    }
    static __device__ __forceinline__ Key none() { return Key{~0ull}; }
    __device__ __forceinline__ bool     lt(const Key &o) const { return k < o.k; }
    __device__ __forceinline__ bool     valid() const { return k != ~0ull; }
    __device__ __forceinline__ uint32_t low() const { return uint32_t(k); }
    __device__ __forceinline__ float    d2() const { return __uint_as_float(uint32_t(k >> 32)); }
};

template <> struct Key<double> {
    double   d;
    uint32_t i;
    static __device__ __forceinline__ Key make(double d2, uint32_t low) { return Key{d2, low}; }
    static __device__ __forceinline__ Key none() { return Key{DBL_MAX, 0xFFFFFFFFu}; }
    __device__ __forceinline__ bool     lt(const Key &o) const { return d < o.d || (d == o.d && i < o.i); }
    __device__ __forceinline__ bool     valid() const { return i != 0xFFFFFFFFu; }
    __device__ __forceinline__ uint32_t low() const { return i; }
    __device__ __forceinline__ double   d2() const { return d; }
};

__device__ __forceinline__ int keySite(uint32_t low) { return int(low >> kLdsSlotBits); }
__device__ __forceinline__ int keySlot(uint32_t low) { return int(low & uint32_t(kLdsNone)); }

// Static-index insertion into an ascending top-K list held in registers.
template <int K, class KeyT> __device__ __forceinline__ void topkInsert(KeyT (&keys)[K], KeyT nk)
{
    if (!nk.lt(keys[K - 1])) return;
#pragma unroll
    for (int k = K - 1; k > 0; --k) {
        bool before = nk.lt(keys[k - 1]);
        keys[k]     = before ? keys[k - 1] : (nk.lt(keys[k]) ? nk : keys[k]);
    }
    keys[0] = nk.lt(keys[0]) ? nk : keys[0];
}

// Per-thread ring search over the global sorted grid of one instance (used by
// the clipping and the diagnostic k-NN kernels).  Sites are stored in R and
// were binned in R arithmetic; distances and keys are in KR (the clipping
// orders its neighbours by FP64 distances).  Same stop rule as the loss kernel
// and the oracle: conservative for R-precision binning, so the list equals
// brute force.
template <int K, class R, class KR = R>
__device__ void knnGlobal(const Inst &in, const V2<R> *spos, const int *sidx, const int *start, V2<R> qr,
                          Key<KR> (&keys)[K])
{
#pragma unroll
    for (int k = 0; k < K; ++k) keys[k] = Key<KR>::none();
    const R      cellR  = R(in.cell);
    const R      inv    = R(1) / cellR;
    const KR     cell   = KR(cellR);
    const KR     rel    = KR(ringRel<R>());
    const V2<KR> q{KR(qr.x), KR(qr.y)};
    const KR     extent = fmax(KR(in.ext), fmax(fabs(q.x), fabs(q.y)));
    const KR     margin = rel * extent;
    const int    cx = cellCoord(qr.x, inv, in.gw), cy = cellCoord(qr.y, inv, in.gh);
    for (int ring = 0;; ++ring) {
        int x0 = cx - ring, x1 = cx + ring, y0 = cy - ring, y1 = cy + ring;
        for (int y = max(y0, 0); y <= min(y1, in.gh - 1); ++y) {
            int stride = (y == y0 || y == y1) ? 1 : max(x1 - x0, 1);
            for (int x = x0; x <= x1; x += stride) {
                if (x < 0 || x >= in.gw) continue;
                int c = y * in.gw + x;
                for (int s = start[c]; s < start[c + 1]; ++s) {
                    V2<KR> p{KR(spos[s].x), KR(spos[s].y)};
                    topkInsert<K>(keys, Key<KR>::make(dist2(p, q), (uint32_t(sidx[s]) << kLdsSlotBits) | kLdsNone));
                }
            }
        }
        if (x0 <= 0 && y0 <= 0 && x1 >= in.gw - 1 && y1 >= in.gh - 1) break;
        KR safe = fmin(fmin(q.x - KR(x0) * cell, KR(x1 + 1) * cell - q.x), fmin(q.y - KR(y0) * cell, KR(y1 + 1) * cell - q.y));
        safe    = fmax(safe - margin, KR(0));
        if (keys[K - 1].valid() && keys[K - 1].d2() * (KR(1) + rel) < safe * safe) break;
    }
}

// ---------------------------------------------------------------------------
// 1. Scan: exclusive scan of the per-cell site counts into cell starts (and
//    zeroes the counts), over tiles of kScanTile cells.  Single pass with
//    decoupled look-back: a workgroup takes the next tile from a counter (so
//    every tile it waits on belongs to a workgroup that is already running),
//    publishes the tile's sum, then walks back over the earlier tiles of the
//    same instance until one with an inclusive prefix.  The first tile of an
//    instance begins its step (iter += 1, clears the loss sums, convergence
//    test); a finished instance's tiles publish "skip" and leave its grid alone.
//    Flag words carry an epoch that the scatter kernel advances after every
//    scan, so nothing needs clearing between launches.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 1024;
constexpr int kScanItems   = 16;
static_assert(kScanThreads * kScanItems == kScanTileCells, "one scan tile per workgroup pass");

// Inclusive prefix sum over the 64 lanes of a wavefront (DPP row shifts, then
// row broadcasts); every lane must be active.
__device__ __forceinline__ int waveInclusiveSum(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true); // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true); // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, true); // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, true); // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false); // row_bcast:15 into rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false); // row_bcast:31 into rows 2, 3
    return v;
}

// Look-back flag: epoch (30 bits) | status (2 bits) | value (32 bits).
constexpr uint32_t kTileAggregate = 1, kTilePrefix = 2, kTileSkip = 3;

__device__ __forceinline__ void publishTile(unsigned long long *f, uint32_t epoch, uint32_t status, int value)
{
    const unsigned long long w = (uint64_t(epoch & 0x3FFFFFFFu) << 34) | (uint64_t(status) << 32) | uint32_t(value);
    __hip_atomic_store(f, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// What the start of a step reads from the instance's record and state, loaded
// in one round (the decision below then waits for nothing else).
struct StepIn {
    int     iter, done, maxIter;
    float   tol;
    int64_t count;
    double  loss, prevMean;
};
__device__ __forceinline__ StepIn stepLoad(const Inst &in, const InstState *s)
{
    return StepIn{s->iter, s->done, in.maxIter, in.tol, s->count, s->loss, s->prevMean};
}

// Begins the step of an instance (first tile); returns 1 when it is skipped.
__device__ __forceinline__ int beginStep(const StepIn &v, InstState *s, int mode)
{
    if (mode != kModeIterate) {
        s->evalLoss  = 0.0;
        s->evalCount = 0;
        return 0;
    }
    int d = v.done;
    if (!d) {
        if (v.iter > 0 && v.count > 0 && v.tol > 0.0f) {
            const double mean = v.loss / double(v.count);
            if (v.iter > 1 && fabs(mean - v.prevMean) < double(v.tol) * fabs(v.prevMean)) d = 1;
            s->prevMean = mean;
        }
        if (v.maxIter > 0 && v.iter >= v.maxIter) d = 1;
        if (!d) {
            s->iter  = v.iter + 1;
            s->count = 0;
            s->loss  = 0.0;
        }
        s->done = d;
    }
    return d;
}

// One scan tile: exclusive scan of counts into starts (and zero the counts).
// Tiles of multi-tile instances find their prefix by look-back (`dyn`); a
// single-tile instance's tile starts at zero and never waits.
__device__ __forceinline__ void scanTile(const ScanTile &tl, int tile, bool dyn, uint32_t epoch, const Inst *insts,
                                         InstState *state, LossPartial *partial, unsigned long long *flags,
                                         int *counts, int *start, int mode, int *waveSum, int *sShared)
{
    const int   lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const Inst &in   = insts[tl.inst];
    int        *cnt  = counts + in.cellOff + tl.begin;
    int        *st   = start + in.cellOff + tl.begin;
    const int   n    = tl.count;

    // Each thread owns kScanItems consecutive cells (16-byte vector loads and
    // stores: cell offsets and tile starts are multiples of 64).
    // The first tile's thread 0 reads the step's inputs now, beside the counts.
    StepIn si{};
    if (threadIdx.x == 0 && tl.first) si = stepLoad(in, state + tl.inst);
    const int  i0   = int(threadIdx.x) * kScanItems;
    const bool full = i0 + kScanItems <= n;
    int        v[kScanItems];
    if (full) {
#pragma unroll
        for (int j = 0; j < kScanItems / 4; ++j) {
            const int4 q = reinterpret_cast<const int4 *>(cnt + i0)[j];
            v[4 * j] = q.x, v[4 * j + 1] = q.y, v[4 * j + 2] = q.z, v[4 * j + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) v[j] = (i0 + j < n) ? cnt[i0 + j] : 0;
    }
    int tot = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) tot += v[j];
    const int inc = waveInclusiveSum(tot);
    if (lane == 63) waveSum[wave] = inc;
    __syncthreads();
    int wpre = 0, btot = 0;
#pragma unroll
    for (int k = 0; k < kScanThreads / 64; ++k) {
        const int s = waveSum[k];
        wpre += (k < wave) ? s : 0;
        btot += s;
    }

    // ---- tile prefix (look-back over the instance's earlier tiles) ---------
    // Wave 0 reads a window of up to 64 predecessor flags at once (lane l the
    // flag of tile j0 - l); the nearest tile with a prefix (or a skip) ends
    // the walk once every flag before it has its aggregate, otherwise the
    // window's aggregates are added and the window moves back.  The instance's
    // first tile always publishes a prefix or a skip.
    if (threadIdx.x < 64) {
        int skip = 0, prefix = 0;
        if (tl.first) {
            if (lane == 0) {
                skip = beginStep(si, state + tl.inst, mode);
                if (dyn) publishTile(&flags[tile], epoch, skip ? kTileSkip : kTilePrefix, btot);
            }
        } else {
            if (lane == 0) publishTile(&flags[tile], epoch, kTileAggregate, btot);
            const int      jmin = tile - tl.begin / kScanTileCells; // the instance's first tile
            const uint32_t ep   = epoch & 0x3FFFFFFFu;
            for (int j0 = tile - 1;;) {
                const int          j  = j0 - lane;
                unsigned long long f  = 0;
                if (j >= jmin) f = __hip_atomic_load(&flags[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t st    = uint32_t(f >> 32) & 3u;
                const bool     valid = j >= jmin && uint32_t(f >> 34) == ep && st != 0u;
                const uint64_t have  = __ballot(valid);
                const uint64_t stop  = __ballot(valid && st != kTileAggregate);
                const uint64_t inWin = __ballot(j >= jmin);
                const uint64_t need  = stop ? (~0ull >> (63 - __builtin_ctzll(stop))) : inWin;
                if ((have & need) != need) { // a flag the walk needs is not published yet
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                const int v   = (need >> lane) & 1ull ? int(uint32_t(f)) : 0;
                const int sum = __builtin_amdgcn_readlane(waveInclusiveSum(v), 63);
                if (!stop) {
                    prefix += sum;
                    j0 -= 64;
                    continue;
                }
                const int p = __builtin_ctzll(stop);
                skip        = __builtin_amdgcn_readlane(int(st), p) == int(kTileSkip);
                prefix += sum;
                break;
            }
            if (lane == 0) publishTile(&flags[tile], epoch, skip ? kTileSkip : kTilePrefix, prefix + btot);
        }
        if (lane == 0) sShared[0] = skip, sShared[1] = prefix;
    }
    __syncthreads();
    if (!sShared[0]) {
        if (tl.first && mode == kModeIterate && threadIdx.x < kLossPartials) {
            partial[int64_t(tl.inst) * kLossPartials + threadIdx.x].loss  = 0.0;
            partial[int64_t(tl.inst) * kLossPartials + threadIdx.x].count = 0;
        }
        int ex = sShared[1] + wpre + inc - tot;
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const int c = v[j];
            v[j]        = ex;
            ex += c;
        }
        if (full) {
#pragma unroll
            for (int j = 0; j < kScanItems / 4; ++j) {
                reinterpret_cast<int4 *>(st + i0)[j]  = int4{v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]};
                reinterpret_cast<int4 *>(cnt + i0)[j] = int4{0, 0, 0, 0};
            }
        } else {
#pragma unroll
            for (int j = 0; j < kScanItems; ++j)
                if (i0 + j < n) st[i0 + j] = v[j], cnt[i0 + j] = 0;
        }
    }
    __syncthreads();
}

// Counting-sort scatter of one instance's sites by the workgroup that just
// scanned its (single) tile: the starts are this workgroup's own writes.
__device__ void scatterInstance(const Inst &in, const ScanScatter &sc, const int *start)
{
    for (int i = int(threadIdx.x); i < in.S; i += kScanThreads) {
        const int64_t g  = in.siteOff + i;
        const int     ce = sc.cellOf[g];
        if (ce < 0) continue; // empty slot
        const int64_t d = in.siteOff + start[in.cellOff + ce] + sc.rank[g];
        if (sc.posBytes == 8)
            reinterpret_cast<int2 *>(sc.spos)[d] = reinterpret_cast<const int2 *>(sc.pos)[g];
        else
            reinterpret_cast<int4 *>(sc.spos)[d] = reinterpret_cast<const int4 *>(sc.pos)[g];
        sc.sidx[d] = i;
    }
}

// Tiles [0, nDyn) belong to multi-tile instances and are taken in order from
// the counter (a workgroup only ever waits on tiles that running workgroups
// took before it); tiles [nDyn, nTiles) are single-tile instances, spread
// statically.  The scatter kernel, which follows every scan, rearms the
// counter and advances the flag epoch.  When no instance has more than one
// tile (sc.posBytes != 0) each workgroup scatters the instances it scanned and
// there is no scatter launch (nothing to rearm: only multi-tile instances use
// the counter and the flags).
__global__ __launch_bounds__(kScanThreads) void scanKernel(const Inst *__restrict__ insts, InstState *state,
                                                          LossPartial *partial, const ScanTile *__restrict__ tiles,
                                                          ScanCtl *ctl, unsigned long long *flags, int *counts,
                                                          int *start, int mode, int *exactLen, int nQueues,
                                                          ScanScatter sc)
{
    __shared__ int waveSum[kScanThreads / 64];
    __shared__ int sShared[3];
    // Empty the exact-selection queues the following loss launches fill.
    if (blockIdx.x == 0 && int(threadIdx.x) < nQueues) exactLen[threadIdx.x] = 0;
    const int      nTiles = ctl->nTiles, nDyn = ctl->nDyn;
    const uint32_t epoch  = ctl->epoch;
    if (nDyn > 0) {
        for (;;) {
            if (threadIdx.x == 0) {
                int t = nDyn;
                if (__hip_atomic_load(&ctl->next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nDyn)
                    t = atomicAdd(&ctl->next, 1);
                sShared[2] = t;
            }
            __syncthreads();
            const int tile = sShared[2];
            __syncthreads();
            if (tile >= nDyn) break;
            scanTile(tiles[tile], tile, true, epoch, insts, state, partial, flags, counts, start, mode, waveSum,
                     sShared);
        }
    }
    for (int tile = nDyn + blockIdx.x; tile < nTiles; tile += gridDim.x) {
        const ScanTile tl = tiles[tile];
        scanTile(tl, tile, false, epoch, insts, state, partial, flags, counts, start, mode, waveSum, sShared);
        if (sc.posBytes && !sShared[0]) scatterInstance(insts[tl.inst], sc, start); // sShared[0]: step skipped
    }
}

// ---------------------------------------------------------------------------
// 2. Scatter into cell order (voro_optimizer.cu:179-187)
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(kSiteChunk) void scatterKernel(const Inst *__restrict__ insts,
                                                            const InstState *__restrict__ state,
                                                            const SiteWork *__restrict__ work, const int *nWorkPtr,
                                                            const V2<R> *__restrict__ pos,
                                                            const int *__restrict__ cellOf,
                                                            const int *__restrict__ rank,
                                                            const int *__restrict__ start, V2<R> *spos, int *sidx,
                                                            int mode, ScanCtl *ctl)
{
    // The scan before this launch is complete: rearm its tile counter and move
    // its look-back flags to a new epoch.
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->next  = 0;
        ctl->epoch = ctl->epoch + 1;
    }
    (void)insts; // the work items carry the offsets
    const int nWork = *nWorkPtr;
    for (int w = blockIdx.x; w < nWork; w += gridDim.x) {
        const SiteWork wk = loadUniform(work + w);
        const int      i  = wk.begin + int(threadIdx.x);
        if (i >= wk.begin + wk.count) continue;
        // The site loads go out with the instance's done flag.
        const int64_t g  = wk.siteOff + i;
        const int     ce = cellOf[g];
        const int     rk = rank[g];
        const V2<R>   p  = pos[g];
        if (mode == kModeIterate && loadUniform(&state[wk.inst].done)) continue;
        if (ce < 0) continue; // empty slot
        const int s = start[wk.cellOff + ce] + rk;
        spos[wk.siteOff + s] = p;
        sidx[wk.siteOff + s] = i;
    }
}

// ---------------------------------------------------------------------------
// 3. VoroLoss + gradient (voro_optimizer.cu:199-266)
//
// A block holds four independent wavefronts; each takes up to 64 targets of
// one hash cell that lie within rho of a chunk centre q (both precomputed on
// the host: targets do not move).  The sites of the (2r+1)^2 cell block around
// the cell are staged in the wavefront's LDS slice.
//
// Fast path.  With D >= the distance from q to its K-th nearest staged site,
// every target p of the chunk has D_K(p) <= D + rho, so its K nearest sites lie
// within D + 2 rho of q.  When the staged block provably contains that disc,
// the candidates (q's own nearest first) are compacted in LDS.  In FP32 each
// lane then resolves its term by screening (screenTerm below) and evaluates the
// exact formula once; lanes the screen cannot prove select their exact top-K
// under the (d2, index) order as an unsorted set with a tracked maximum.
// Otherwise the lanes fall back to a ring search.  Wave-wide reductions use DPP
// (no LDS round trips).
// ---------------------------------------------------------------------------
constexpr int kLossWaves     = 4;
constexpr int kLossThreads   = 64 * kLossWaves;
// Workgroups go to the 8 XCDs in turn (workgroup b to XCD b % 8), and each XCD
// has its own L2.  Work items are handed out in runs of kXcdRun consecutive
// items per XCD, runs of successive XCDs interleaved, so neighbouring chunks
// (whose staged blocks share sites and cell starts) meet in one L2 while every
// XCD still gets an even share.  The launch rounds the grid up to a multiple
// of 2 kXcdRun workgroups, which makes the mapping a permutation of each
// round of gridDim.x * kLossWaves items.
constexpr int kXcds          = 8;
constexpr int kXcdRun        = 16;
constexpr int kLossStage     = 128; // staged sites per wavefront, two per lane (block of <= 7 x 7 cells)
// Bisection of q's K-th squared distance over the bit patterns of the staged
// squared distances (non-negative floats order as their patterns): from
// [largest / 2^8, largest] (or [0, largest] when K sites lie below the lower
// end), 10 halvings leave a bracket of 2^16 patterns, 1/128 of a binade.
constexpr int      kBisectSteps = 10;
constexpr uint32_t kBisectSpan  = 8u << 23; // eight binades

// Screen table of a staged chunk: one row per distinct nearest site c_r of the
// chunk's targets, one column per inner-band candidate c_i.  Entry (r, i) holds
// the unit vector n from c_r to c_i and -|c_i - c_r| / 2, so the signed
// distance of a target p with nearest site c_r to their bisector is
// fma(ux, nx, fma(uy, ny, off)) with u = p - c_r.  Rows are kTblPad entries
// longer than the columns, which puts the rows a wavefront reads at once in
// different LDS banks.
constexpr int kTblEntries = 128;
constexpr int kTblPad     = 2;

template <class R> struct LossLds {
    V2<R>    pos[kLossStage];
    int      idx[kLossStage];
    GradT<R> grad[2 * kLossStage];
    // The table lives while the chunk's terms are screened; the exact
    // selection inside the chunk runs afterwards and reuses the space.
    union {
        struct {
            uint64_t xkey[kLossStage]; // (d2, index) keys of the candidates
            int      xtop[32];         // slots of the K nearest, in order
        } x;
        struct {
            float2 n[kTblEntries];
            float  off[kTblEntries];
        } t;
    } u;
};

// Outcome of one lane's VoroLoss term: nearest site c0 (instance index i0, LDS
// slot sl0 or kLdsNone), bisector partner j (bj, slj), d = signed distance to
// their bisector, ev = cj - c0 and l = |ev|.
template <class R> struct LaneTerm {
    int   i0 = -1, sl0 = kLdsNone, bj = -1, slj = kLdsNone, used = 0;
    V2<R> c0{R(0), R(0)}, ev{R(0), R(0)};
    R     d = R(0), l = R(0), rl = R(0); // rl = RN(1 / l)
};

// Screening walks the compacted candidates in groups of kScreenGroup; the list
// is padded to a whole group with sentinel sites far outside any instance.
constexpr int   kScreenGroup = 4;
constexpr float kFarSite     = 1e30f;

// Wave-uniform bounds for the screen (computed from the chunk centre q).
struct ScreenBounds {
    int   mp1;  // candidates [0, mp1) hold every site that can be a target's nearest
    int   mp2;  // candidates [mp2, mp) lie farther than r2 from q (outer band)
    int   mp;   // all candidates, padded to whole groups
    int   nb;   // sites in the two inner bands (candidates [0, nb))
    float r2;   // inner radius of the outer band (lower bound)
    float rho;  // >= |p - q| for every target of the chunk
    float qx, qy; // chunk centre q
    float mg;   // absolute rounding margin
    float sl;   // sqrt(lo) (1 - 4e-6): fewer than K staged sites lie within it of q
    // Screen table (rows = the distinct nearest sites, as a mask of their
    // slots; cols = the row stride in entries, the first mp2 of each row
    // used); rowMask == 0: no table, every candidate is screened with the
    // per-candidate formula.
    uint32_t rowMask;
    int      cols;
};

// Packed screening keys: |v| (a 32-bit float) with its low 7 bits replaced by
// the candidate's slot, so unsigned min / median-of-three order them by
// magnitude and carry the slot.  With the mask in a vector register
// (packMask()) the packing is one v_and_or_b32 per key.
constexpr uint32_t kPackMask = 0x7FFFFFFFu & ~uint32_t(kLossStage - 1);
constexpr uint32_t kPackInf  = 0x7F800000u; // packed keys below this hold finite values

__device__ __forceinline__ uint32_t packMask()
{
    uint32_t m;
    asm volatile("v_mov_b32 %0, %1" : "=v"(m) : "i"(kPackMask));
    return m;
}
__device__ __forceinline__ uint32_t packKey(uint32_t mask, float v, int slot)
{
    return (__float_as_uint(v) & mask) | uint32_t(slot);
}
// Two smallest keys: a2 = median(a1, k, a2) then a1 = min(a1, k).
__device__ __forceinline__ void keepTwo(uint32_t &a1, uint32_t &a2, uint32_t k)
{
    a2 = max(min(a1, k), min(max(a1, k), a2));
    a1 = min(a1, k);
}

// Nearest candidate of a target p among the first band of the wave's compacted
// candidates (sites within D1 + 2 rho of q, D1 = q's nearest distance: a
// target's nearest site is no farther from q), by packed keys of d2.  The
// packing truncates d2 by at most 2^-16 relative: when the two smallest keys
// differ above the slot bits, the smallest is strictly the nearest (d2 below
// the other's truncated value), which is the (d2, index) order's first;
// otherwise `ok` is false and the lane takes the exact selection.
__device__ __forceinline__ int nearestCandidate(const LossLds<float> &lds, const ScreenBounds &sb, V2<float> p,
                                                uint32_t mask, bool &ok)
{
    constexpr uint32_t kIdxMask = uint32_t(kLossStage - 1);
    uint32_t b1 = ~0u, b2 = ~0u;
#pragma unroll 1
    for (int g = 0; g < sb.mp1; g += kScreenGroup) {
#pragma unroll
        for (int k = 0; k < kScreenGroup; ++k) keepTwo(b1, b2, packKey(mask, dist2(lds.pos[g + k], p), g + k));
    }
    ok = (b1 ^ b2) > kIdxMask;
    return int(b1 & kIdxMask);
}

// FP32 screening of one lane's VoroLoss term, given its nearest candidate c0
// (slot s0, certified when nearOk), over the wave's compacted candidates (a
// superset of the lane's K nearest sites).
//   1. For every candidate c_i, an approximation a_i of |distance from p to
//      the bisector of (c0, c_i)|, packed with the slot in its low 7 bits; the
//      two smallest are kept.  c0 itself, sites on top of it and the padding
//      give NaN or infinity and never win.
//      The first mp2 candidates, when the chunk has a screen table: a_i =
//      |fma(ux, nx, fma(uy, ny, off))| from the row of c0 (u = p - c0).
//      Without a table, and for the outer band: a_i = rsq(ee) (ue - ee / 2)
//      with ee = |e|^2 and ue = <u, e> (e = c_i - c0) rounded as the reference
//      rounds them.
//      Candidates of the outer band (farther than r2 from q) are skipped when
//      their lower bound (r2 - |p - q| - |p - c0|) / 2 exceeds every active
//      lane's runner-up.
//   2. The winner w must be one of the K nearest: |p - w| + |p - q| below
//      sqrt(lo) proves it (fewer than K sites lie within sqrt(lo) of q), else
//      at most K candidates with d2 <= d2(w).
// Error bounds (eps = 2^-24).  Let delta_i = <p - m_i, n_i> be the exact signed
// distance (m_i the exact midpoint of c0 and c_i, n_i the exact unit normal).
// The reference's d_i = RN(RN(ue / l) - RN(l / 2)) has |d_i - delta_i| <=
// 11 eps |u| + 4 eps |delta_i| <= 2^-20 (|u| + |delta_i|): u and e carry
// relative errors eps, ue and ee 2 eps more each, l = RN(sqrt(ee)) is within
// 3 eps l of |e|, and |e| <= 2 (|u| + |delta_i|).
//  * Per-candidate formula: the fma forms ue - ee / 2 with one rounding from
//    the same ee and ue, and with rsq accurate to 2 ulp |a_i - |d_i|| <=
//    2^-20 (|u| + |d_i|).
//  * Table: e rsq(ee) is within 1.5 2^-22 + 2^-23 < 2^-21 of n_i, the half
//    length 0.5 ee rsq(ee) within 2^-21 L / 2 of L / 2 (L = |c_i - c0|), u
//    within eps |u|, and the two fma add eps (|u| + L / 2) + eps |a_i|; with
//    L <= 2 (|u| + |delta_i|), |a_i - delta_i| <= 2^-19.9 |u| + 2^-20.9
//    |delta_i| + eps |a_i|.
// Packing truncates a_i by at most 2^-16 relative.  So (both forms)
//   a_2 (1 - 2^-13) > a_1 (1 + 2^-13) + 2^-17 |u|
// implies |d_j| > |d_w| (1 + 2^-22) for every other candidate j: RN(d_j^2) >
// RN(d_w^2) strictly, and the reference's loop over the K nearest (smallest
// d^2, ties to the earlier key) picks w.  The exact term is then computed for w
// alone with the reference's formula.  Returns false when any condition fails;
// the caller runs the exact selection for that lane.
__device__ __forceinline__ bool screenTerm(const LossLds<float> &lds, const ScreenBounds &sb, V2<float> p, bool active,
                                           uint32_t mask, int s0, bool nearOk, int K, LaneTerm<float> &o)
{
    constexpr uint32_t kIdxMask = uint32_t(kLossStage - 1);
    static_assert((kLossStage & (kLossStage - 1)) == 0, "slot mask");
    const V2<float> c0 = lds.pos[s0];
    const V2<float> u{p.x - c0.x, p.y - c0.y};
    // uB >= |u|: |u|^2 <= RN(|u|^2) (1 + 2^-23), and v_sqrt_f32 is within
    // 1 ulp on normal inputs (smaller ones are raised to 2^-100).
    const float     uB = __builtin_amdgcn_sqrtf(fmaxf(dotR(u, u), 0x1p-100f)) * (1.0f + 0x1p-20f);
    // dq >= |p - q| for this target (at most rho): the rounded distance, raised
    // past the errors of dist2 and v_sqrt_f32, plus the margin.  For K >= 16 the
    // chunk-wide rho is used: the outer band and the count are rarely needed
    // there, and the per-target bound cost more than it saved (measured).
    float dq = sb.rho;
    if (K < 16) // uniform
        dq = fminf(__builtin_amdgcn_sqrtf(dist2(p, V2<float>{sb.qx, sb.qy})) * (1.0f + 0x1p-20f) + sb.mg, sb.rho);
    uint32_t        a1 = ~0u, a2 = ~0u;
    auto            screenGroups = [&](int g0, int g1) {
#pragma unroll 1
        for (int g = g0; g < g1; g += kScreenGroup) {
#pragma unroll
            for (int k = 0; k < kScreenGroup; ++k) {
                const V2<float> c  = lds.pos[g + k];
                const float     ex = c.x - c0.x, ey = c.y - c0.y;
                const float     ee = __builtin_fmaf(ey, ey, ex * ex);
                const float     ue = __builtin_fmaf(u.y, ey, u.x * ex);
                keepTwo(a1, a2, packKey(mask, __builtin_amdgcn_rsqf(ee) * __builtin_fmaf(-0.5f, ee, ue), g + k));
            }
        }
    };
    if (sb.rowMask) {
        // The row of c0: lanes whose target is not taken read row 0.
        const int     row = active ? __builtin_popcount(sb.rowMask & ((1u << s0) - 1u)) * sb.cols : 0;
        const float2 *tn  = lds.u.t.n + row;
        const float  *to  = lds.u.t.off + row;
#pragma unroll 1
        for (int g = 0; g < sb.mp2; g += kScreenGroup) {
#pragma unroll
            for (int k = 0; k < kScreenGroup; ++k) {
                const float2 n = tn[g + k];
                keepTwo(a1, a2, packKey(mask, __builtin_fmaf(u.x, n.x, __builtin_fmaf(u.y, n.y, to[g + k])), g + k));
            }
        }
    } else {
        screenGroups(0, sb.mp2);
    }
    // Outer band: |d_i| >= (|p - c_i| - |p - c0|) / 2 >= (r2 - |p - q| - |u|) / 2.
    // When that exceeds the runner-up (with room for the screen's error) in
    // every lane, no outer candidate can be the winner or the runner-up.
    {
        const float v2  = __uint_as_float(a2 & ~kIdxMask);
        const float lb  = 0.5f * (sb.r2 * (1.0f - 1e-5f) - dq - uB * (1.0f + 1e-5f));
        const bool  out = v2 * (1.0f + 0x1p-10f) + uB * 0x1p-14f < lb; // false when v2 is NaN (no runner-up)
        if (__ballot(active && !out)) screenGroups(sb.mp2, sb.mp);
    }
    const int   w  = int(a1 & kIdxMask);
    const float v1 = __uint_as_float(a1 & ~kIdxMask);
    const float v2 = __uint_as_float(a2 & ~kIdxMask);
    bool        ok = nearOk && a1 < kPackInf &&
              (a2 >= kPackInf ||
               __builtin_fmaf(v2, -0x1p-13f, v2) > __builtin_fmaf(v1, 0x1p-13f, v1) + uB * 0x1p-17f);
    const V2<float> cw  = lds.pos[w];
    const float     d2w = dist2(cw, p);
    // Every site within |p - w| of p lies within |p - w| + |p - q| of q, and
    // fewer than K staged sites lie within sqrt(lo) of q.
    const float     tq     = sb.sl - dq - sb.mg;
    bool            member = d2w * (1.0f + 4e-6f) < (tq > 0.0f ? tq * tq : 0.0f);
    if (__ballot(ok && !member)) {
        // w's rank in p's (d2, index) order is at most the number of candidates
        // with d2 <= d2(w).  The two inner bands hold nb sites; when nb < K only
        // the outer band is counted, and rank <= nb + (outer ones with d2 <=
        // d2(w)) (w itself is among the nb or among those).
        const bool outerOnly = sb.nb < K;
        int        cnt       = outerOnly ? sb.nb : 0;
#pragma unroll 1
        for (int g = outerOnly ? sb.nb & ~(kScreenGroup - 1) : 0; g < sb.mp; g += kScreenGroup) {
#pragma unroll
            for (int k = 0; k < kScreenGroup; ++k)
                cnt += (!outerOnly || g + k >= sb.nb) && dist2(lds.pos[g + k], p) <= d2w ? 1 : 0;
        }
        member = cnt <= K;
    }
    ok = ok && member;
    // The reference's term for (c0, w).  Below 2^-90 the length is under
    // 2^-45 < 1e-12 and the reference skips the pair.
    o.ev           = V2<float>{cw.x - c0.x, cw.y - c0.y};
    const float ee = dotR(o.ev, o.ev);
    o.l            = sqrtNormal(ee);
    ok             = ok && ee >= 0x1p-90f && !(o.l < 1e-12f);
    // ue / l and 1 / l from one refined reciprocal (1e-12 <= l <= 2 ext
    // wherever the term is used); a numerator below 2^-70 other than 0 takes
    // the full division.
    const float ue = dotR(u, o.ev);
    const float r1 = recipRefined(o.l);
    if (__ballot(ok && ue != 0.0f && fabsf(ue) < 0x1p-70f)) o.d = fmaR(-0.5f, o.l, ue / o.l);
    else o.d = fmaR(-0.5f, o.l, divCore(ue, o.l, r1));
    o.rl  = divCore(1.0f, o.l, r1);
    o.i0  = lds.idx[s0], o.sl0 = s0, o.c0 = c0;
    o.bj  = lds.idx[w], o.slj = w, o.used = 1;
    return ok;
}

// Screen table rows for the distinct nearest sites of a staged chunk (slots in
// rowMask, all below 32) over its first `cols` candidates, rows `stride`
// entries apart; candidates from slot m on are padding and get NaN entries.
// Every lane of the wave takes part.
__device__ __forceinline__ void buildScreenTable(LossLds<float> &lds, uint32_t rowMask, int cols, int stride, int m,
                                                 int lane)
{
    const float nan = __builtin_nanf("");
    int         row = 0;
    for (uint32_t rm = rowMask; rm; rm &= rm - 1u, row += stride) {
        const V2<float> c0 = lds.pos[__builtin_ctz(rm)];
        for (int i = lane; i < cols; i += 64) {
            float nx = nan, ny = nan, off = nan;
            if (i < m) {
                const V2<float> c  = lds.pos[i];
                const float     ex = c.x - c0.x, ey = c.y - c0.y;
                const float     ee = __builtin_fmaf(ey, ey, ex * ex);
                const float     rs = __builtin_amdgcn_rsqf(ee); // c_i == c0: infinity, and the entry is NaN
                nx = ex * rs, ny = ey * rs;
                off = -0.5f * (ee * rs);
            }
            lds.u.t.n[row + i]   = float2{nx, ny};
            lds.u.t.off[row + i] = off;
        }
    }
}

// Gradient of the term d^2 for the pair (c0, cj) (voro_optimizer.cu:244-254).
// With u = p - c0, e = cj - c0, l = |e|, n = e / l:
//   dd/dcj = (u - <u, n> n) / l - n / 2,   dd/dc0 = -n - dd/dcj,
// scaled by 2d and added to the accumulators of cj (dj) and c0 (d0).
// rl is the correctly rounded 1 / l.
template <class R>
__device__ __forceinline__ int lossGrad(V2<R> u, V2<R> ev, R rl, R d, GradT<R> *dj, GradT<R> *d0, double gscale)
{
    const V2<R> n{rl * ev.x, rl * ev.y};
    const R     un = dotR(u, n);
    const R     tx = fmaR(-un, n.x, u.x), ty = fmaR(-un, n.y, u.y);
    const V2<R> dcj{fmaR(rl, tx, R(-0.5) * n.x), fmaR(rl, ty, R(-0.5) * n.y)};
    const R     s = R(2) * d;
    const R     gx = s * dcj.x, gy = s * dcj.y, hx = -(s * (n.x + dcj.x)), hy = -(s * (n.y + dcj.y));
    GradOps<R>::add(dj, gx, gscale);
    GradOps<R>::add(dj + 1, gy, gscale);
    GradOps<R>::add(d0, hx, gscale);
    GradOps<R>::add(d0 + 1, hy, gscale);
    // 1 when a component was clamped by the FP32 fixed-point scaling (the
    // instance's `saturated` count reports it); FP64 sums are not clamped.
    if constexpr (sizeof(R) == 4)
        return int(fmax(fmax(fabs(gx), fabs(gy)), fmax(fabs(hx), fabs(hy))) > quantLimit(gscale));
    return 0;
}

template <class R> __device__ __forceinline__ R relMargin();
template <> __device__ __forceinline__ float  relMargin<float>() { return 4e-6f; }
template <> __device__ __forceinline__ double relMargin<double>() { return 1e-12; }

template <class R> __device__ __forceinline__ R rInf();
template <> __device__ __forceinline__ float  rInf<float>() { return __builtin_huge_valf(); }
template <> __device__ __forceinline__ double rInf<double>() { return __builtin_huge_val(); }

// Orders LDS traffic of one wavefront (its lanes execute in lockstep).
__device__ __forceinline__ void waveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// DPP controls (gfx9 encoding).
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;
constexpr int kDppBcast15 = 0x142, kDppBcast31 = 0x143;

template <int Ctrl, int RowMask> __device__ __forceinline__ uint32_t dpp32(uint32_t identity, uint32_t v)
{
    return uint32_t(__builtin_amdgcn_update_dpp(int(identity), int(v), Ctrl, RowMask, 0xF, false));
}

// Wave-wide reduction of a 32-bit value with a commutative op; every lane of
// the wave must be active.  Result valid in all lanes (read from lane 63).
template <class T, class Op> __device__ __forceinline__ T waveReduce32(T v, T identity, Op op)
{
    static_assert(sizeof(T) == 4, "32-bit values only");
    auto bits = [](T x) { return __builtin_bit_cast(uint32_t, x); };
    auto val  = [](uint32_t b) { return __builtin_bit_cast(T, b); };
    const uint32_t id = bits(identity);
    v = op(v, val(dpp32<kDppXor1, 0xF>(id, bits(v))));
    v = op(v, val(dpp32<kDppXor2, 0xF>(id, bits(v))));
    v = op(v, val(dpp32<kDppHalfMirror, 0xF>(id, bits(v))));
    v = op(v, val(dpp32<kDppMirror, 0xF>(id, bits(v))));
    v = op(v, val(dpp32<kDppBcast15, 0xA>(id, bits(v))));
    v = op(v, val(dpp32<kDppBcast31, 0xC>(id, bits(v))));
    return val(uint32_t(__builtin_amdgcn_readlane(int(bits(v)), 63)));
}

__device__ __forceinline__ double dppDouble(double identity, double v, int which)
{
    uint64_t b = __builtin_bit_cast(uint64_t, v), i = __builtin_bit_cast(uint64_t, identity);
    uint32_t lo = uint32_t(b), hi = uint32_t(b >> 32), ilo = uint32_t(i), ihi = uint32_t(i >> 32);
    switch (which) {
    case 0: lo = dpp32<kDppXor1, 0xF>(ilo, lo), hi = dpp32<kDppXor1, 0xF>(ihi, hi); break;
    case 1: lo = dpp32<kDppXor2, 0xF>(ilo, lo), hi = dpp32<kDppXor2, 0xF>(ihi, hi); break;
    case 2: lo = dpp32<kDppHalfMirror, 0xF>(ilo, lo), hi = dpp32<kDppHalfMirror, 0xF>(ihi, hi); break;
    case 3: lo = dpp32<kDppMirror, 0xF>(ilo, lo), hi = dpp32<kDppMirror, 0xF>(ihi, hi); break;
    case 4: lo = dpp32<kDppBcast15, 0xA>(ilo, lo), hi = dpp32<kDppBcast15, 0xA>(ihi, hi); break;
    default: lo = dpp32<kDppBcast31, 0xC>(ilo, lo), hi = dpp32<kDppBcast31, 0xC>(ihi, hi); break;
    }
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

__device__ __forceinline__ double waveSumDouble(double v)
{
#pragma unroll
    for (int k = 0; k < 6; ++k) v = v + dppDouble(0.0, v, k);
    uint64_t b = __builtin_bit_cast(uint64_t, v);
    uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b)), 63));
    uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b >> 32)), 63));
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

template <class R> __device__ __forceinline__ R waveMaxR(R v);
template <> __device__ __forceinline__ float waveMaxR<float>(float v)
{
    return waveReduce32<float>(v, -__builtin_huge_valf(), [](float a, float b) { return fmaxf(a, b); });
}
template <> __device__ __forceinline__ double waveMaxR<double>(double v)
{
#pragma unroll
    for (int k = 0; k < 6; ++k) v = fmax(v, dppDouble(-__builtin_huge_val(), v, k));
    uint64_t b = __builtin_bit_cast(uint64_t, v);
    uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b)), 63));
    uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b >> 32)), 63));
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

template <class R> __device__ __forceinline__ R waveMinR(R v)
{
    return -waveMaxR<R>(-v);
}

// The LDS gradient of a wavefront holds `copies` (1, 2 or 4) interleaved
// copies of the accumulators of its staged sites: slot s * copies + c, 16
// bytes (x, y) each.  Lane l adds into copy l % copies, so lanes whose terms
// touch the same site collide less, and the copies of one site lie in
// different LDS banks.  The flush sums them.  A lane's accumulator of site s is
// one shift-add from its copy's base: gradAt(base, s, shift), shift =
// log2(copies) + 4.
template <class T> __device__ __forceinline__ T *gradAt(T *laneBase, int s, int shift)
{
    return reinterpret_cast<T *>(reinterpret_cast<char *>(laneBase) + (s << shift));
}

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ int laneRank(uint64_t mask)
{
    return int(__builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u)));
}

// The exact selection for one target pt of a staged chunk, by the whole
// wavefront, when the screen left it open.  The chunk's m candidates hold pt's
// K nearest sites (the candidate disc), so they are ranked by the (d2, index)
// keys of the exact kernel and lane k evaluates the reference's term for
// neighbour k; the smallest (d^2, k) wins (voro_optimizer.cu:225-242), exactly
// as in the exact kernel.  The winner's gradient goes to the chunk's LDS
// accumulators.  Returns the winning lane (-1: every pair degenerate); that
// lane's outputs are the term's loss, the nearest site's and the partner's
// instance indices and whether a gradient component was clamped.
__device__ __forceinline__ int exactInChunk(LossLds<float> &lds, int m, V2<float> pt, int lane, int K, long long *gLane,
                                            int gShift, bool grad, double gscale, float &lossT, int &i0, int &j, int &sat)
{
    for (int s = lane; s < m; s += 64)
        lds.u.x.xkey[s] = Key<float>::make(dist2(lds.pos[s], pt), (uint32_t(lds.idx[s]) << kLdsSlotBits) | uint32_t(s)).k;
    waveSync();
    for (int s = lane; s < m; s += 64) {
        const uint64_t ks   = lds.u.x.xkey[s];
        int            rank = 0;
#pragma unroll 4
        for (int t = 0; t < m; ++t) rank += lds.u.x.xkey[t] < ks ? 1 : 0;
        if (rank < K) lds.u.x.xtop[rank] = s;
    }
    waveSync();
    const int       s0 = lds.u.x.xtop[0];
    const V2<float> c0 = lds.pos[s0];
    const V2<float> u{pt.x - c0.x, pt.y - c0.y};
    float           dd = rInf<float>(), d = 0.0f, l = 0.0f;
    V2<float>       ev{0.0f, 0.0f};
    int             sj = 0;
    bool            ok = false;
    if (lane >= 1 && lane < K) {
        sj                 = lds.u.x.xtop[lane];
        const V2<float> cj = lds.pos[sj];
        ev                 = V2<float>{cj.x - c0.x, cj.y - c0.y};
        l                  = sqrtR(dotR(ev, ev));
        if (!(l < 1e-12f)) {
            d  = fmaR(-0.5f, l, dotR(u, ev) / l);
            dd = d * d;
            ok = true;
        }
    }
    const float    best = waveMinR<float>(dd);
    const uint64_t win  = __ballot(ok && dd == best);
    const int      wl   = win ? __builtin_ctzll(win) : -1;
    if (lane == wl) {
        lossT = best, i0 = lds.idx[s0], j = lds.idx[sj];
        sat   = grad ? lossGrad(u, ev, 1.0f / l, d, gradAt(gLane, sj, gShift), gradAt(gLane, s0, gShift), gscale) : 0;
    }
    waveSync();
    return wl;
}

// At most 64 VGPRs, so 8 wavefronts share a SIMD: the kernel is bound by issue
// and by the latency of its per-chunk loads, and occupancy hides the latter.
// k comes from each work item's instance record, so one launch serves fits of
// every k.
template <int TPL, class R, bool Diag>
__global__ __launch_bounds__(kLossThreads) __attribute__((amdgpu_waves_per_eu(8))) void lossKernel(LossArgs<R> a)
{
    // Diagnostics (path counters, k-NN lists, per-target terms, exact-only
    // runs) exist only in the Diag instantiation.
    unsigned long long *const stats     = Diag ? a.stats : nullptr;
    int *const                dbgKnn    = Diag ? a.dbgKnn : nullptr;
    TermDbg *const            dbgTerm   = Diag ? a.dbgTerm : nullptr;
    const bool                exactOnly = Diag && a.exactOnly;
    // The plain instantiation runs optimizer iterations with a gradient only
    // (the launcher sends evaluation passes to the other one), so these are
    // constants there.
    const bool iterate  = !Diag || a.mode == kModeIterate;
    const bool wantGrad = !Diag || a.grad != nullptr;
    __shared__ LossLds<R> ldsAll[kLossWaves];
    const int  lane  = threadIdx.x & 63;
    // The wavefront's index is uniform: read it into a scalar register so that
    // everything derived from the work item (chunk, instance, cell block,
    // bounds) is scalar arithmetic.
    const int  wave  = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    LossLds<R> &lds  = ldsAll[wave];
    const int  nWork = *a.nWork;
    const uint32_t mask = sizeof(R) == 4 ? packMask() : 0u;
    const int  xcdWave = (int(blockIdx.x) / kXcds) * kLossWaves + wave; // the wavefront's rank on its XCD
    const int  first   = gridDim.x % (2 * kXcdRun) == 0
                             ? ((xcdWave / kXcdRun) * kXcds + int(blockIdx.x) % kXcds) * kXcdRun + xcdWave % kXcdRun
                             : int(blockIdx.x) * kLossWaves + wave;
    for (int w = first; w < nWork; w += gridDim.x * kLossWaves) {
        const TargetWork wk = loadUniform(a.work + w);
        const int64_t    so = wk.siteOff;

        // Every load that needs only the work item goes out at once: the
        // chunk's targets (lane + 64 h, h < TPL; their slot numbers
        // when subsampled) and the cell starts of the staged block.  The
        // instance record and state arrive meanwhile.
        V2<R>    p[TPL];
        uint32_t tsl[TPL];
        bool     take[TPL];
        const bool sub = (wk.flags & kWorkSubsample) != 0;
#pragma unroll
        for (int h = 0; h < TPL; ++h) {
            const int64_t ti = wk.tBegin + (lane + 64 * h < wk.count ? lane + 64 * h : 0);
            p[h]   = a.tpos[ti];
            tsl[h] = sub ? uint32_t(a.tslot[ti]) : 0u;
        }
        // The block of rings 0..rs around the chunk's cell, [sx0,sx1] x [sy0,sy1]:
        // lane r*snx + i reads the range of cell (sx0 + i, sy0 + r); lane / snx is
        // a multiply and shift (exact for lane < 7 snx, snx <= 7).
        const int *strt = a.start + wk.cellOff;
        int        rs   = wk.flags & kWorkRingsMask;
        int        sx0 = 0, sx1 = 0, sy0 = 0, sy1 = 0, snx = 1, sny = 1, cb = 0, ce = 0;
        auto       blockStarts = [&]() {
            sx0 = max(wk.cx - rs, 0), sx1 = min(wk.cx + rs, wk.gw - 1);
            sy0 = max(wk.cy - rs, 0), sy1 = min(wk.cy + rs, wk.gh - 1);
            snx = sx1 - sx0 + 1, sny = sy1 - sy0 + 1;
            const int rowMul = snx == 1 ? 256 : snx == 2 ? 128 : snx == 3 ? 86 : snx == 4 ? 64 : snx == 5 ? 52
                               : snx == 6 ? 43 : 37; // ceil(256 / snx)
            cb = 0, ce = 0;
            if (lane < snx * sny) {
                const int r = (lane * rowMul) >> 8;
                const int c = (sy0 + r) * wk.gw + sx0 + (lane - r * snx);
                cb          = strt[c];
                ce          = strt[c + 1];
            }
        };
        if constexpr (sizeof(R) == 4) blockStarts();

        const Inst in = loadUniform(a.insts + wk.inst);
        const int  K  = in.K; // neighbours of the instance's terms
        InstState *st = a.state + wk.inst;
        if (iterate && loadUniform(&st->done)) continue;
        bool anyTake = false;
#pragma unroll
        for (int h = 0; h < TPL; ++h) {
            const bool valid = lane + 64 * h < wk.count;
            if (!valid) p[h] = V2<R>{R(-1), R(-1)};
            take[h] = valid && p[h].x >= R(0);
            if (take[h] && in.useBox && iterate)
                take[h] = p[h].x >= R(in.bx0) && p[h].x <= R(in.bx1) && p[h].y >= R(in.by0) && p[h].y <= R(in.by1);
            if (take[h] && iterate && in.fraction < 1.0f)
                take[h] = rand01_loss(tsl[h], uint32_t(loadUniform(&st->iter)), in.seed) < in.fraction;
            anyTake = anyTake || take[h];
        }
        if (__ballot(anyTake) == 0) {
            if (stats && lane == 0) atomicAdd(&stats[8 * (blockIdx.x & 63) + 3], 1ull);
            continue;
        }

        bool         fast  = false;
        int          m     = 0; // compacted candidates (fast path)
        int          gcopies = 1, gShift = 4; // gradient copies, log2(copies) + 4
        int          total = 0; // sites of the staged block
        ScreenBounds sb{};
        if constexpr (sizeof(R) == 4) {
            const int   gw = wk.gw, gh = wk.gh;
            const R     cell = R(in.cell);
            const R     ext  = R(in.ext);
            const V2<R> q{R(wk.qx), R(wk.qy)};
            const R     mg  = relMargin<R>() * fmax(ext, fmax(fabs(q.x), fabs(q.y)) + R(wk.rho));
            const R     rho = R(wk.rho) + mg;
            const int   e0 = lane, e1 = lane + 64;
            R           d0 = rInf<R>(), d1 = rInf<R>(), lo = R(0), hi = R(0), thr = R(0), thr2 = R(0);
            // One retry: a block that misses the candidate disc is staged again
            // with kStageRetry rings, a block of kStageRetry rings that holds
            // too many sites with two.
            bool retried = false;
            for (;;) {
                // ---- stage the block into LDS ----------------------------------
                // Rows of the block are contiguous ranges of the sorted arrays:
                // row r covers entries [rb[r], rb[r] + len) and LDS slots
                // [roff[r], roff[r+1]).
                constexpr int kRows = 2 * kStageRetry + 1;
                int           rb[kRows], roff[kRows + 1];
                roff[0] = 0;
#pragma unroll
                for (int r = 0; r < kRows; ++r) {
                    const int rr = min(r, sny - 1);
                    rb[r]        = __builtin_amdgcn_readlane(cb, rr * snx);
                    roff[r + 1]  = roff[r] + (r < sny ? __builtin_amdgcn_readlane(ce, rr * snx + snx - 1) - rb[r] : 0);
                }
                total = roff[kRows];
                fast  = total <= kLossStage && total >= K;
                if (fast) {
                    // Slot d of the block (row-major over its cells) holds sorted
                    // entry d + rb[r] - roff[r] of its row r (the last row with
                    // roff[r] <= d).  Each lane takes slots lane and lane + 64 and
                    // issues their four loads together; a lane past the end reads
                    // slot 0's entry (total >= K >= 2) and keeps nothing.
                    int s0 = lane + rb[0], s1 = lane + 64 + rb[0];
#pragma unroll
                    for (int r = 1; r < kRows; ++r) {
                        const int step = (rb[r] - roff[r]) - (rb[r - 1] - roff[r - 1]);
                        s0 += lane >= roff[r] ? step : 0;
                        s1 += lane + 64 >= roff[r] ? step : 0;
                    }
                    const int     first = __builtin_amdgcn_readfirstlane(s0);
                    const bool    in0 = e0 < total, in1 = e1 < total;
                    const int64_t g0 = so + (in0 ? s0 : first), g1 = so + (in1 ? s1 : first);
                    const V2<R>   q0 = a.spos[g0], q1 = a.spos[g1];
                    const int     x0 = a.sidx[g0], x1 = a.sidx[g1];
                    if (in0) lds.pos[e0] = q0, lds.idx[e0] = x0;
                    if (in1) lds.pos[e1] = q1, lds.idx[e1] = x1;
                    waveSync();
                    // Squared distance from q to each staged site (two per lane).
                    d0 = in0 ? dist2(q0, q) : rInf<R>();
                    d1 = in1 ? dist2(q1, q) : rInf<R>();
                    // Bisection: lo < D^2 <= hi for D^2 = q's K-th smallest, i.e.
                    // count(d2 <= lo) < K <= count(d2 <= hi) over the staged sites,
                    // on the squared distances' bit patterns (uniform: scalar
                    // arithmetic, two vector compares per step).
                    {
                        const uint32_t b0 = __builtin_bit_cast(uint32_t, float(d0));
                        const uint32_t b1 = __builtin_bit_cast(uint32_t, float(d1));
                        auto count = [&](uint32_t x) {
                            return __popcll(__ballot(b0 <= x)) + __popcll(__ballot(b1 <= x));
                        };
                        uint32_t hiB = waveReduce32<uint32_t>(max(in0 ? b0 : 0u, in1 ? b1 : 0u), 0u,
                                                              [](uint32_t x, uint32_t y) { return max(x, y); });
                        uint32_t loB = hiB > kBisectSpan ? hiB - kBisectSpan : 0u;
                        if (count(loB) >= K) loB = 0u;
                        for (int it = 0; it < kBisectSteps; ++it) {
                            const uint32_t mid = loB + ((hiB - loB) >> 1);
                            if (count(mid) >= K) hiB = mid;
                            else loB = mid;
                        }
                        lo = R(__builtin_bit_cast(float, loB));
                        hi = R(__builtin_bit_cast(float, hiB));
                    }
                    const R D = sqrtR(hi) + mg;
                    thr       = D + R(2) * rho + mg;
                    thr2      = thr * thr * (R(1) + relMargin<R>());
                    // The block must contain the disc of radius thr around q: then
                    // it holds the K nearest sites of every target of the chunk.
                    R edge = rInf<R>();
                    if (sx0 > 0) edge = fmin(edge, q.x - R(sx0) * cell);
                    if (sx1 < gw - 1) edge = fmin(edge, R(sx1 + 1) * cell - q.x);
                    if (sy0 > 0) edge = fmin(edge, q.y - R(sy0) * cell);
                    if (sy1 < gh - 1) edge = fmin(edge, R(sy1 + 1) * cell - q.y);
                    fast = edge - R(2) * mg > thr;
                }
                if (fast || retried) break;
                if (total > kLossStage) {
                    if (rs < kStageRetry) break;
                    rs = 2;
                } else {
                    if (rs >= kStageRetry) break;
                    rs = kStageRetry;
                }
                retried = true;
                waveSync();
                blockStarts();
            }
            if (fast) {
                // Compact the candidates (d2 <= thr^2) into LDS, the first band
                // (d2 <= r1^2, r1 = q's nearest distance + 2 rho: it holds every
                // target's nearest site) first.
                const R        r1   = sqrtR(waveMinR<R>(fmin(d0, d1))) + R(2) * rho + R(2) * mg;
                const R        r1sq = fmin(r1 * r1 * (R(1) + relMargin<R>()), thr2);
                const R        r2sq = fmin(fmax(lo, r1sq), thr2); // outer band: r2sq < d2 <= thr2
                const uint64_t a0 = __ballot(d0 <= r1sq), a1 = __ballot(d1 <= r1sq);
                const uint64_t b0 = __ballot(d0 <= r2sq && !(d0 <= r1sq)), b1 = __ballot(d1 <= r2sq && !(d1 <= r1sq));
                const uint64_t c0b = __ballot(d0 <= thr2 && !(d0 <= r2sq)), c1b = __ballot(d1 <= thr2 && !(d1 <= r2sq));
                const int      nA = __popcll(a0) + __popcll(a1);
                const int      nB = nA + __popcll(b0) + __popcll(b1);
                m                 = nB + __popcll(c0b) + __popcll(c1b);
                int dst0 = -1, dst1 = -1;
                if (d0 <= r1sq) dst0 = laneRank(a0);
                else if (d0 <= r2sq) dst0 = nA + laneRank(b0);
                else if (d0 <= thr2) dst0 = nB + laneRank(c0b);
                if (d1 <= r1sq) dst1 = __popcll(a0) + laneRank(a1);
                else if (d1 <= r2sq) dst1 = nA + __popcll(b0) + laneRank(b1);
                else if (d1 <= thr2) dst1 = nB + __popcll(c0b) + laneRank(c1b);
                V2<R> p0{}, p1{};
                int   i0v = 0, i1v = 0;
                if (dst0 >= 0) p0 = lds.pos[e0], i0v = lds.idx[e0];
                if (dst1 >= 0) p1 = lds.pos[e1], i1v = lds.idx[e1];
                waveSync();
                if (dst0 >= 0) lds.pos[dst0] = p0, lds.idx[dst0] = i0v;
                if (dst1 >= 0) lds.pos[dst1] = p1, lds.idx[dst1] = i1v;
                // Pad to whole screening groups with far sentinels.
                const int mp = (m + kScreenGroup - 1) & ~(kScreenGroup - 1);
                if (lane < mp - m) lds.pos[m + lane] = V2<R>{R(kFarSite), R(kFarSite)};
                // Clear the gradient copies.
                gcopies = m <= kLossStage / 4 ? 4 : (m <= kLossStage / 2 ? 2 : 1);
                gShift  = m <= kLossStage / 4 ? 6 : (m <= kLossStage / 2 ? 5 : 4);
                if (wantGrad) {
#pragma unroll
                    for (int i = lane; i < 2 * kLossStage; i += 64) lds.grad[i] = GradT<R>(0);
                }
                waveSync();
                // Screen bounds (see screenTerm).  Fewer than K staged sites
                // lie within sqrt(lo) of q.
                sb.mp1     = __builtin_amdgcn_readfirstlane((nA + kScreenGroup - 1) & ~(kScreenGroup - 1));
                sb.mp2     = __builtin_amdgcn_readfirstlane((nB + kScreenGroup - 1) & ~(kScreenGroup - 1));
                sb.mp      = __builtin_amdgcn_readfirstlane(mp);
                sb.nb      = __builtin_amdgcn_readfirstlane(nB);
                sb.r2      = sqrtR(r2sq);
                sb.rho     = rho;
                sb.qx = q.x, sb.qy = q.y, sb.mg = mg, sb.sl = sqrtR(lo) * R(1 - 4e-6);
            }
        }

        // Open targets of a staged chunk take the exact selection inside the
        // chunk; the k-NN list diagnostics and exact-only runs queue them.
        const bool inl = fast && !exactOnly && !dbgKnn;
        const bool scr = fast && !exactOnly; // screened terms (FP32 only)
        // ---- nearest candidate of every target, then the screen table over
        //      the chunk's distinct nearest sites (when their slots are below
        //      32 and the rows fit) ---------------------------------------------
        uint32_t nearPack = 0; // pass h: bits 8h..8h+6 the nearest slot, bit 8h+7 set when certified
        if constexpr (sizeof(R) == 4) {
            if (scr) {
                uint32_t rows = 0;
                bool     high = false;
#pragma unroll
                for (int h = 0; h < TPL; ++h) {
                    if (__ballot(take[h]) == 0) continue;
                    bool      nok;
                    const int s0 = nearestCandidate(lds, sb, p[h], mask, nok);
                    nearPack |= uint32_t(s0 | (nok ? 0x80 : 0)) << (8 * h);
                    if (take[h]) {
                        if (s0 < 32) rows |= 1u << s0;
                        else high = true;
                    }
                }
                const uint32_t rowMask = waveReduce32<uint32_t>(rows, 0u, [](uint32_t x, uint32_t y) { return x | y; });
                const int      nrows   = __builtin_popcount(rowMask);
                if (__ballot(high) == 0 && nrows * (sb.mp2 + kTblPad) <= kTblEntries) {
                    buildScreenTable(lds, rowMask, sb.mp2, sb.mp2 + kTblPad, m, lane);
                    sb.rowMask = rowMask, sb.cols = sb.mp2 + kTblPad;
                    waveSync();
                }
            }
        }
        // ---- each target: screened term, or queued for the exact selection
        //      (for the k-NN list diagnostics every taken target is queued; bit
        //      31 of the instance marks a term already settled here) ------------
        double   lossSum = 0.0;
        int      used    = 0;
        int      openCnt = 0;
        int      sat     = 0; // targets whose gradient terms were clamped
        uint64_t openM[TPL]; // open targets of each pass (exact selection inside the chunk)
        GradT<R> *const gLane = lds.grad + 2 * (lane & (gcopies - 1)); // this lane's gradient copy
#pragma unroll
        for (int h = 0; h < TPL; ++h) {
            openM[h] = 0;
            if (__ballot(take[h] || (dbgKnn && lane + 64 * h < wk.count)) == 0) continue;
            const int64_t t  = wk.tBegin + lane + 64 * h;
            LaneTerm<R> tr;
            bool        open = take[h];
            if constexpr (sizeof(R) == 4) {
                if (scr) {
                    const uint32_t np = nearPack >> (8 * h);
                    if (screenTerm(lds, sb, p[h], take[h], mask, int(np & 0x7Fu), (np & 0x80u) != 0, K, tr)) open = false;
                    tr.used = take[h] && !open ? tr.used : 0;
                }
            }
            const bool     push = (open && !inl) || (dbgKnn && take[h]);
            const uint64_t pm   = __ballot(push);
            if (pm) {
                int base = 0;
                if (lane == 0) base = atomicAdd(a.exactLen, __popcll(pm));
                base = __builtin_amdgcn_readfirstlane(base);
                if (push)
                    a.exactQ[base + laneRank(pm)] =
                        ExactItem{open ? wk.inst : int(uint32_t(wk.inst) | 0x80000000u), int(t)};
            }
            openCnt += __popcll(__ballot(open));
            const R lossT = tr.used ? tr.d * tr.d : R(0);
            if (tr.used && wantGrad)
                sat += lossGrad(V2<R>{p[h].x - tr.c0.x, p[h].y - tr.c0.y}, tr.ev, tr.rl, tr.d,
                         gradAt(gLane, tr.slj, gShift), gradAt(gLane, tr.sl0, gShift), __builtin_ldexp(1.0, in.gradShift));
            if (dbgKnn && wk.inst == a.dbgInst && lane + 64 * h < wk.count && !take[h]) {
                for (int k = 0; k < K; ++k) dbgKnn[(t - in.targetOff) * K + k] = -1;
            }
            if (dbgTerm && lane + 64 * h < wk.count && !open) {
                const int64_t o   = t;
                dbgTerm[o].i0   = tr.used ? tr.i0 : -1;
                dbgTerm[o].j    = tr.used ? tr.bj : -1;
                dbgTerm[o].loss = double(lossT);
            }
            lossSum += double(lossT);
            used += tr.used;
            if constexpr (sizeof(R) == 4) openM[h] = inl ? __ballot(open) : 0ull;
        }
        // ---- targets the screen left open: the exact selection inside the
        //      chunk (its keys take the table's space) --------------------------
        if constexpr (sizeof(R) == 4) {
#pragma unroll
            for (int h = 0; h < TPL; ++h) {
                uint64_t om = openM[h];
                while (om) {
                    const int l = __builtin_ctzll(om);
                    om &= om - 1;
                    const V2<float> pt{__int_as_float(__builtin_amdgcn_readlane(__float_as_int(p[h].x), l)),
                                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(p[h].y), l))};
                    float     xl = 0.0f;
                    int       xi0 = -1, xj = -1, xs = 0;
                    const int wl = exactInChunk(lds, m, pt, lane, K, gLane, gShift, wantGrad,
                                                   __builtin_ldexp(1.0, in.gradShift), xl, xi0, xj, xs);
                    if (lane == wl) lossSum += double(xl), used += 1, sat += xs;
                    if (dbgTerm && lane == (wl >= 0 ? wl : 0)) {
                        const int64_t o = wk.tBegin + l + 64 * h;
                        dbgTerm[o].i0   = wl >= 0 ? xi0 : -1;
                        dbgTerm[o].j    = wl >= 0 ? xj : -1;
                        dbgTerm[o].loss = wl >= 0 ? double(xl) : 0.0;
                    }
                }
            }
        }
        if (stats && lane == 0) {
            unsigned long long *sp = stats + 8 * (blockIdx.x & 63); // 64 lines, summed on the host
            atomicAdd(&sp[fast ? 0 : (total <= kLossStage ? 1 : 2)], 1ull);
            atomicAdd(&sp[4], (unsigned long long)m);
            atomicAdd(&sp[5], (unsigned long long)total);
            if (fast && openCnt) {
                atomicAdd(&sp[6], 1ull);
                atomicAdd(&sp[7], (unsigned long long)openCnt);
            }
        }

        if (__ballot(sat != 0)) { // clamped gradient terms (rare): count them
            const int ns = waveReduce32<int>(sat, 0, [](int x, int y) { return x + y; });
            if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long *>(&st->saturated), (unsigned long long)ns);
        }
        // ---- FP64 loss reduction: one atomic pair per wavefront ----------
        const double ls = waveSumDouble(lossSum);
        const int    uc = waveReduce32<int>(used, 0, [](int x, int y) { return x + y; });
        if (lane == 0 && uc > 0) {
            if (iterate) {
                LossPartial *pp = a.partial + int64_t(wk.inst) * kLossPartials + (blockIdx.x & (kLossPartials - 1));
                unsafeAtomicAdd(&pp->loss, ls);
                atomicAdd(&pp->count, (unsigned long long)uc);
            } else {
                unsafeAtomicAdd(&st->evalLoss, ls);
                atomicAdd((unsigned long long *)&st->evalCount, (unsigned long long)uc);
            }
        }

        // ---- flush the LDS gradient: one atomic per touched site ----------
        if (fast && wantGrad) {
            waveSync();
            for (int s = lane; s < m; s += 64) {
                GradT<R> gx = GradT<R>(0), gy = GradT<R>(0);
                const GradT<R> *gs = gradAt(lds.grad, s, gShift);
                for (int c = 0; c < gcopies; ++c) gx += gs[2 * c], gy += gs[2 * c + 1];
                if (gx != GradT<R>(0) || gy != GradT<R>(0)) {
                    const int i = lds.idx[s];
                    GradOps<R>::addRaw(&a.grad[2 * (so + i)], gx);
                    GradOps<R>::addRaw(&a.grad[2 * (so + i) + 1], gy);
                }
            }
        }
        waveSync();
    }
}

// ---------------------------------------------------------------------------
// 3b. Exact selection for the targets the loss kernel queued, one wavefront
//     per target.  The wavefront gathers the sites of successive rings of cells
//     around the target's cell into LDS and keeps the K smallest (d2, index)
//     keys by rank counting; it stops with the same conservative rule as the
//     per-thread ring search (knnGlobal), so the list is the exact one.  Lane k
//     then evaluates the reference's term for neighbour k and the smallest
//     (d^2, k) wins (voro_optimizer.cu:225-242).  Gradients go to global memory
//     as fixed-point atomics, so the sums equal the staged ones.  The scan
//     kernel empties the queue before each step.
// ---------------------------------------------------------------------------
constexpr int kExactWaves      = 4;
constexpr int kExactCand       = 256; // LDS keys per wavefront: the current K nearest + a batch of new sites
constexpr int kExactFirstRings = 2;   // the first round gathers rings 0..2 (25 cells <= 64 lanes)

// P: each candidate's position (from the sorted copy) is kept beside its key, so
// the term reads no position from HBM.  The clip and the FP64 exact selection,
// whose occupancy is LDS-bound, go without.
template <class R, bool P> struct ExactLds {
    Key<R> cand[kExactCand];
    Key<R> top[32];
    V2<R>  pos[P ? kExactCand : 1];
    V2<R>  topPos[P ? 32 : 1];
    int    cb[64];   // a round of ring cells: first sorted entry of each
    int    co[65];   // and the prefix of their site counts
};

// Sorts the K smallest of the distinct keys cand[0, nc) into cand[0, min(K, nc)).
template <class R, bool P> __device__ __forceinline__ int keepNearest(ExactLds<R, P> &lds, int nc, int lane, int K)
{
    for (int j = lane; j < nc; j += 64) {
        const Key<R> kj   = lds.cand[j];
        int          rank = 0;
#pragma unroll 8
        for (int i = 0; i < nc; ++i) rank += lds.cand[i].lt(kj) ? 1 : 0;
        if (rank < K) {
            lds.top[rank] = kj;
            if constexpr (P) lds.topPos[rank] = lds.pos[j];
        }
    }
    waveSync();
    const int nk = min(nc, K);
    if (lane < nk) {
        lds.cand[lane] = lds.top[lane];
        if constexpr (P) lds.pos[lane] = lds.topPos[lane];
    }
    waveSync();
    return nk;
}

// Cell q (0 <= q < 8r) of the ring at Chebyshev distance r >= 1 around (cx, cy).
__device__ __forceinline__ void ringCell(int r, int q, int cx, int cy, int &x, int &y)
{
    const int side = 2 * r + 1;
    if (q < side) x = cx - r + q, y = cy - r;
    else if (q < 2 * side) x = cx - r + (q - side), y = cy + r;
    else if (q < 2 * side + side - 2) x = cx - r, y = cy - r + 1 + (q - 2 * side);
    else x = cx + r, y = cy - r + 1 + (q - 3 * side + 2);
}

// The K nearest sites of p under the (d2, index) order, d2 in the key type KR,
// gathered by one wavefront into lds.cand[0, nc) (sorted; nc = min(K, sites)):
// the 5 x 5 block (rings 0-2) in one round, then ring by ring in rounds of 64
// cells; a ring ends the search when the K-th distance is below the distance to
// the searched block's edge (less a rounding margin).
template <class R, class KR, bool P>
__device__ __forceinline__ int gatherKnn(ExactLds<KR, P> &lds, const Inst &in, const V2<R> *spos, const int *sidx,
                                         const int *strt, V2<R> pr, int lane, int K)
{
    const int    gw = in.gw, gh = in.gh;
    const R      inv    = R(1) / R(in.cell);
    const KR     cell   = KR(R(in.cell));
    const KR     rel    = KR(ringRel<R>());
    const V2<KR> p{KR(pr.x), KR(pr.y)};
    const KR     margin = rel * fmax(KR(in.ext), fmax(fabs(p.x), fabs(p.y)));
    const int    cx = cellCoord(pr.x, inv, gw), cy = cellCoord(pr.y, inv, gh);
    int          nc = 0;
    for (int r = kExactFirstRings;; ++r) {
        const int first      = r == kExactFirstRings;
        const int roundCells = first ? (2 * r + 1) * (2 * r + 1) : 8 * r;
        for (int q0 = 0; q0 < roundCells; q0 += 64) {
            int b = 0, cnt = 0;
            if (q0 + lane < roundCells) {
                int x, y;
                if (first) x = cx - r + lane % (2 * r + 1), y = cy - r + lane / (2 * r + 1);
                else ringCell(r, q0 + lane, cx, cy, x, y);
                if (x >= 0 && x < gw && y >= 0 && y < gh) {
                    b   = strt[y * gw + x];
                    cnt = strt[y * gw + x + 1] - b;
                }
            }
            const int incl   = waveInclusiveSum(cnt);
            lds.cb[lane]     = b;
            lds.co[lane + 1] = incl;
            if (lane == 0) lds.co[0] = 0;
            const int total = __builtin_amdgcn_readlane(incl, 63);
            waveSync();
            for (int s0 = 0; s0 < total;) {
                const int slice = min(total - s0, kExactCand - nc);
                for (int s = lane; s < slice; s += 64) {
                    // Cell of entry s0 + s: the last c with co[c] <= s0 + s.
                    const int g  = s0 + s;
                    int       lo = 0;
#pragma unroll
                    for (int step = 32; step > 0; step >>= 1)
                        if (lds.co[lo + step] <= g) lo += step;
                    const int    site = lds.cb[lo] + (g - lds.co[lo]);
                    const V2<KR> q{KR(spos[site].x), KR(spos[site].y)};
                    lds.cand[nc + s] = Key<KR>::make(dist2(q, p), (uint32_t(sidx[site]) << kLdsSlotBits) | uint32_t(kLdsNone));
                    if constexpr (P) lds.pos[nc + s] = q;
                }
                waveSync();
                nc += slice;
                s0 += slice;
                if (nc == kExactCand) nc = keepNearest(lds, nc, lane, K);
            }
        }
        nc = keepNearest(lds, nc, lane, K);
        const int x0 = cx - r, x1 = cx + r, y0 = cy - r, y1 = cy + r;
        if (x0 <= 0 && y0 <= 0 && x1 >= gw - 1 && y1 >= gh - 1) break;
        if (nc == K) {
            KR safe = fmin(fmin(p.x - KR(x0) * cell, KR(x1 + 1) * cell - p.x),
                           fmin(p.y - KR(y0) * cell, KR(y1 + 1) * cell - p.y));
            safe    = fmax(safe - margin, KR(0));
            if (lds.cand[K - 1].d2() * (KR(1) + rel) < safe * safe) break;
        }
    }
    return nc;
}

// The exact selection for queue entry e (item `it`) by one wavefront, with the
// K nearest sites of the entry's instance.
template <class R, bool P>
__device__ __forceinline__ void exactItem(const LossArgs<R> &a, ExactLds<R, P> &lds, const ExactItem it, int e, int lane)
{
    const bool      listOnly = it.inst < 0;
    const int       inst     = it.inst & 0x7FFFFFFF;
    const Inst      in       = loadUniform(a.insts + inst);
    const int       K        = in.K;
    const int64_t   so       = in.siteOff;
    const V2<R>    *spos     = a.spos + so;
    const int      *sidx     = a.sidx + so;
    const int      *strt     = a.start + in.cellOff;
    const V2<R>     p        = a.tpos[it.tg];
    const int       nc       = gatherKnn<R, R, P>(lds, in, spos, sidx, strt, p, lane, K);
    if (a.dbgKnn && inst == a.dbgInst && lane < K)
        a.dbgKnn[(int64_t(it.tg) - in.targetOff) * K + lane] = lane < nc ? keySite(lds.cand[lane].low()) : -1;
    if (listOnly) {
        waveSync();
        return;
    }

    // ---- the term: lane k evaluates neighbour k ----------------------------
    LaneTerm<R> tr;
    bool        ok = false;
    R           dd = rInf<R>();
    if (nc > 0) {
        tr.i0 = keySite(lds.cand[0].low());
        // The sorted copy of the positions holds the same values as a.pos.
        if constexpr (P) tr.c0 = lds.pos[0];
        else tr.c0 = a.pos[so + tr.i0];
        if (lane >= 1 && lane < nc) {
            const int   j  = keySite(lds.cand[lane].low());
            V2<R>       cj;
            if constexpr (P) cj = lds.pos[lane];
            else cj = a.pos[so + j];
            const V2<R> ev{cj.x - tr.c0.x, cj.y - tr.c0.y};
            const R     l = sqrtR(dotR(ev, ev));
            if (!(l < (sizeof(R) == 4 ? R(1e-12f) : R(1e-12)))) {
                tr.d = fmaR(R(-0.5), l, dotR(V2<R>{p.x - tr.c0.x, p.y - tr.c0.y}, ev) / l);
                tr.ev = ev, tr.l = l, tr.bj = j;
                dd    = tr.d * tr.d;
                ok    = true;
            }
        }
    }
    waveSync();
    const R        best = waveMinR<R>(dd);
    const uint64_t win  = __ballot(ok && dd == best);
    const int      wl   = win ? __builtin_ctzll(win) : -1;
    const R        lossT = win ? best : R(0);
    if (lane == wl) {
        if (a.grad &&
            lossGrad(V2<R>{p.x - tr.c0.x, p.y - tr.c0.y}, tr.ev, R(1) / tr.l, tr.d, &a.grad[2 * (so + tr.bj)],
                     &a.grad[2 * (so + tr.i0)], __builtin_ldexp(1.0, in.gradShift)))
            atomicAdd(reinterpret_cast<unsigned long long *>(&a.state[inst].saturated), 1ull);
        if (a.mode == kModeIterate) {
            LossPartial *pp = a.partial + int64_t(inst) * kLossPartials + (e & (kLossPartials - 1));
            unsafeAtomicAdd(&pp->loss, double(lossT));
            atomicAdd(&pp->count, 1ull);
        } else {
            InstState *st = a.state + inst;
            unsafeAtomicAdd(&st->evalLoss, double(lossT));
            atomicAdd((unsigned long long *)&st->evalCount, 1ull);
        }
    }
    if (a.dbgTerm && lane == (win ? wl : 0)) {
        a.dbgTerm[it.tg].i0   = win ? tr.i0 : -1;
        a.dbgTerm[it.tg].j    = win ? tr.bj : -1;
        a.dbgTerm[it.tg].loss = double(lossT);
    }
}

template <class R>
__global__ __launch_bounds__(64 * kExactWaves) void exactKernel(LossArgs<R> a)
{
    // FP32: positions beside the keys (a short queue, latency-bound).  FP64 runs
    // send every target here, and the smaller LDS footprint (more workgroups per
    // CU) is worth more than the dependent loads it saves.
    constexpr bool P = sizeof(R) == 4;
    __shared__ ExactLds<R, P> ldsAll[kExactWaves];
    const int          wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    ExactLds<R, P>    &lds  = ldsAll[wave];
    const int    lane = threadIdx.x & 63;
    const int    n    = *a.exactLen;
    const int    nw   = gridDim.x * kExactWaves;
    // Entry e goes to wavefront e % nw, numbered across workgroups first, so a
    // short queue spreads over the chip.
    for (int e = blockIdx.x + gridDim.x * wave; e < n; e += nw) exactItem<R, P>(a, lds, loadUniform(a.exactQ + e), e, lane);
}
// ---------------------------------------------------------------------------
// 4. Adam + MultiStepLR + clamp (voro_optimizer.cu:268-301), then bin the new
//    position for the next step (voro_optimizer.cu:168-177).
// ---------------------------------------------------------------------------
// Cells a site chunk bins through LDS: when the chunk's new cells span at most
// this many consecutive cell indices (sites given in a spatially coherent
// order, as grid slots are), the chunk counts its sites per cell in LDS and
// takes one global atomic per occupied cell; otherwise one per site.
constexpr int kBinSpan = 2048;

template <class R>
__global__ __launch_bounds__(kSiteChunk) void adamBinKernel(AdamArgs<R> a)
{
    __shared__ long long sCount;
    __shared__ int       sBin[kBinSpan];
    __shared__ int       sLo[kSiteChunk / 64], sHi[kSiteChunk / 64];
    const int nWork = *a.nWork;
    for (int w = blockIdx.x; w < nWork; w += gridDim.x) {
        const SiteWork wk = loadUniform(a.work + w);
        const Inst     in = loadUniform(a.insts + wk.inst);
        InstState     *st = a.state + wk.inst;
        if (a.doAdam && loadUniform(&st->done)) continue;
        const int     i    = wk.begin + int(threadIdx.x);
        const bool    mine = i < wk.begin + wk.count;
        const int64_t g    = wk.siteOff + i;
        // The site's state is loaded before the loss partials are folded, so
        // the two latencies overlap.
        V2<R>    p{}, mo{}, vo{};
        GradT<R> gx{}, gy{};
        if (mine) {
            p = a.pos[g];
            if (a.doAdam) gx = a.grad[2 * g], gy = a.grad[2 * g + 1], mo = a.m[g], vo = a.v[g];
        }
        // An empty slot (x < 0 or NaN: slot layouts keep them in place) is
        // never moved and never binned (voro_optimizer.cu:168-177, 268-301).
        const bool site = mine && p.x >= R(0);
        if (a.doAdam) {
            // Fold the step's loss partials: every chunk needs the target count,
            // the instance's first chunk also publishes the totals.
            if (threadIdx.x < 64) {
                const LossPartial &pp = a.partial[int64_t(wk.inst) * kLossPartials + threadIdx.x];
                const uint32_t     c  = waveReduce32<uint32_t>(uint32_t(pp.count), 0u,
                                                               [](uint32_t x, uint32_t y) { return x + y; });
                const double       l  = waveSumDouble(pp.loss);
                if (threadIdx.x == 0) {
                    sCount = (long long)c;
                    if (wk.begin == 0) st->count = (long long)c, st->loss = l;
                }
            }
            __syncthreads();
        }
        if (mine) {
            if (a.doAdam) {
                const long long cnt  = sCount;
                const double    ginv = __builtin_ldexp(1.0, -in.gradShift);
                V2<R>           gr{GradOps<R>::toReal(gx, ginv), GradOps<R>::toReal(gy, ginv)};
                a.grad[2 * g]     = GradT<R>(0);
                a.grad[2 * g + 1] = GradT<R>(0);
                bool movable      = site && (in.allMovable ||
                                        (in.rectW > 0 ? (i % in.rectW >= in.rx0 && i % in.rectW < in.rx1 &&
                                                         i / in.rectW >= in.ry0 && i / in.rectW < in.ry1)
                                                      : !a.frozen[g]));
                if (movable) {
                    const int t  = loadUniform(&st->iter);
                    R         lr = R(in.lr);
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        if (t - 1 >= in.milestones[k]) lr = lr * R(in.gamma);
                    const R b1 = R(in.beta1), b2 = R(in.beta2), eps = R(in.eps);
                    const R bc1 = R(1.0 - powi(double(b1), t));
                    const R bc2 = R(1.0 - powi(double(b2), t));
                    const R inv = R(1) / R(cnt > 1 ? cnt : 1);
                    const R ob1 = R(1) - b1, ob2 = R(1) - b2;
                    V2<R>   gg{inv * gr.x, inv * gr.y};
                    V2<R>   mi{b1 * mo.x + ob1 * gg.x, b1 * mo.y + ob1 * gg.y};
                    V2<R>   vi{b2 * vo.x + ob2 * (gg.x * gg.x), b2 * vo.y + ob2 * (gg.y * gg.y)};
                    a.m[g] = mi;
                    a.v[g] = vi;
                    p.x    = p.x - (lr * (mi.x / bc1)) / (sqrtR(vi.x / bc2) + eps);
                    p.y    = p.y - (lr * (mi.y / bc1)) / (sqrtR(vi.y / bc2) + eps);
                    p.x    = fmin(fmax(p.x, R(0)), R(in.W));
                    p.y    = fmin(fmax(p.y, R(0)), R(in.H));
                    a.pos[g] = p;
                }
            }
        }
        int c = 0;
        if (site) {
            const R inv = R(1) / R(in.cell);
            c           = cellCoord(p.y, inv, in.gh) * in.gw + cellCoord(p.x, inv, in.gw);
            a.cellOf[g] = c;
        } else if (mine) {
            a.cellOf[g] = -1;
        }
        // Rank of the site within its cell for the counting sort (any order
        // within a cell gives the same results: the k-NN order is (d2, index)).
        const int wv = int(threadIdx.x) >> 6;
        const int lo = waveReduce32<int>(site ? c : INT_MAX, INT_MAX, [](int x, int y) { return min(x, y); });
        const int hi = waveReduce32<int>(site ? c : -1, -1, [](int x, int y) { return max(x, y); });
        if ((threadIdx.x & 63) == 0) sLo[wv] = lo, sHi[wv] = hi;
        __syncthreads();
        int cLo = sLo[0], cHi = sHi[0];
#pragma unroll
        for (int k = 1; k < kSiteChunk / 64; ++k) cLo = min(cLo, sLo[k]), cHi = max(cHi, sHi[k]);
        if (cHi >= cLo && cHi - cLo < kBinSpan) {
            const int span = cHi - cLo + 1;
            for (int b = int(threadIdx.x); b < span; b += kSiteChunk) sBin[b] = 0;
            __syncthreads();
            const int local = site ? atomicAdd(&sBin[c - cLo], 1) : 0;
            __syncthreads();
            for (int b = int(threadIdx.x); b < span; b += kSiteChunk) {
                const int n = sBin[b];
                if (n) sBin[b] = atomicAdd(&a.counts[in.cellOff + cLo + b], n);
            }
            __syncthreads();
            if (site) a.rank[g] = sBin[c - cLo] + local;
        } else if (site) {
            a.rank[g] = atomicAdd(&a.counts[in.cellOff + c], 1);
        }
        __syncthreads(); // sCount, sBin and the bounds are rewritten by the next chunk
    }
}

// ---------------------------------------------------------------------------
// Diagnostic k-NN of arbitrary queries against one instance's current grid.
// ---------------------------------------------------------------------------
template <int K, class R>
__global__ void knnQueryKernel(const Inst *__restrict__ insts, int inst, const V2<R> *__restrict__ spos,
                               const int *__restrict__ sidx, const int *__restrict__ start,
                               const V2<R> *__restrict__ q, int nq, int kOut, int *outIdx, R *outD2)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const Inst &in = insts[inst];
    Key<R>      keys[K];
    knnGlobal<K>(in, spos + in.siteOff, sidx + in.siteOff, start + in.cellOff, q[i], keys);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k < kOut) {
            outIdx[int64_t(i) * kOut + k] = keys[k].valid() ? keySite(keys[k].low()) : -1;
            outD2[int64_t(i) * kOut + k]  = keys[k].valid() ? keys[k].d2() : R(FLT_MAX);
        }
    }
}

// ---------------------------------------------------------------------------
// Cell clipping (voro_optimizer.cu:312-437).  FP64 predicates and
// intersections; the polygon is rebuilt in place through a small ring buffer.
// ---------------------------------------------------------------------------
__device__ bool occupied(const uint64_t *masks, int tilesX, int tilesY, float invVoxel, float px, float py)
{
    if (!masks) return false;
    int i = int(floorf(px * invVoxel)), j = int(floorf(py * invVoxel));
    if (i < 0 || j < 0) return false;
    int tx = i >> 3, ty = j >> 3;
    if (tx >= tilesX || ty >= tilesY) return false;
    return (masks[int64_t(ty) * tilesX + tx] >> (((j & 7) << 3) | (i & 7))) & 1ull;
}

// One wavefront per site (4 per workgroup).  The cell polygon is held one
// vertex per lane; LDS holds the k-NN gather and one clip's output.
constexpr int kClipWaves = 4;

struct ClipLds {
    ExactLds<double, false> knn;
    double           px[kMaxCellVerts], py[kMaxCellVerts];
    int              pe[kMaxCellVerts];
};

// Clips the polygon (vertex k in lane k < n, relative to the site) to the
// half-plane <x, r> <= |r|^2 / 2 of the bisector with neighbour `label` at
// offset r (voro_optimizer.cu:312-347).  The output vertices are placed by a
// prefix sum in the order a sequential Sutherland-Hodgman pass emits them and
// computed with the same FP64 expressions, so the polygon is bit-identical; an
// exit crossing takes the new label, an entry keeps the old edge's.  Vertices
// beyond kMaxCellVerts are dropped and the call returns false.
__device__ bool clipWave(ClipLds &lds, double &vx, double &vy, int &ve, int &n, double rx, double ry, int label,
                         int lane)
{
    const double c   = 0.5 * (rx * rx + ry * ry);
    const bool   act = lane < n;
    const double f   = act ? vx * rx + vy * ry - c : 0.0;
    if (!__ballot(act && f > 0.0)) return true;
    const int    k1  = lane + 1 == n ? 0 : lane + 1;
    const double bx  = __shfl(vx, k1), by = __shfl(vy, k1), fb = __shfl(f, k1);
    const bool   inA = f <= 0.0, inB = fb <= 0.0;
    const int    cnt = act ? int(inA) + int(inA != inB) : 0;
    const int    inc = waveInclusiveSum(cnt);
    const int    pos = inc - cnt, total = __builtin_amdgcn_readlane(inc, 63);
    if (act && inA && pos < kMaxCellVerts) lds.px[pos] = vx, lds.py[pos] = vy, lds.pe[pos] = ve;
    if (act && inA != inB) {
        const int q = pos + int(inA);
        if (q < kMaxCellVerts) {
            const double t = f / (f - fb);
            lds.px[q] = vx + t * (bx - vx), lds.py[q] = vy + t * (by - vy), lds.pe[q] = inA ? label : ve;
        }
    }
    waveSync();
    n = min(total, kMaxCellVerts);
    if (lane < n) vx = lds.px[lane], vy = lds.py[lane], ve = lds.pe[lane];
    waveSync();
    return total <= kMaxCellVerts;
}

// Clipped Voronoi cell of every site of one instance (voro_optimizer.cu:361-437):
// the box clipped by the bisectors of the 32 nearest sites in FP64 order,
// certified once the next neighbour is farther than twice the cell's radius,
// else by brute force over every site within that distance (ascending index);
// FP64 shoelace centroid -> occupancy.
template <class R>
__global__ __launch_bounds__(64 * kClipWaves) void clipKernel(ClipArgs<R> a)
{
    __shared__ ClipLds ldsAll[kClipWaves];
    const int    wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    ClipLds     &lds  = ldsAll[wave];
    const int    lane = threadIdx.x & 63;
    const Inst   in   = loadUniform(a.insts + a.inst);
    const V2<R> *pos  = a.pos + in.siteOff;
    const V2<R> *spos = a.spos + in.siteOff;
    const int   *sidx = a.sidx + in.siteOff;
    const int   *strt = a.start + in.cellOff;
    for (int i = blockIdx.x * kClipWaves + wave; i < in.S; i += gridDim.x * kClipWaves) {
        const V2<R>  ci  = pos[i];
        if (!(ci.x >= R(0))) { // empty slot: no cell (voro_optimizer.cu:361-437 skips it)
            if (lane == 0) a.inside[i] = 0, a.count[i] = 0;
            continue;
        }
        const int    nc  = gatherKnn<R, double, false>(lds.knn, in, spos, sidx, strt, ci, lane, kCellK);
        const double cxd = double(ci.x), cyd = double(ci.y);
        const double x0 = -cxd, y0 = -cyd, x1 = double(in.W) - cxd, y1 = double(in.H) - cyd;
        // The box: lanes 0-3 hold (x0,y0), (x1,y0), (x1,y1), (x0,y1), edges -1..-4.
        double vx = (lane == 0 || lane == 3) ? x0 : x1, vy = lane < 2 ? y0 : y1;
        int    ve = -1 - lane, n = 4;
        bool   ok = true, certified = false;
        double r2max = waveMaxR<double>(lane < n ? vx * vx + vy * vy : 0.0);
        for (int k = 0; k < kCellK; ++k) {
            const Key<double> key = lds.knn.cand[k < nc ? k : 0];
            const int         j   = k < nc ? keySite(key.low()) : -1;
            if (j < 0) {
                certified = true;
                break;
            }
            if (j == i) continue;
            if (key.d2() > 4.0 * r2max) {
                certified = true;
                break;
            }
            const V2<R> cj = pos[j];
            ok    = clipWave(lds, vx, vy, ve, n, double(cj.x) - cxd, double(cj.y) - cyd, j, lane) && ok;
            r2max = waveMaxR<double>(lane < n ? vx * vx + vy * vy : 0.0);
        }
        if (!certified && lds.knn.cand[kCellK - 1].d2() > 4.0 * r2max) certified = true;
        if (!certified) {
            if (lane == 0) atomicAdd(&a.flags[0], 1);
            for (int jb = 0; jb < in.S; jb += 64) {
                const int j    = jb + lane;
                bool      cand = false;
                if (j < in.S && j != i) {
                    bool seen = false;
                    for (int k = 0; k < kCellK; ++k) seen |= k < nc && keySite(lds.knn.cand[k].low()) == j;
                    if (!seen) {
                        const V2<R>  cj = pos[j];
                        const double rx = double(cj.x) - cxd, ry = double(cj.y) - cyd;
                        // r2max only shrinks: rechecked below.  Empty slots are no neighbours.
                        cand = cj.x >= R(0) && !(rx * rx + ry * ry > 4.0 * r2max);
                    }
                }
                for (uint64_t m = __ballot(cand); m; m &= m - 1) {
                    const int    jj = jb + __builtin_ctzll(m);
                    const V2<R>  cj = pos[jj];
                    const double rx = double(cj.x) - cxd, ry = double(cj.y) - cyd;
                    if (rx * rx + ry * ry > 4.0 * r2max) continue;
                    ok    = clipWave(lds, vx, vy, ve, n, rx, ry, jj, lane) && ok;
                    r2max = waveMaxR<double>(lane < n ? vx * vx + vy * vy : 0.0);
                }
            }
        }
        if (!ok && lane == 0) atomicAdd(&a.flags[1], 1);
        // The polygon in LDS for the sequential centroid sums and the output.
        if (lane < n) lds.px[lane] = vx, lds.py[lane] = vy, lds.pe[lane] = ve;
        waveSync();
        if (lane == 0) {
            double area = 0.0, sx = 0.0, sy = 0.0;
            for (int k = 0; k < n; ++k) {
                const int    k1 = (k + 1 == n) ? 0 : k + 1;
                const double cr = lds.px[k] * lds.py[k1] - lds.px[k1] * lds.py[k];
                area += cr;
                sx += (lds.px[k] + lds.px[k1]) * cr;
                sy += (lds.py[k] + lds.py[k1]) * cr;
            }
            float cenx = float(ci.x), ceny = float(ci.y);
            if (fabs(area) > 1e-12) {
                cenx = cenx + float(sx / (3.0 * area));
                ceny = ceny + float(sy / (3.0 * area));
            }
            a.inside[i] = occupied(a.masks, a.tilesX, a.tilesY, a.invVoxel, cenx, ceny) ? 1 : 0;
            a.count[i]  = n;
        }
        if (lane < n) {
            a.verts[(int64_t(i) * kMaxCellVerts + lane) * 2 + 0] = float(cxd + vx);
            a.verts[(int64_t(i) * kMaxCellVerts + lane) * 2 + 1] = float(cyd + vy);
            a.labels[int64_t(i) * kMaxCellVerts + lane]          = ve;
        }
        waveSync();
    }
}

// ---------------------------------------------------------------------------
// 5b. Boundary assembly (voro_optimizer.cu:798-854) on the device, by one
//     workgroup in phases separated by barriers (boundaries hold O(sqrt S)
//     edges, and one launch costs less than a chain of them):
//   count   boundary edges per inside cell, scanned into offsets;
//   emit    each edge's two end points in the order (cell, edge): key = the
//           sorted label triple (cell, previous/next label, edge label);
//   insert  end points into an open-addressing table keyed by the triple; a
//           slot keeps the smallest end point index of its key (atomicMin);
//   flag    an end point is its key's representative when it is that minimum;
//           vertex ids = exclusive scan of the flags (first-appearance order);
//   final   vertex positions from the representatives, edge ids, in/out
//           degrees and each vertex's successor;
//   loops   watertight iff every vertex has one in- and one out-edge; loops =
//           cycles of the successor map, by pointer jumping (after round r a
//           vertex holds the smallest id among its next 2^r successors, so
//           with 2^r >= V every vertex holds its cycle's minimum; loops =
//           vertices that are their own minimum).
// More than `cap` end points: result[3] = 1 and nothing past the count.
// ---------------------------------------------------------------------------
constexpr int      kBndThreads  = 1024;
constexpr int      kBndLdsVerts = 4096;
constexpr uint32_t kBndEmpty    = ~0u;

__device__ __forceinline__ bool facesOut(int l, const unsigned char *inside) { return l < 0 || !inside[l]; }

__device__ __forceinline__ BndKey bndKey(int x, int y, int z)
{
    uint32_t v0 = uint32_t(x + kBoxLabels), v1 = uint32_t(y + kBoxLabels), v2 = uint32_t(z + kBoxLabels), t;
    if (v0 > v1) t = v0, v0 = v1, v1 = t;
    if (v1 > v2) t = v1, v1 = v2, v2 = t;
    if (v0 > v1) t = v0, v0 = v1, v1 = t;
    return BndKey{v0, v1, v2, 0u};
}

__device__ __forceinline__ uint32_t bndHash(const BndKey &k)
{
    uint32_t h = k.a * 0x9E3779B1u;
    h ^= k.b * 0x85EBCA77u + (h << 6) + (h >> 2);
    h ^= k.c * 0xC2B2AE3Du + (h << 6) + (h >> 2);
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    return h ^ (h >> 12);
}

// Exclusive scan of the values val(0..n-1) by the workgroup (contiguous
// segments per thread) into out; returns the total to every thread.
template <class F> __device__ int bndScan(F val, int *out, int n, int *waveTot)
{
    const int t = int(threadIdx.x), per = (n + kBndThreads - 1) / kBndThreads;
    const int b = min(t * per, n), e = min(b + per, n);
    int       s = 0;
    for (int i = b; i < e; ++i) s += val(i);
    const int incl = waveInclusiveSum(s);
    if ((t & 63) == 63) waveTot[t >> 6] = incl;
    __syncthreads();
    int run = incl - s, total = 0;
    for (int w = 0; w < kBndThreads / 64; ++w) {
        if (w < (t >> 6)) run += waveTot[w];
        total += waveTot[w];
    }
    for (int i = b; i < e; ++i) {
        const int v = val(i);
        out[i]      = run;
        run += v;
    }
    __syncthreads();
    return total;
}

// Pointer jumping over the successor map nxt0 of V vertices (in LDS or in
// global scratch); returns how many of this thread's vertices are the
// smallest of their cycle.
template <class P>
__device__ int bndJump(P lab, P nx, P labN, P nxN, const int *nxt0, int V)
{
    const int t = int(threadIdx.x);
    for (int v = t; v < V; v += kBndThreads) lab[v] = v, nx[v] = nxt0[v];
    __syncthreads();
    for (int span = 1; span < V; span *= 2) {
        for (int v = t; v < V; v += kBndThreads) {
            const int w = nx[v];
            labN[v]     = min(lab[v], lab[w]);
            nxN[v]      = nx[w];
        }
        __syncthreads();
        P tl = lab, tn = nx;
        lab = labN, nx = nxN, labN = tl, nxN = tn;
    }
    int own = 0;
    for (int v = t; v < V; v += kBndThreads) own += lab[v] == v ? 1 : 0;
    return own;
}

__global__ __launch_bounds__(kBndThreads) void boundaryKernel(BoundaryArgs a)
{
    __shared__ int waveTot[kBndThreads / 64];
    __shared__ int sBad, sLoops;
    __shared__ int sJump[4][kBndLdsVerts]; // pointer-jumping labels and successors, double-buffered
    const int      t = int(threadIdx.x);
    if (t == 0) sBad = 0, sLoops = 0;
    // count
    for (int i = t; i < a.S; i += kBndThreads) {
        int n = 0;
        if (a.inside[i]) {
            const int *lab = a.labels + size_t(i) * kMaxCellVerts;
            for (int k = 0; k < a.count[i]; ++k) n += facesOut(lab[k], a.inside) ? 1 : 0;
        }
        a.ecount[i] = n;
    }
    __syncthreads();
    const int E = bndScan([&](int i) { return a.ecount[i]; }, a.eoff, a.S, waveTot);
    const int n = 2 * E;
    if (n > a.cap) {
        if (t == 0) a.result[0] = 0, a.result[1] = 0, a.result[2] = 0, a.result[3] = 1, a.result[4] = E;
        return;
    }
    for (uint32_t h = uint32_t(t); h <= a.tableMask; h += kBndThreads) a.table[h] = kBndEmpty;
    for (int v = t; v < n; v += kBndThreads) a.outDeg[v] = 0, a.inDeg[v] = 0;
    // emit
    for (int i = t; i < a.S; i += kBndThreads) {
        if (!a.inside[i]) continue;
        const int    cc  = a.count[i];
        const int   *lab = a.labels + size_t(i) * kMaxCellVerts;
        const float *vp  = a.verts + size_t(i) * kMaxCellVerts * 2;
        int          j   = 2 * a.eoff[i];
        for (int k = 0; k < cc; ++k) {
            const int l = lab[k];
            if (!facesOut(l, a.inside)) continue;
            const int kp = k == 0 ? cc - 1 : k - 1, kn = k + 1 == cc ? 0 : k + 1;
            a.key[j]         = bndKey(i, lab[kp], l);
            a.key[j + 1]     = bndKey(i, l, lab[kn]);
            a.pos[2 * j]     = vp[2 * k];
            a.pos[2 * j + 1] = vp[2 * k + 1];
            a.pos[2 * j + 2] = vp[2 * kn];
            a.pos[2 * j + 3] = vp[2 * kn + 1];
            j += 2;
        }
    }
    __syncthreads();
    // insert (the table has at least twice as many slots as end points, so
    // probing ends)
    for (int j = t; j < n; j += kBndThreads) {
        const BndKey k = a.key[j];
        uint32_t     h = bndHash(k) & a.tableMask;
        for (;;) {
            const uint32_t v = atomicCAS(&a.table[h], kBndEmpty, uint32_t(j));
            if (v == kBndEmpty) break;
            const BndKey o = a.key[v];
            if (o.a == k.a && o.b == k.b && o.c == k.c) {
                atomicMin(&a.table[h], uint32_t(j));
                break;
            }
            h = (h + 1) & a.tableMask;
        }
        a.slot[j] = int(h);
    }
    __syncthreads();
    // flag
    for (int j = t; j < n; j += kBndThreads) a.rep[j] = int(a.table[a.slot[j]]);
    __syncthreads();
    const int V = bndScan([&](int j) { return a.rep[j] == j ? 1 : 0; }, a.vid, n, waveTot);
    // final
    for (int j = t; j < n; j += kBndThreads) {
        const int r = a.rep[j], id = a.vid[r];
        a.edges[j]  = id;
        if (r == j) a.vertices[2 * id] = a.pos[2 * j], a.vertices[2 * id + 1] = a.pos[2 * j + 1];
        if ((j & 1) == 0) {
            atomicAdd(&a.outDeg[id], 1);
            a.nxt[id] = a.vid[a.rep[j + 1]];
        } else {
            atomicAdd(&a.inDeg[id], 1);
        }
    }
    __syncthreads();
    // loops
    for (int v = t; v < V; v += kBndThreads)
        if (a.outDeg[v] != 1 || a.inDeg[v] != 1) atomicOr(&sBad, 1);
    __syncthreads();
    const bool wt = sBad == 0;
    if (wt) {
        // Up to kBndLdsVerts vertices the rounds run in LDS.
        const int own = V <= kBndLdsVerts ? bndJump(sJump[0], sJump[1], sJump[2], sJump[3], a.nxt, V)
                                          : bndJump(a.lab0, a.nxt, a.lab1, a.nxt1, a.nxt, V);
        if (own) atomicAdd(&sLoops, own);
        __syncthreads();
    }
    if (t == 0) a.result[0] = V, a.result[1] = wt ? 1 : 0, a.result[2] = wt ? sLoops : 0, a.result[3] = 0, a.result[4] = E;
}

// ---------------------------------------------------------------------------
// Poisson-disk dart throwing over a table of instance grids
// (poisson_disk.cu:138-177).  Cells of one phase group are >= 3 apart, so
// every thread writes only its own cell and the result is deterministic.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void poissonClearKernel(const PoiInst *__restrict__ insts,
                                                          const PoiWork *__restrict__ work, int nWork, float *grid)
{
    // Clear chunks cover the whole region, one cell per thread.
    for (int w = blockIdx.x; w < nWork; w += gridDim.x) {
        PoiWork        wk = work[w];
        const PoiInst &pi = insts[wk.inst];
        int            rw = pi.x1 - pi.x0, rh = pi.y1 - pi.y0;
        int            c  = wk.base + int(threadIdx.x);
        if (c >= rw * rh) continue;
        int     cx = pi.x0 + c % rw, cy = pi.y0 + c / rw;
        float2 *g  = reinterpret_cast<float2 *>(grid) + pi.gridOff;
        g[int64_t(cy) * pi.gw + cx] = make_float2(-1.0f, -1.0f);
    }
}

__global__ __launch_bounds__(256) void poissonPhaseKernel(const PoiInst *__restrict__ insts,
                                                          const PoiWork *__restrict__ work, int nWork, float *grid,
                                                          int px, int py, int round)
{
    for (int w = blockIdx.x; w < nWork; w += gridDim.x) {
        PoiWork        wk = work[w];
        const PoiInst &pi = insts[wk.inst];
        int            rw = pi.x1 - pi.x0, rh = pi.y1 - pi.y0;
        int            nx = (rw - px + 2) / 3, ny = (rh - py + 2) / 3;
        int            k  = wk.base + int(threadIdx.x);
        if (nx <= 0 || ny <= 0 || k >= nx * ny) continue;
        int     cx = pi.x0 + 3 * (k % nx) + px, cy = pi.y0 + 3 * (k / nx) + py;
        float2 *g  = reinterpret_cast<float2 *>(grid) + pi.gridOff;
        int     cellId = cy * pi.gw + cx;
        if (g[cellId].x >= 0.0f) continue;
        uint32_t rnd = pi.roundBase + uint32_t(round);
        float    x   = (float(cx) + rand01_poisson(uint32_t(cellId), rnd, pi.seed, 0)) * pi.cellSize;
        float    y   = (float(cy) + rand01_poisson(uint32_t(cellId), rnd, pi.seed, 1)) * pi.cellSize;
        if (x >= pi.W || y >= pi.H) continue;
        float r2 = pi.radius * pi.radius;
        int   nx0 = max(cx - 2, 0), nx1 = min(cx + 2, pi.gw - 1);
        int   ny0 = max(cy - 2, 0), ny1 = min(cy + 2, pi.gh - 1);
        bool  ok  = true;
        for (int yy = ny0; yy <= ny1 && ok; ++yy)
            for (int xx = nx0; xx <= nx1; ++xx) {
                float2 q = g[yy * pi.gw + xx];
                if (q.x < 0.0f) continue;
                float dx = q.x - x, dy = q.y - y;
                if (__builtin_fmaf(dy, dy, dx * dx) < r2) { ok = false; break; }
            }
        if (ok) g[cellId] = make_float2(x, y);
    }
}

// The whole resample of one region in one workgroup: the region (cleared) and
// its 2-cell halo (kept: the dart test reads cells within +-2) are staged in
// LDS, every (round, phase) step of the launch sequence above runs between
// block barriers in the same order and with the same arithmetic, and the region
// is written back.  One launch instead of 9 x trials; the grids are identical.
__global__ __launch_bounds__(kPoiRegionThreads) void poissonRegionKernel(PoiInst pi, float *grid, int trials)
{
    __shared__ float2 cells[kPoiRegionCells];
    const int hx0 = max(pi.x0 - 2, 0), hy0 = max(pi.y0 - 2, 0);
    const int hx1 = min(pi.x1 + 2, pi.gw), hy1 = min(pi.y1 + 2, pi.gh);
    const int hw = hx1 - hx0, hh = hy1 - hy0;
    float2   *g  = reinterpret_cast<float2 *>(grid) + pi.gridOff;
    for (int c = int(threadIdx.x); c < hw * hh; c += kPoiRegionThreads) {
        const int  cx = hx0 + c % hw, cy = hy0 + c / hw;
        const bool in = cx >= pi.x0 && cx < pi.x1 && cy >= pi.y0 && cy < pi.y1;
        cells[c]      = in ? make_float2(-1.0f, -1.0f) : g[int64_t(cy) * pi.gw + cx];
    }
    __syncthreads();
    const int   rw = pi.x1 - pi.x0, rh = pi.y1 - pi.y0;
    const float r2 = pi.radius * pi.radius;
    for (int round = 0; round < trials; ++round) {
        const uint32_t rnd = pi.roundBase + uint32_t(round);
        for (int py = 0; py < 3; ++py) {
            for (int px = 0; px < 3; ++px) {
                const int nx = (rw - px + 2) / 3, ny = (rh - py + 2) / 3;
                for (int k = int(threadIdx.x); k < nx * ny; k += kPoiRegionThreads) {
                    const int cx = pi.x0 + 3 * (k % nx) + px, cy = pi.y0 + 3 * (k / nx) + py;
                    const int lc = (cy - hy0) * hw + (cx - hx0);
                    if (cells[lc].x >= 0.0f) continue;
                    const int   cellId = cy * pi.gw + cx;
                    const float x      = (float(cx) + rand01_poisson(uint32_t(cellId), rnd, pi.seed, 0)) * pi.cellSize;
                    const float y      = (float(cy) + rand01_poisson(uint32_t(cellId), rnd, pi.seed, 1)) * pi.cellSize;
                    if (x >= pi.W || y >= pi.H) continue;
                    // The 5x5 neighbourhood (clipped to the grid) read at once:
                    // a conflict with any occupied cell rejects the dart.
                    bool ok = true;
#pragma unroll
                    for (int dyc = -2; dyc <= 2; ++dyc) {
#pragma unroll
                        for (int dxc = -2; dxc <= 2; ++dxc) {
                            const int    xx = cx + dxc, yy = cy + dyc;
                            const bool   in = xx >= 0 && xx < pi.gw && yy >= 0 && yy < pi.gh;
                            const float2 q  = cells[in ? (yy - hy0) * hw + (xx - hx0) : lc];
                            const float  dx = q.x - x, dy = q.y - y;
                            ok = ok && !(in && q.x >= 0.0f && __builtin_fmaf(dy, dy, dx * dx) < r2);
                        }
                    }
                    if (ok) cells[lc] = make_float2(x, y);
                }
                __syncthreads();
            }
        }
    }
    for (int c = int(threadIdx.x); c < rw * rh; c += kPoiRegionThreads) {
        const int cx = pi.x0 + c % rw, cy = pi.y0 + c / rw;
        g[int64_t(cy) * pi.gw + cx] = cells[(cy - hy0) * hw + (cx - hx0)];
    }
}

__global__ void fillIntKernel(int *p, int64_t n, int v)
{
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// ===========================================================================
// Host-side launchers
// ===========================================================================
hipError_t launchScan(int grid, hipStream_t s, const Inst *insts, InstState *state, LossPartial *partial,
                      const ScanTile *tiles, ScanCtl *ctl, unsigned long long *flags, int *counts, int *start,
                      int mode, int *exactLen, int nQueues, const ScanScatter &sc)
{
    if (nQueues > kScanThreads) return hipErrorInvalidValue;
    if (sc.posBytes != 0 && sc.posBytes != 8 && sc.posBytes != 16) return hipErrorInvalidValue;
    hipLaunchKernelGGL(scanKernel, dim3(grid), dim3(kScanThreads), 0, s, insts, state, partial, tiles, ctl, flags,
                       counts, start, mode, exactLen, nQueues, sc);
    return hipGetLastError();
}

template <class R>
hipError_t launchScatter(int grid, hipStream_t s, const Inst *insts, const InstState *state, const SiteWork *work,
                         const int *nWork, const void *pos, const int *cellOf, const int *rank, const int *start,
                         void *spos, int *sidx, int mode, ScanCtl *ctl)
{
    hipLaunchKernelGGL(scatterKernel<R>, dim3(grid), dim3(kSiteChunk), 0, s, insts, state, work, nWork,
                       static_cast<const V2<R> *>(pos), cellOf, rank, start, static_cast<V2<R> *>(spos), sidx, mode,
                       ctl);
    return hipGetLastError();
}

template <class R> hipError_t launchLoss(int tpl, int grid, hipStream_t s, const LossArgs<R> &a)
{
    // Evaluation passes and passes without a gradient take the instantiation
    // that also serves the diagnostics.
    const bool diag   = a.stats || a.dbgKnn || a.dbgTerm || a.exactOnly || a.mode != kModeIterate || !a.grad;
    const int  blocks = ((grid + kLossWaves - 1) / kLossWaves + 2 * kXcdRun - 1) / (2 * kXcdRun) * (2 * kXcdRun);
    if (!lossTplSupported(tpl, sizeof(R) == 8)) return hipErrorInvalidValue;
#define VM_T(t)                                                                                                   \
    if (tpl == t) {                                                                                               \
        if (diag)                                                                                                 \
            hipLaunchKernelGGL((lossKernel<t, R, true>), dim3(blocks), dim3(kLossThreads), 0, s, a);              \
        else                                                                                                      \
            hipLaunchKernelGGL((lossKernel<t, R, false>), dim3(blocks), dim3(kLossThreads), 0, s, a);             \
        return hipGetLastError();                                                                                 \
    }
    VM_T(3)
    if constexpr (sizeof(R) == 4) {
        VM_T(2) VM_T(4)
    }
#undef VM_T
    return hipErrorInvalidValue;
}
template <class R> hipError_t launchExact(hipStream_t s, const LossArgs<R> &a)
{
    hipLaunchKernelGGL((exactKernel<R>), dim3(a.exactGrid), dim3(64 * kExactWaves), 0, s, a);
    return hipGetLastError();
}

template <class R> hipError_t launchAdamBin(int grid, hipStream_t s, const AdamArgs<R> &a)
{
    hipLaunchKernelGGL(adamBinKernel<R>, dim3(grid), dim3(kSiteChunk), 0, s, a);
    return hipGetLastError();
}

template <class R>
hipError_t launchKnnQuery(int K, hipStream_t s, const Inst *insts, int inst, const void *spos, const int *sidx,
                          const int *start, const void *q, int nq, int *outIdx, void *outD2)
{
    dim3 grid((nq + 63) / 64), block(64);
    if (nq <= 0) return hipSuccess;
    auto *sp = static_cast<const V2<R> *>(spos);
    auto *qq = static_cast<const V2<R> *>(q);
    auto *od = static_cast<R *>(outD2);
    if (K <= 8) hipLaunchKernelGGL((knnQueryKernel<8, R>), grid, block, 0, s, insts, inst, sp, sidx, start, qq, nq, K, outIdx, od);
    else if (K <= 16) hipLaunchKernelGGL((knnQueryKernel<16, R>), grid, block, 0, s, insts, inst, sp, sidx, start, qq, nq, K, outIdx, od);
    else if (K <= 32) hipLaunchKernelGGL((knnQueryKernel<32, R>), grid, block, 0, s, insts, inst, sp, sidx, start, qq, nq, K, outIdx, od);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

template <class R> hipError_t launchClip(int nSites, hipStream_t s, const ClipArgs<R> &a)
{
    if (nSites <= 0) return hipSuccess;
    const int grid = std::min((nSites + kClipWaves - 1) / kClipWaves, 65536);
    hipLaunchKernelGGL(clipKernel<R>, dim3(grid), dim3(64 * kClipWaves), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Clipped cells -> CSR: exclusive scan of the vertex counts (tiles of
// kCsrTile cells, then the tile totals), and a gather of every cell's vertices
// and labels to its offset, so only the vertices that exist leave the device.
// ---------------------------------------------------------------------------
constexpr int kCsrThreads = 256, kCsrItems = 16, kCsrTile = kCsrThreads * kCsrItems;

__global__ __launch_bounds__(kCsrThreads) void csrScanKernel(const int *__restrict__ count, int n, int *start,
                                                             int *tileSum)
{
    __shared__ int wsum[kCsrThreads / 64];
    const int      base = blockIdx.x * kCsrTile + int(threadIdx.x) * kCsrItems;
    int            v[kCsrItems], tot = 0;
#pragma unroll
    for (int j = 0; j < kCsrItems; ++j) v[j] = base + j < n ? count[base + j] : 0, tot += v[j];
    const int inc = waveInclusiveSum(tot);
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = inc;
    __syncthreads();
    int pre = 0, btot = 0;
#pragma unroll
    for (int k = 0; k < kCsrThreads / 64; ++k) pre += k < int(threadIdx.x >> 6) ? wsum[k] : 0, btot += wsum[k];
    int ex = pre + inc - tot;
#pragma unroll
    for (int j = 0; j < kCsrItems; ++j) {
        if (base + j < n) start[base + j] = ex;
        ex += v[j];
    }
    if (threadIdx.x == 0) tileSum[blockIdx.x] = btot;
}

// Exclusive scan of the (few) tile totals in place; the grand total goes to start[n].
__global__ void csrTileKernel(int *tileSum, int nTiles, int *start, int n)
{
    if (threadIdx.x != 0) return;
    int s = 0;
    for (int t = 0; t < nTiles; ++t) {
        const int c = tileSum[t];
        tileSum[t]  = s;
        s += c;
    }
    start[n] = s;
}

__global__ void csrGatherKernel(const float *__restrict__ verts, const int *__restrict__ labels,
                                const int *__restrict__ count, int n, int *start, const int *__restrict__ tileSum,
                                float *cverts, int *clabels)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int o = start[i] + tileSum[i / kCsrTile];
        start[i]    = o;
        for (int k = 0; k < count[i]; ++k) {
            const int64_t src = int64_t(i) * kMaxCellVerts + k;
            cverts[2 * int64_t(o + k)]     = verts[2 * src];
            cverts[2 * int64_t(o + k) + 1] = verts[2 * src + 1];
            clabels[o + k]                 = labels[src];
        }
    }
}

int csrTiles(int n) { return (n + kCsrTile - 1) / kCsrTile; }

hipError_t launchClipCsr(hipStream_t s, const float *verts, const int *labels, const int *count, int n, int *start,
                         int *tileSum, float *cverts, int *clabels)
{
    if (n <= 0) return hipSuccess;
    const int nt = csrTiles(n);
    hipLaunchKernelGGL(csrScanKernel, dim3(nt), dim3(kCsrThreads), 0, s, count, n, start, tileSum);
    hipLaunchKernelGGL(csrTileKernel, dim3(1), dim3(64), 0, s, tileSum, nt, start, n);
    const int g = std::min((n + 255) / 256, 8192);
    hipLaunchKernelGGL(csrGatherKernel, dim3(g), dim3(256), 0, s, verts, labels, count, n, start, tileSum, cverts,
                       clabels);
    return hipGetLastError();
}

hipError_t launchBoundary(hipStream_t s, const BoundaryArgs &a)
{
    if (a.S <= 0) return hipSuccess;
    hipLaunchKernelGGL(boundaryKernel, dim3(1), dim3(kBndThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launchPoissonClear(int grid, hipStream_t s, const PoiInst *insts, const PoiWork *work, int nWork, float *g)
{
    if (nWork <= 0) return hipSuccess;
    hipLaunchKernelGGL(poissonClearKernel, dim3(grid), dim3(256), 0, s, insts, work, nWork, g);
    return hipGetLastError();
}

bool poissonRegionFits(const PoiInst &pi)
{
    const int64_t hw = int64_t(std::min(pi.x1 + 2, pi.gw) - std::max(pi.x0 - 2, 0));
    const int64_t hh = int64_t(std::min(pi.y1 + 2, pi.gh) - std::max(pi.y0 - 2, 0));
    return pi.x1 > pi.x0 && pi.y1 > pi.y0 && hw * hh <= kPoiRegionCells;
}

hipError_t launchPoissonRegion(hipStream_t s, const PoiInst &pi, float *grid, int trials)
{
    if (!poissonRegionFits(pi)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(poissonRegionKernel, dim3(1), dim3(kPoiRegionThreads), 0, s, pi, grid, trials);
    return hipGetLastError();
}

hipError_t launchPoissonPhase(int grid, hipStream_t s, const PoiInst *insts, const PoiWork *work, int nWork, float *g,
                              int px, int py, int round)
{
    if (nWork <= 0) return hipSuccess;
    hipLaunchKernelGGL(poissonPhaseKernel, dim3(grid), dim3(256), 0, s, insts, work, nWork, g, px, py, round);
    return hipGetLastError();
}

// Marks pooled targets dead (position (-1, -1), slot -1): the loss kernel takes
// no target with x < 0.  Local edits retire a target's sorted entry this way.
template <class R> __global__ void killTargetsKernel(V2<R> *tpos, int *tslot, const int64_t *idx, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int64_t t = idx[i];
        tpos[t]         = V2<R>{R(-1), R(-1)};
        tslot[t]        = -1;
    }
}

template <class R> hipError_t launchKillTargets(hipStream_t s, V2<R> *tpos, int *tslot, const int64_t *idx, int n)
{
    if (n <= 0) return hipSuccess;
    const int blocks = (n + 255) / 256;
    hipLaunchKernelGGL(killTargetsKernel<R>, dim3(unsigned(blocks > 1024 ? 1024 : blocks)), dim3(256), 0, s, tpos,
                       tslot, idx, n);
    return hipGetLastError();
}
template hipError_t launchKillTargets<float>(hipStream_t, V2<float> *, int *, const int64_t *, int);
template hipError_t launchKillTargets<double>(hipStream_t, V2<double> *, int *, const int64_t *, int);

hipError_t launchFillInt(hipStream_t s, int *p, int64_t n, int v)
{
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(fillIntKernel, dim3(unsigned(blocks > 4096 ? 4096 : blocks)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

template hipError_t launchScatter<float>(int, hipStream_t, const Inst *, const InstState *, const SiteWork *,
                                         const int *, const void *, const int *, const int *, const int *, void *,
                                         int *, int, ScanCtl *);
template hipError_t launchScatter<double>(int, hipStream_t, const Inst *, const InstState *, const SiteWork *,
                                          const int *, const void *, const int *, const int *, const int *, void *,
                                          int *, int, ScanCtl *);
template hipError_t launchLoss<float>(int, int, hipStream_t, const LossArgs<float> &);
template hipError_t launchLoss<double>(int, int, hipStream_t, const LossArgs<double> &);
template hipError_t launchExact<float>(hipStream_t, const LossArgs<float> &);
template hipError_t launchExact<double>(hipStream_t, const LossArgs<double> &);
template hipError_t launchAdamBin<float>(int, hipStream_t, const AdamArgs<float> &);
template hipError_t launchAdamBin<double>(int, hipStream_t, const AdamArgs<double> &);
template hipError_t launchKnnQuery<float>(int, hipStream_t, const Inst *, int, const void *, const int *, const int *,
                                          const void *, int, int *, void *);
template hipError_t launchKnnQuery<double>(int, hipStream_t, const Inst *, int, const void *, const int *,
                                           const int *, const void *, int, int *, void *);
template hipError_t launchClip<float>(int, hipStream_t, const ClipArgs<float> &);
template hipError_t launchClip<double>(int, hipStream_t, const ClipArgs<double> &);

} // namespace vmk
