// voro_oracle.cpp -- CPU restatement of the VoroMesh optimizer step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library: only
// the tests, the smoke check and the CPU-baseline leg of the benchmark may load
// it.  The product path never links or calls it.
//
// It restates, function by function, the algorithm of the reference's CUDA
// implementation (src/voromesh/voro_optimizer.cu, src/poisson/poisson_disk.cu)
// with three deliberate, documented choices that make it deterministic:
//   * k-nearest lists are ordered by the total order (d2, index) with
//     d2 = fma(dy, dy, dx*dx); the reference orders ties by visit order
//     (voro_optimizer.cu:149-151), which depends on atomic ranks (:176).
//   * gradient contributions are summed order-independently (the reference
//     uses FP32 atomics in arbitrary order, :250-253): in FP32 mode each term
//     c becomes round(c * 2^g) in 64-bit integers (g from the target count and
//     domain extent, see gradShift), and the sum is read back as sum * 2^-g;
//     in FP64 mode terms are summed in double in target order.
//   * the Adam bias corrections 1 - beta^t are evaluated in double by
//     exponentiation by squaring and rounded once (the reference uses powf,
//     :287-288, whose rounding is library specific).
// Every arithmetic expression is written out with explicit fma() where one is
// wanted; the file is compiled with -ffp-contract=off so that no other fused
// operation appears.  Templated on the real type: float gives the FP32
// semantics the GPU must reproduce bit for bit (k-NN lists, per-target loss
// terms, clipped cells); double gives the FP64 error oracle.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Stateless RNG (voro_optimizer.cu:69-78, poisson_disk.cu:47-58)
// ---------------------------------------------------------------------------
inline uint32_t pcg_hash(uint32_t v)
{
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word  = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// Target subsampling draw, voro_optimizer.cu:75-78.
inline float rand01_loss(uint32_t index, uint32_t iter, uint32_t seed)
{
    uint32_t h = pcg_hash(index ^ pcg_hash(iter * 2654435761u ^ pcg_hash(seed)));
    return float(h & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}

// Poisson dart draw, poisson_disk.cu:52-58.
inline float rand01_poisson(uint32_t cell, uint32_t round, uint32_t seed, uint32_t stream)
{
    uint32_t h = pcg_hash(cell ^ pcg_hash(round * 2654435761u ^ pcg_hash(seed + stream * 0x9E3779B9u)));
    return float(h & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}

template <class R> struct V2 { R x, y; };

template <class R> inline R fma_(R a, R b, R c);
template <> inline float fma_<float>(float a, float b, float c) { return std::fmaf(a, b, c); }
template <> inline double fma_<double>(double a, double b, double c) { return std::fma(a, b, c); }

template <class R> inline R sqrt_(R a);
template <> inline float sqrt_<float>(float a) { return std::sqrt(a); }
template <> inline double sqrt_<double>(double a) { return std::sqrt(a); }

// Squared distance as both the oracle and the GPU define it.
template <class R> inline R dist2(V2<R> p, V2<R> q)
{
    R dx = p.x - q.x, dy = p.y - q.y;
    return fma_<R>(dy, dy, dx * dx);
}

template <class R> inline R dot(V2<R> a, V2<R> b) { return fma_<R>(a.y, b.y, a.x * b.x); }

// Exponentiation by squaring in double (same sequence of products on every
// platform, so the result is bit-reproducible).
inline double powi(double b, int e)
{
    double r = 1.0;
    while (e > 0) {
        if (e & 1) r = r * b;
        b = b * b;
        e >>= 1;
    }
    return r;
}

// ---------------------------------------------------------------------------
// Uniform hash grid (voro_optimizer.cu:115-125, 168-187), built as a stable
// counting sort so the within-cell order is by site index.
// ---------------------------------------------------------------------------
template <class R> struct Grid {
    int              w = 1, h = 1;
    R                cell = 1, inv = 1;
    R                ext = 1; // max domain extent, scales the stop margin
    std::vector<int> start;   // w*h + 1
    std::vector<int> idx;     // site index, sorted by cell
    std::vector<V2<R>> pos;
};

template <class R> inline int cellCoord(R v, R inv, int n)
{
    R   f = v * inv;
    int c = (f >= R(0)) ? (f < R(2147483520.0) ? int(f) : n - 1) : 0; // int() truncates; NaN -> 0
    return std::min(std::max(c, 0), n - 1);
}

template <class R>
Grid<R> buildGrid(const V2<R> *sites, int n, int w, int h, R cell, R ext)
{
    Grid<R> g;
    g.w = w, g.h = h, g.cell = cell, g.inv = R(1) / cell, g.ext = ext;
    std::vector<int> cellOf(n, -1);
    g.start.assign(size_t(w) * h + 1, 0);
    for (int i = 0; i < n; ++i) {
        if (!(sites[i].x >= R(0))) continue; // empty slot (x < 0)
        int c = cellCoord(sites[i].y, g.inv, h) * w + cellCoord(sites[i].x, g.inv, w);
        cellOf[i] = c;
        g.start[c + 1]++;
    }
    for (size_t c = 0; c + 1 < g.start.size(); ++c) g.start[c + 1] += g.start[c];
    g.idx.resize(g.start.back());
    g.pos.resize(g.start.back());
    std::vector<int> fill(g.start.begin(), g.start.end() - 1);
    for (int i = 0; i < n; ++i) {
        if (cellOf[i] < 0) continue;
        int s = fill[cellOf[i]]++;
        g.idx[s] = i;
        g.pos[s] = sites[i];
    }
    return g;
}

// Sorted top-K buffer under the total order (d2, index).
template <class R> struct TopK {
    int              K;
    std::vector<R>   d2;
    std::vector<int> idx;
    explicit TopK(int k) : K(k), d2(k, std::numeric_limits<R>::max()), idx(k, -1) {}
    static bool less(R a, int ia, R b, int ib) { return a < b || (a == b && unsigned(ia) < unsigned(ib)); }
    void insert(R d, int i)
    {
        if (!less(d, i, d2[K - 1], idx[K - 1])) return;
        int k = K - 1;
        while (k > 0 && less(d, i, d2[k - 1], idx[k - 1])) {
            d2[k]  = d2[k - 1];
            idx[k] = idx[k - 1];
            --k;
        }
        d2[k]  = d;
        idx[k] = i;
    }
};

// Square-ring search outward from q's cell (voro_optimizer.cu:130-163).  The
// stop test is made conservative by a relative margin so that the result
// always equals the brute-force list under the (d2, index) order, whatever the
// grid cell size.
template <class R> void knnGrid(const Grid<R> &g, V2<R> q, TopK<R> &top)
{
    const R rel    = sizeof(R) == 4 ? R(1e-6) : R(1e-13);
    R       extent = std::max(g.ext, std::max(std::fabs(q.x), std::fabs(q.y)));
    const R margin = rel * extent;
    int     cx = cellCoord(q.x, g.inv, g.w), cy = cellCoord(q.y, g.inv, g.h);
    for (int ring = 0;; ++ring) {
        int x0 = cx - ring, x1 = cx + ring, y0 = cy - ring, y1 = cy + ring;
        for (int y = std::max(y0, 0); y <= std::min(y1, g.h - 1); ++y) {
            int stride = (y == y0 || y == y1) ? 1 : std::max(x1 - x0, 1);
            for (int x = x0; x <= x1; x += stride) {
                if (x < 0 || x >= g.w) continue;
                int c = y * g.w + x;
                for (int s = g.start[c]; s < g.start[c + 1]; ++s) top.insert(dist2(g.pos[s], q), g.idx[s]);
            }
        }
        if (x0 <= 0 && y0 <= 0 && x1 >= g.w - 1 && y1 >= g.h - 1) break;
        R safe = std::min(std::min(q.x - R(x0) * g.cell, R(x1 + 1) * g.cell - q.x),
                          std::min(q.y - R(y0) * g.cell, R(y1 + 1) * g.cell - q.y));
        safe   = std::max(safe - margin, R(0));
        if (top.idx[top.K - 1] >= 0 && top.d2[top.K - 1] * (R(1) + rel) < safe * safe) break;
    }
}

template <class R> void knnBrute(const V2<R> *sites, int n, V2<R> q, TopK<R> &top)
{
    for (int i = 0; i < n; ++i)
        if (sites[i].x >= R(0)) top.insert(dist2(sites[i], q), i);
}

// ---------------------------------------------------------------------------
// VoroLoss for one target and its gradient (voro_optimizer.cu:199-256).
// ---------------------------------------------------------------------------
template <class R> struct TargetTerm {
    int   i0 = -1, j = -1;
    R     loss = 0;
    V2<R> g0{0, 0}, gj{0, 0};
};

template <class R>
bool lossTerm(const V2<R> *sites, const TopK<R> &top, V2<R> p, bool wantGrad, TargetTerm<R> &out)
{
    const int K  = top.K;
    int       i0 = top.idx[0];
    if (i0 < 0) return false;
    V2<R> c0 = sites[i0];
    V2<R> u{p.x - c0.x, p.y - c0.y};
    R     best = std::numeric_limits<R>::max(), bestD = 0;
    int   bestJ = -1;
    const R tiny = sizeof(R) == 4 ? R(1e-12f) : R(1e-12);
    for (int k = 1; k < K; ++k) {
        int j = top.idx[k];
        if (j < 0) break;
        V2<R> e{sites[j].x - c0.x, sites[j].y - c0.y};
        R     l = sqrt_<R>(dot(e, e));
        if (l < tiny) continue;
        R q = dot(u, e) / l;
        R d = fma_<R>(R(-0.5), l, q);
        R dd = d * d;
        if (dd < best) { best = dd; bestD = d; bestJ = j; }
    }
    if (bestJ < 0) return false;
    out.i0 = i0, out.j = bestJ, out.loss = best;
    if (wantGrad) {
        V2<R> e{sites[bestJ].x - c0.x, sites[bestJ].y - c0.y};
        R     l  = sqrt_<R>(dot(e, e));
        R     rl = R(1) / l;
        V2<R> n{rl * e.x, rl * e.y};
        R     un = dot(u, n);
        R     tx = fma_<R>(-un, n.x, u.x), ty = fma_<R>(-un, n.y, u.y);
        V2<R> dcj{fma_<R>(rl, tx, R(-0.5) * n.x), fma_<R>(rl, ty, R(-0.5) * n.y)};
        R     s = R(2) * bestD;
        out.gj  = {s * dcj.x, s * dcj.y};
        out.g0  = {-(s * (n.x + dcj.x)), -(s * (n.y + dcj.y))};
    }
    return true;
}

// Whether target t (slot index `slot`) takes part in iteration `iter`
// (voro_optimizer.cu:214-220).
template <class R>
bool takeTarget(V2<R> p, uint32_t slot, const double *box, float fraction, uint32_t iter, uint32_t seed)
{
    bool take = p.x >= R(0);
    if (take && box) take = p.x >= R(box[0]) && p.x <= R(box[2]) && p.y >= R(box[1]) && p.y <= R(box[3]);
    if (take && fraction < 1.0f) take = rand01_loss(slot, iter, seed) < fraction;
    return take;
}

// ceil(log2(x)) for x > 0, exact.
inline int ceilLog2(double x)
{
    int    e;
    double m = std::frexp(x, &e);
    return m == 0.5 ? e - 1 : e;
}

// Fixed-point shift of the FP32 gradient sums: a site collects at most T
// terms of magnitude ~ few * extent, so the sums stay below 2^55; each scaled
// term is clamped to +-2^51.
inline int gradShift(int liveTargets, float extent)
{
    int g = 53 - ceilLog2(double(std::max(liveTargets, 1))) - ceilLog2(double(std::max(extent, 1e-30f)));
    return std::min(std::max(g, 3), 47);
}

inline int64_t quantizeTerm(float c, int shift)
{
    double x = std::ldexp(double(c), shift);
    x        = std::fmin(std::fmax(x, -2251799813685248.0), 2251799813685248.0); // +-2^51
    return std::llrint(x);
}

struct LossOut {
    double  lossSum = 0;
    int64_t count   = 0;
};

// One full VoroLoss pass: per-target terms in parallel, then a serial sum in
// target order (deterministic for any thread count).
template <class R>
LossOut lossPass(const V2<R> *sites, int S, const Grid<R> &g, const V2<R> *targets, int T,
                 const int32_t *slotOf, int K, const double *box, float fraction, uint32_t iter, uint32_t seed,
                 float extent, double *grad /* 2*S or null */, int32_t *c0Out, int32_t *bjOut, R *lossOut,
                 R *termGrad /* 4*T */)
{
    std::vector<TargetTerm<R>> terms(T);
    std::vector<uint8_t>       used(T, 0);
    bool                       wantGrad = grad != nullptr || termGrad != nullptr;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int t = 0; t < T; ++t) {
        V2<R>    p    = targets[t];
        uint32_t slot = slotOf ? uint32_t(slotOf[t]) : uint32_t(t);
        if (!takeTarget(p, slot, box, fraction, iter, seed)) continue;
        TopK<R> top(K);
        knnGrid(g, p, top);
        used[t] = lossTerm(sites, top, p, wantGrad, terms[t]) ? 1 : 0;
    }
    LossOut out;
    if (grad) std::fill(grad, grad + 2 * size_t(S), 0.0);
    const bool           fixedPoint = sizeof(R) == 4;
    int                  live       = 0;
    for (int t = 0; t < T; ++t) live += targets[t].x >= R(0);
    const int            shift = gradShift(live, extent);
    std::vector<int64_t> isum(fixedPoint && grad ? 2 * size_t(S) : 0, 0);
    for (int t = 0; t < T; ++t) {
        if (c0Out) c0Out[t] = used[t] ? terms[t].i0 : -1;
        if (bjOut) bjOut[t] = used[t] ? terms[t].j : -1;
        if (lossOut) lossOut[t] = used[t] ? terms[t].loss : R(0);
        if (termGrad) {
            termGrad[4 * size_t(t) + 0] = used[t] ? terms[t].g0.x : R(0);
            termGrad[4 * size_t(t) + 1] = used[t] ? terms[t].g0.y : R(0);
            termGrad[4 * size_t(t) + 2] = used[t] ? terms[t].gj.x : R(0);
            termGrad[4 * size_t(t) + 3] = used[t] ? terms[t].gj.y : R(0);
        }
        if (!used[t]) continue;
        out.lossSum += double(terms[t].loss);
        out.count += 1;
        if (grad && fixedPoint) {
            isum[2 * size_t(terms[t].j) + 0] += quantizeTerm(float(terms[t].gj.x), shift);
            isum[2 * size_t(terms[t].j) + 1] += quantizeTerm(float(terms[t].gj.y), shift);
            isum[2 * size_t(terms[t].i0) + 0] += quantizeTerm(float(terms[t].g0.x), shift);
            isum[2 * size_t(terms[t].i0) + 1] += quantizeTerm(float(terms[t].g0.y), shift);
        } else if (grad) {
            grad[2 * size_t(terms[t].j) + 0] += double(terms[t].gj.x);
            grad[2 * size_t(terms[t].j) + 1] += double(terms[t].gj.y);
            grad[2 * size_t(terms[t].i0) + 0] += double(terms[t].g0.x);
            grad[2 * size_t(terms[t].i0) + 1] += double(terms[t].g0.y);
        }
    }
    if (grad && fixedPoint)
        for (size_t k = 0; k < isum.size(); ++k) grad[k] = std::ldexp(double(isum[k]), -shift);
    return out;
}

} // namespace

// ===========================================================================
// C ABI used by the Python tests (ctypes).  `real` selects the arithmetic:
// 0 = FP32 semantics, 1 = FP64.  Point arrays are interleaved (x, y) pairs of
// that type.
// ===========================================================================
#define VO_API extern "C"

struct vo_sched {
    double lr, gamma, beta1, beta2, eps; // rounded to the real type before use
    int    milestones[4];
    double width, height;
};

VO_API uint32_t vo_pcg_hash(uint32_t v) { return pcg_hash(v); }
VO_API float    vo_rand01_loss(uint32_t index, uint32_t iter, uint32_t seed) { return rand01_loss(index, iter, seed); }
VO_API float    vo_rand01_poisson(uint32_t cell, uint32_t round, uint32_t seed, uint32_t stream)
{
    return rand01_poisson(cell, round, seed, stream);
}

VO_API int vo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

VO_API void vo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

template <class R>
static int knnImpl(const void *sitesV, int S, const void *queriesV, int Q, int K, double W, double H, double cell,
                   int brute, int32_t *outIdx, void *outD2V)
{
    auto   *sites   = static_cast<const V2<R> *>(sitesV);
    auto   *queries = static_cast<const V2<R> *>(queriesV);
    R      *outD2   = static_cast<R *>(outD2V);
    int     gw = std::max(1, int(std::ceil(W / cell))), gh = std::max(1, int(std::ceil(H / cell)));
    Grid<R> g;
    if (!brute) g = buildGrid(sites, S, gw, gh, R(cell), R(std::max(W, H)));
#pragma omp parallel for schedule(dynamic, 256)
    for (int q = 0; q < Q; ++q) {
        TopK<R> top(K);
        if (brute) knnBrute(sites, S, queries[q], top);
        else knnGrid(g, queries[q], top);
        for (int k = 0; k < K; ++k) {
            outIdx[size_t(q) * K + k] = top.idx[k];
            if (outD2) outD2[size_t(q) * K + k] = top.d2[k];
        }
    }
    return 0;
}

// k nearest sites of each query (ring search over a grid of `cell`, or brute
// force when brute != 0).  Missing entries get index -1 and d2 = max.
VO_API int vo_knn(int real, const void *sites, int S, const void *queries, int Q, int K, double W, double H, double cell,
           int brute, int32_t *outIdx, void *outD2)
{
    if (K <= 0 || S < 0 || Q < 0 || cell <= 0) return 1;
    return real ? knnImpl<double>(sites, S, queries, Q, K, W, H, cell, brute, outIdx, outD2)
                : knnImpl<float>(sites, S, queries, Q, K, W, H, cell, brute, outIdx, outD2);
}

template <class R>
static int lossImpl(const void *sitesV, int S, const void *targetsV, int T, const int32_t *slotOf, int K, double W,
                    double H, double cell, float fraction, uint32_t seed, uint32_t iter, const double *box,
                    double *grad, double *lossSum, int64_t *count, int32_t *c0Out, int32_t *bjOut, void *lossT,
                    void *termGrad)
{
    auto   *sites   = static_cast<const V2<R> *>(sitesV);
    auto   *targets = static_cast<const V2<R> *>(targetsV);
    int     gw = std::max(1, int(std::ceil(W / cell))), gh = std::max(1, int(std::ceil(H / cell)));
    Grid<R> g = buildGrid(sites, S, gw, gh, R(cell), R(std::max(W, H)));
    LossOut o = lossPass(sites, S, g, targets, T, slotOf, K, box, fraction, iter, seed, float(std::max(W, H)), grad,
                         c0Out, bjOut, static_cast<R *>(lossT), static_cast<R *>(termGrad));
    if (lossSum) *lossSum = o.lossSum;
    if (count) *count = o.count;
    return 0;
}

// One VoroLoss pass at step `iter` (1-based, used by the subsampling hash).
// grad (2*S doubles) receives the summed, un-normalised gradient.  Optional
// per-target outputs: nearest site, bisector partner, loss term, and the four
// gradient components (g0.x, g0.y, gj.x, gj.y) in the real type.
VO_API int vo_loss_grad(int real, const void *sites, int S, const void *targets, int T, const int32_t *slotOf, int K,
                 double W, double H, double cell, float fraction, uint32_t seed, uint32_t iter, const double *box,
                 double *grad, double *lossSum, int64_t *count, int32_t *c0Out, int32_t *bjOut, void *lossT,
                 void *termGrad)
{
    if (K < 2 || cell <= 0) return 1;
    return real ? lossImpl<double>(sites, S, targets, T, slotOf, K, W, H, cell, fraction, seed, iter, box, grad,
                                   lossSum, count, c0Out, bjOut, lossT, termGrad)
                : lossImpl<float>(sites, S, targets, T, slotOf, K, W, H, cell, fraction, seed, iter, box, grad,
                                  lossSum, count, c0Out, bjOut, lossT, termGrad);
}

// Adam + MultiStepLR + clamp for every movable live site (voro_optimizer.cu:268-301).
template <class R>
static void adamImpl(V2<R> *sites, V2<R> *m, V2<R> *v, const double *grad, const uint8_t *movable, int S, int t,
                     int64_t count, const vo_sched &sc)
{
    R lr = R(sc.lr);
    for (int k = 0; k < 4; ++k)
        if (t - 1 >= sc.milestones[k]) lr = lr * R(sc.gamma);
    const R b1 = R(sc.beta1), b2 = R(sc.beta2), eps = R(sc.eps);
    const R bc1 = R(1.0 - powi(double(b1), t));
    const R bc2 = R(1.0 - powi(double(b2), t));
    const R inv = R(1) / R(std::max<int64_t>(count, 1));
    const R ob1 = R(1) - b1, ob2 = R(1) - b2;
    const R W = R(sc.width), H = R(sc.height);
    for (int i = 0; i < S; ++i) {
        if (!(sites[i].x >= R(0))) continue;
        if (movable && !movable[i]) continue;
        V2<R> g{inv * R(grad[2 * size_t(i)]), inv * R(grad[2 * size_t(i) + 1])};
        V2<R> mi{b1 * m[i].x + ob1 * g.x, b1 * m[i].y + ob1 * g.y};
        V2<R> vi{b2 * v[i].x + ob2 * (g.x * g.x), b2 * v[i].y + ob2 * (g.y * g.y)};
        m[i]    = mi;
        v[i]    = vi;
        V2<R> p = sites[i];
        p.x     = p.x - (lr * (mi.x / bc1)) / (sqrt_<R>(vi.x / bc2) + eps);
        p.y     = p.y - (lr * (mi.y / bc1)) / (sqrt_<R>(vi.y / bc2) + eps);
        p.x     = std::min(std::max(p.x, R(0)), W);
        p.y     = std::min(std::max(p.y, R(0)), H);
        sites[i] = p;
    }
}

// Applies one Adam step at step number t with the given (un-normalised)
// gradient sums and target count.
VO_API int vo_adam(int real, void *sites, void *m, void *v, const double *grad, const uint8_t *movable, int S, int t,
            int64_t count, const vo_sched *sc)
{
    if (!sc || t < 1) return 1;
    if (real)
        adamImpl<double>(static_cast<V2<double> *>(sites), static_cast<V2<double> *>(m),
                         static_cast<V2<double> *>(v), grad, movable, S, t, count, *sc);
    else
        adamImpl<float>(static_cast<V2<float> *>(sites), static_cast<V2<float> *>(m), static_cast<V2<float> *>(v),
                        grad, movable, S, t, count, *sc);
    return 0;
}

template <class R>
static int fitImpl(void *sitesV, void *mV, void *vV, const uint8_t *movable, int S, const void *targetsV, int T,
                   const int32_t *slotOf, int K, double cell, float fraction, uint32_t seed, const double *box,
                   int tStart, int iters, const vo_sched *sc, double *lossHist, int64_t *countHist)
{
    auto *sites   = static_cast<V2<R> *>(sitesV);
    auto *targets = static_cast<const V2<R> *>(targetsV);
    auto *m       = static_cast<V2<R> *>(mV);
    auto *v       = static_cast<V2<R> *>(vV);
    int   gw = std::max(1, int(std::ceil(sc->width / cell))), gh = std::max(1, int(std::ceil(sc->height / cell)));
    std::vector<double> grad(2 * size_t(S));
    for (int it = 0; it < iters; ++it) {
        int     t = tStart + it;
        Grid<R> g = buildGrid(sites, S, gw, gh, R(cell), R(std::max(sc->width, sc->height)));
        LossOut o = lossPass(sites, S, g, targets, T, slotOf, K, box, fraction, uint32_t(t), seed,
                             float(std::max(sc->width, sc->height)), grad.data(), nullptr, nullptr,
                             static_cast<R *>(nullptr), static_cast<R *>(nullptr));
        if (lossHist) lossHist[it] = o.lossSum;
        if (countHist) countHist[it] = o.count;
        adamImpl<R>(sites, m, v, grad.data(), movable, S, t, o.count, *sc);
    }
    return 0;
}

// `iters` optimizer iterations with step numbers tStart, tStart+1, ...  Sites,
// m and v are updated in place.  lossHist/countHist receive the per-iteration
// loss sum and target count.
VO_API int vo_fit(int real, void *sites, void *m, void *v, const uint8_t *movable, int S, const void *targets, int T,
           const int32_t *slotOf, int K, double cell, float fraction, uint32_t seed, const double *box, int tStart,
           int iters, const vo_sched *sc, double *lossHist, int64_t *countHist)
{
    if (K < 2 || cell <= 0 || !sc || tStart < 1) return 1;
    return real ? fitImpl<double>(sites, m, v, movable, S, targets, T, slotOf, K, cell, fraction, seed, box, tStart,
                                  iters, sc, lossHist, countHist)
                : fitImpl<float>(sites, m, v, movable, S, targets, T, slotOf, K, cell, fraction, seed, box, tStart,
                                 iters, sc, lossHist, countHist);
}

// ---------------------------------------------------------------------------
// Cell clipping (voro_optimizer.cu:312-437), FP64 predicates.
// ---------------------------------------------------------------------------
static const int kMaxCellVerts = 24;

struct CellPoly {
    double vx[kMaxCellVerts], vy[kMaxCellVerts];
    int    e[kMaxCellVerts];
    int    n = 0;
};

// Keep <x, r> <= |r|^2/2 (the side of the generator at the origin).
static bool clipByBisector(CellPoly &poly, double rx, double ry, int label)
{
    double c      = 0.5 * (rx * rx + ry * ry);
    bool   anyOut = false;
    for (int k = 0; k < poly.n; ++k) anyOut |= poly.vx[k] * rx + poly.vy[k] * ry - c > 0.0;
    if (!anyOut) return true;
    CellPoly out;
    bool     ok   = true;
    auto     push = [&](double x, double y, int e) {
        if (out.n < kMaxCellVerts) {
            out.vx[out.n] = x, out.vy[out.n] = y, out.e[out.n] = e;
            ++out.n;
        } else
            ok = false;
    };
    for (int k = 0; k < poly.n; ++k) {
        int    k1 = (k + 1) % poly.n;
        double ax = poly.vx[k], ay = poly.vy[k], bx = poly.vx[k1], by = poly.vy[k1];
        double fa = ax * rx + ay * ry - c;
        double fb = bx * rx + by * ry - c;
        bool   inA = fa <= 0.0, inB = fb <= 0.0;
        if (inA) push(ax, ay, poly.e[k]);
        if (inA != inB) {
            double t = fa / (fa - fb);
            push(ax + t * (bx - ax), ay + t * (by - ay), inA ? label : poly.e[k]);
        }
    }
    poly = out;
    return ok;
}

static double maxRadius2(const CellPoly &p)
{
    double r2 = 0.0;
    for (int k = 0; k < p.n; ++k) r2 = std::fmax(r2, p.vx[k] * p.vx[k] + p.vy[k] * p.vy[k]);
    return r2;
}

static bool occupiedAt(const uint64_t *masks, int tilesX, int tilesY, float invVoxel, float px, float py)
{
    if (!masks) return false;
    int i = int(std::floor(px * invVoxel)), j = int(std::floor(py * invVoxel));
    if (i < 0 || j < 0) return false;
    int tx = i >> 3, ty = j >> 3;
    if (tx >= tilesX || ty >= tilesY) return false;
    return (masks[size_t(ty) * tilesX + tx] >> (((j & 7) << 3) | (i & 7))) & 1ull;
}

template <class R>
static int clipImpl(const void *sitesV, int S, double W, double H, double cell, int Kc, const uint64_t *masks,
                    int tilesX, int tilesY, float voxel, float *outVerts, int32_t *outLabels, int32_t *outCount,
                    uint8_t *outInside, int32_t *flags)
{
    auto *sites = static_cast<const V2<R> *>(sitesV);
    // Neighbours are ordered by FP64 distance in both modes (the clip is FP64
    // throughout), so FP32 and FP64 sites with equal values clip identically.
    std::vector<V2<double>> sd(S);
    for (int i = 0; i < S; ++i) sd[i] = V2<double>{double(sites[i].x), double(sites[i].y)};
    int          gw = std::max(1, int(std::ceil(W / cell))), gh = std::max(1, int(std::ceil(H / cell)));
    Grid<double> g  = buildGrid(sd.data(), S, gw, gh, cell, std::max(W, H));
    float   invVoxel = voxel > 0 ? 1.0f / voxel : 0.0f;
    int     unsafeN = 0, overflowN = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : unsafeN, overflowN)
    for (int i = 0; i < S; ++i) {
        V2<R> ci = sites[i];
        if (!(ci.x >= R(0))) {
            outCount[i]  = 0;
            outInside[i] = 0;
            continue;
        }
        TopK<double> top(Kc);
        knnGrid(g, sd[i], top);
        CellPoly poly;
        double   x0 = -double(ci.x), y0 = -double(ci.y), x1 = W - double(ci.x), y1 = H - double(ci.y);
        poly.vx[0] = x0, poly.vy[0] = y0, poly.e[0] = -1;
        poly.vx[1] = x1, poly.vy[1] = y0, poly.e[1] = -2;
        poly.vx[2] = x1, poly.vy[2] = y1, poly.e[2] = -3;
        poly.vx[3] = x0, poly.vy[3] = y1, poly.e[3] = -4;
        poly.n     = 4;
        bool   ok = true, certified = false;
        double r2max = maxRadius2(poly);
        for (int k = 0; k < Kc; ++k) {
            int j = top.idx[k];
            if (j < 0) { certified = true; break; }
            if (j == i) continue;
            if (double(top.d2[k]) > 4.0 * r2max) { certified = true; break; }
            ok &= clipByBisector(poly, double(sites[j].x) - double(ci.x), double(sites[j].y) - double(ci.y), j);
            r2max = maxRadius2(poly);
        }
        if (!certified && double(top.d2[Kc - 1]) > 4.0 * r2max) certified = true;
        if (!certified) {
            ++unsafeN;
            for (int j = 0; j < S; ++j) {
                if (j == i || !(sites[j].x >= R(0))) continue;
                bool seen = false;
                for (int k = 0; k < Kc; ++k) seen |= top.idx[k] == j;
                if (seen) continue;
                double rx = double(sites[j].x) - double(ci.x), ry = double(sites[j].y) - double(ci.y);
                if (rx * rx + ry * ry > 4.0 * r2max) continue;
                ok &= clipByBisector(poly, rx, ry, j);
                r2max = maxRadius2(poly);
            }
        }
        if (!ok) ++overflowN;
        double area = 0.0, cx = 0.0, cy = 0.0;
        for (int k = 0; k < poly.n; ++k) {
            int    k1 = (k + 1) % poly.n;
            double cr = poly.vx[k] * poly.vy[k1] - poly.vx[k1] * poly.vy[k];
            area += cr;
            cx += (poly.vx[k] + poly.vx[k1]) * cr;
            cy += (poly.vy[k] + poly.vy[k1]) * cr;
        }
        float cenx = float(ci.x), ceny = float(ci.y);
        if (std::fabs(area) > 1e-12) {
            cenx = cenx + float(cx / (3.0 * area));
            ceny = ceny + float(cy / (3.0 * area));
        }
        outInside[i] = occupiedAt(masks, tilesX, tilesY, invVoxel, cenx, ceny) ? 1 : 0;
        outCount[i]  = poly.n;
        for (int k = 0; k < poly.n; ++k) {
            outVerts[(size_t(i) * kMaxCellVerts + k) * 2 + 0] = float(double(ci.x) + poly.vx[k]);
            outVerts[(size_t(i) * kMaxCellVerts + k) * 2 + 1] = float(double(ci.y) + poly.vy[k]);
            outLabels[size_t(i) * kMaxCellVerts + k]          = poly.e[k];
        }
    }
    if (flags) flags[0] = unsafeN, flags[1] = overflowN;
    return 0;
}

// Clipped Voronoi cell of every live site against [0,W]x[0,H]: up to 24
// vertices (x, y floats) and edge labels (>= 0 neighbour index, -1..-4 box
// sides) per site, vertex counts, centroid occupancy, and the
// {unsafe, overflow} counters.
VO_API int vo_clip(int real, const void *sites, int S, double W, double H, double cell, int Kc, const uint64_t *masks,
            int tilesX, int tilesY, float voxel, float *outVerts, int32_t *outLabels, int32_t *outCount,
            uint8_t *outInside, int32_t *flags)
{
    if (Kc < 1 || cell <= 0) return 1;
    return real ? clipImpl<double>(sites, S, W, H, cell, Kc, masks, tilesX, tilesY, voxel, outVerts, outLabels,
                                   outCount, outInside, flags)
                : clipImpl<float>(sites, S, W, H, cell, Kc, masks, tilesX, tilesY, voxel, outVerts, outLabels,
                                  outCount, outInside, flags);
}

// Boundary assembly (voro_optimizer.cu:806-854): edges of inside cells facing
// an outside cell or the box; vertices keyed by the sorted label triple.
// Returns the vertex and edge counts; writes at most maxV / maxE entries.
VO_API int vo_boundary(int n, const float *verts, const int32_t *labels, const int32_t *counts, const uint8_t *inside,
                float *outV, int maxV, int32_t *outE, int maxE, int32_t *nV, int32_t *nE, int32_t *watertight,
                int32_t *loops)
{
    struct Key {
        int64_t a, b, c;
        bool    operator<(const Key &o) const
        {
            return a != o.a ? a < o.a : (b != o.b ? b < o.b : c < o.c);
        }
    };
    auto key = [](int a, int b, int c) {
        int64_t v[3] = {int64_t(a) + 4, int64_t(b) + 4, int64_t(c) + 4};
        std::sort(v, v + 3);
        return Key{v[0], v[1], v[2]};
    };
    std::map<Key, int>               vertexOf;
    std::vector<float>               vx;
    std::vector<std::pair<int, int>> edges;
    auto vertex = [&](Key k, float x, float y) {
        auto it = vertexOf.find(k);
        if (it != vertexOf.end()) return it->second;
        int id = int(vx.size() / 2);
        vertexOf.emplace(k, id);
        vx.push_back(x), vx.push_back(y);
        return id;
    };
    for (int i = 0; i < n; ++i) {
        if (!inside[i]) continue;
        int c = counts[i];
        const int32_t *lab = labels + size_t(i) * kMaxCellVerts;
        const float   *pos = verts + size_t(i) * kMaxCellVerts * 2;
        for (int k = 0; k < c; ++k) {
            int l = lab[k];
            if (l >= 0 && inside[l]) continue;
            int prev = lab[(k + c - 1) % c], next = lab[(k + 1) % c];
            int a = vertex(key(i, prev, l), pos[2 * k], pos[2 * k + 1]);
            int b = vertex(key(i, l, next), pos[2 * ((k + 1) % c)], pos[2 * ((k + 1) % c) + 1]);
            edges.push_back({a, b});
        }
    }
    int              V = int(vx.size() / 2);
    std::vector<int> nxt(V, -1), inDeg(V, 0);
    bool             wt = true;
    for (auto [a, b] : edges) {
        if (nxt[a] != -1) wt = false;
        nxt[a] = b;
        ++inDeg[b];
    }
    for (int v = 0; v < V; ++v)
        if (nxt[v] == -1 || inDeg[v] != 1) wt = false;
    int nl = 0;
    if (wt) {
        std::vector<uint8_t> seen(V, 0);
        for (int v = 0; v < V; ++v) {
            if (seen[v]) continue;
            ++nl;
            for (int w = v; !seen[w]; w = nxt[w]) seen[w] = 1;
        }
    }
    for (int v = 0; v < std::min(V, maxV); ++v) outV[2 * v] = vx[2 * v], outV[2 * v + 1] = vx[2 * v + 1];
    for (int e = 0; e < std::min(int(edges.size()), maxE); ++e)
        outE[2 * e] = edges[e].first, outE[2 * e + 1] = edges[e].second;
    *nV         = V;
    *nE         = int(edges.size());
    *watertight = wt ? 1 : 0;
    *loops      = nl;
    return 0;
}

// ---------------------------------------------------------------------------
// Poisson-disk dart throwing (poisson_disk.cu:74-177, 179-199).
// ---------------------------------------------------------------------------
VO_API void vo_poisson_dims(float W, float H, float radius, int32_t *gw, int32_t *gh, float *cellSize)
{
    float cs  = radius / std::sqrt(2.0f);
    *cellSize = cs;
    *gw       = int(std::ceil(W / cs));
    *gh       = int(std::ceil(H / cs));
}

// Clears cells [cx0,cx1) x [cy0,cy1) of `grid` (gw*gh interleaved floats) and
// runs `trials` rounds of the nine phases, round numbers roundBase + r.
VO_API int vo_poisson(float W, float H, float radius, int trials, uint32_t seed, int cx0, int cy0, int cx1, int cy1,
               uint32_t roundBase, float *grid)
{
    float cs;
    int   gw, gh;
    vo_poisson_dims(W, H, radius, &gw, &gh, &cs);
    int x0 = std::max(cx0, 0), y0 = std::max(cy0, 0), x1 = std::min(cx1, gw), y1 = std::min(cy1, gh);
    if (x0 >= x1 || y0 >= y1) return 0;
    for (int y = y0; y < y1; ++y)
        for (int x = x0; x < x1; ++x) grid[2 * (size_t(y) * gw + x)] = -1.0f, grid[2 * (size_t(y) * gw + x) + 1] = -1.0f;
    float r2 = radius * radius;
    for (int round = 0; round < trials; ++round) {
        uint32_t rnd = roundBase + uint32_t(round);
        for (int py = 0; py < 3; ++py)
            for (int px = 0; px < 3; ++px)
                for (int cy = y0 + py; cy < y1; cy += 3)
                    for (int cx = x0 + px; cx < x1; cx += 3) {
                        int cell = cy * gw + cx;
                        if (grid[2 * size_t(cell)] >= 0.0f) continue;
                        float x = (float(cx) + rand01_poisson(uint32_t(cell), rnd, seed, 0)) * cs;
                        float y = (float(cy) + rand01_poisson(uint32_t(cell), rnd, seed, 1)) * cs;
                        if (x >= W || y >= H) continue;
                        bool ok = true;
                        for (int ny = std::max(cy - 2, 0); ok && ny <= std::min(cy + 2, gh - 1); ++ny)
                            for (int nx = std::max(cx - 2, 0); nx <= std::min(cx + 2, gw - 1); ++nx) {
                                float qx = grid[2 * (size_t(ny) * gw + nx)], qy = grid[2 * (size_t(ny) * gw + nx) + 1];
                                if (qx < 0.0f) continue;
                                float dx = qx - x, dy = qy - y;
                                if (std::fmaf(dy, dy, dx * dx) < r2) { ok = false; break; }
                            }
                        if (!ok) continue;
                        grid[2 * size_t(cell)]     = x;
                        grid[2 * size_t(cell) + 1] = y;
                    }
    }
    return 0;
}

// No two samples touching the region (expanded by two cells) closer than r.
VO_API int vo_poisson_verify(float W, float H, float radius, const float *grid, int cx0, int cy0, int cx1, int cy1)
{
    float cs;
    int   gw, gh;
    vo_poisson_dims(W, H, radius, &gw, &gh, &cs);
    float r2  = radius * radius;
    int   vx0 = std::max(cx0 - 2, 0), vx1 = std::min(cx1 + 2, gw);
    int   vy0 = std::max(cy0 - 2, 0), vy1 = std::min(cy1 + 2, gh);
    for (int cy = vy0; cy < vy1; ++cy)
        for (int cx = vx0; cx < vx1; ++cx) {
            float ax = grid[2 * (size_t(cy) * gw + cx)], ay = grid[2 * (size_t(cy) * gw + cx) + 1];
            if (ax < 0) continue;
            for (int ny = std::max(cy - 2, 0); ny <= std::min(cy + 2, gh - 1); ++ny)
                for (int nx = std::max(cx - 2, 0); nx <= std::min(cx + 2, gw - 1); ++nx) {
                    if (nx == cx && ny == cy) continue;
                    float bx = grid[2 * (size_t(ny) * gw + nx)], by = grid[2 * (size_t(ny) * gw + nx) + 1];
                    if (bx < 0) continue;
                    float dx = ax - bx, dy = ay - by;
                    if (dx * dx + dy * dy < r2) return 0;
                }
        }
    return 1;
}

