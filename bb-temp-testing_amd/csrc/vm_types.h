// vm_types.h -- layouts shared by the HIP kernels and the host runtime.
//
// Every optimizer instance ("fit") lives in pooled device arrays and is
// described by one Inst record in a packed instance table.  The per-iteration
// kernels walk work tables (site chunks, target chunks, instance ids), so the
// same four launches serve one fit or hundreds of them.
#pragma once
#include <stdint.h>

namespace vmk {

constexpr int kMaxCellVerts = 24; // clipped cell vertex budget (voro_optimizer.cu:56)
constexpr int kCellK        = 32; // clip neighbours before brute force (voro_optimizer.cu:55)
constexpr int kBoxLabels    = 4;  // edge labels -1..-4 are the box sides (voro_optimizer.cu:58-59)
constexpr int kSiteChunk    = 256;
// Loss kernel: targets per lane (a wavefront takes a chunk of up to 64 * tpl
// targets of one hash cell); a per-fit choice of the tuner among 2, 3 and 4.
constexpr int kTargetsPerLane    = 3;
constexpr int kMaxTargetsPerLane = 4;
constexpr int kLdsSlotBits  = 10;  // low bits of a k-NN key: slot of the site in the staged LDS block
constexpr int kLdsNone      = (1 << kLdsSlotBits) - 1;
constexpr int kStageCap     = kLdsNone - 1; // staged sites per target chunk
constexpr int kMaxSitesPerInst = 1 << 22;  // instance-local site index fits above the LDS slot bits

enum : int { kModeIterate = 0, kModeEvaluate = 1 };

// One optimizer instance.  Offsets are element offsets into the pooled arrays.
struct Inst {
    float   W, H;         // domain [0,W] x [0,H]
    float   cell, inv;    // hash-grid cell size and 1/cell (FP32, as the kernels bin with it)
    float   ext;          // max(W, H): scale of the conservative ring-stop margin
    int     gw, gh;       // hash grid dimensions
    int     S, T;         // live sites, live targets
    int     K;            // loss neighbours
    int     stageR;       // staged ring radius of the loss kernel (block of (2r+1)^2 cells)
    int     allMovable;   // 1: every site moves; 0: consult the frozen mask
    int64_t siteOff;      // sites, m, v, grad, cellOf, rank, sorted pos/idx, frozen
    int64_t cellOff;      // counts and start, gw*gh + 1 entries
    int64_t targetOff;    // sorted targets and their slot numbers
    double  lr, gamma, beta1, beta2, eps; // rounded to the kernel's real type before use
    int     milestones[4];
    float   fraction;     // target subsampling fraction
    uint32_t seed;
    int     useBox;       // targets outside [bx0,bx1]x[by0,by1] are ignored
    float   bx0, by0, bx1, by1;
    int     maxIter;      // 0: no limit
    float   tol;          // relative loss-change convergence threshold (0: off)
    int     gradShift;    // FP32 runs: gradient terms are summed as round(term * 2^gradShift) in int64
    // Sites in slot layout (slot i at (i % rectW, i / rectW)): only slots in
    // [rx0, rx1) x [ry0, ry1) move.  rectW == 0: the frozen mask decides.
    int     rectW, rx0, ry0, rx1, ry1;
    int     pad;
};

// Written by the device every iteration (voro_optimizer.cu:81-85 IterState,
// with the loss sum kept in FP64 and convergence bookkeeping added).
struct InstState {
    int     iter;      // Adam step count, 1-based once a step has started
    int     done;      // converged or reached maxIter: kernels skip the instance
    int64_t count;     // targets used by the loss of the current step
    double  loss;      // sum of their VoroLoss
    double  prevMean;  // mean loss of the previous step (convergence test)
    double  evalLoss;  // EvaluateLoss accumulators
    int64_t evalCount;
    int64_t saturated; // FP32: targets whose gradient terms hit the fixed-point clamp (cumulative)
};

// Per-instance loss accumulators of one step, spread over kLossPartials cache
// lines so the loss kernel's wavefronts of one large instance do not all hit
// one address.  The Adam kernel folds them into InstState (it needs the
// count); the next scan clears them.
constexpr int kLossPartials = 64;
struct LossPartial {
    double             loss;
    unsigned long long count;
    char               pad[112];
};

// Sites [begin, begin+count) of instance `inst`; its pooled site and cell
// offsets are repeated from the record, so a workgroup can issue its site
// loads as soon as it holds the work item.
struct SiteWork {
    int     inst, begin, count, pad;
    int64_t siteOff, cellOff;
};
static_assert(sizeof(SiteWork) == 32, "site work items are two 16-byte words");

// Targets [tBegin, tBegin+count) of the pooled sorted target array, all of
// instance `inst` and in its hash cell (cx, cy); every one lies within `rho`
// of (qx, qy) (static: targets do not move).  The instance's grid width and
// height, pooled offsets and staging rings are repeated from its record, so a
// wavefront can issue its target and cell-start loads as soon as it holds the
// work item, without waiting for the record.
constexpr int kWorkRingsMask = 3;  // flags bits 0-1: staged cell rings
constexpr int kStageRetry    = 3;  // the most rings a block has (7 x 7 cells <= 64 lanes)
constexpr int kWorkSubsample = 4;  // flags bit 2: subsampled targets (read their slot numbers)
struct TargetWork {
    int     inst, count;
    int     cx, cy;
    float   qx, qy, rho;
    int     gw, gh;
    int     flags;
    int64_t tBegin;
    int64_t cellOff, siteOff;
};
static_assert(sizeof(TargetWork) == 64, "work items are four 16-byte words");

// Scan tile: cells [begin, begin+count) of instance `inst` (first: begin == 0).
// The tiles of one instance are consecutive in the table.
struct ScanTile {
    int inst, begin, count, first;
};

// Scan control block: counter over the look-back tiles and flag epoch (the
// scatter kernel after every scan rearms them), the number of tiles and of
// look-back tiles (the tiles of multi-tile instances, first in the table).
struct ScanCtl {
    int      next;
    uint32_t epoch;
    int      nTiles;
    int      nDyn;
};

// A target queued by the loss kernel for the exact k-NN selection: instance
// (bit 31 set: only its k-NN list is wanted) and packed target index.
struct ExactItem {
    int inst, tg;
};

// Poisson dart thrower: one record per instance grid being (re)sampled.
struct PoiInst {
    float    W, H, radius, cellSize;
    int      gw, gh;
    int      x0, y0, x1, y1; // region being sampled
    uint32_t seed, roundBase;
    int64_t  gridOff;        // float2 offset of the instance grid
};

struct PoiWork {
    int inst, base; // phase-cell indices [base, base + blockDim) of instance `inst`
};

} // namespace vmk
