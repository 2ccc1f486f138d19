// vm_noalloc.cpp -- vm_poisson_batch_noalloc: the batched Poisson-disk dart
// thrower (vm_poisson_batch) with every device buffer it needs -- the
// instance table, the clear and phase work tables and the instance grids --
// placed in one fixed array of device memory that the library owns, so a call
// allocates nothing on the device (no hipMalloc / hipFree) or on the host
// heap.  For callers that may not allocate: tight serving loops and harnesses
// that forbid allocation.
//
// The array is a file-scope __device__ variable.  Its host-side name is only a
// stand-in for the symbol: the device address comes from hipGetSymbolAddress,
// looked up once per device (with that device current) on the first call made
// on it, and only that address is handed to copies and kernels.  Each device
// holds its own copy of the array until the process ends; the stream and the
// two timing events of a device are created at the same one-time set-up.
//
// Calls from one process are serialized by a mutex (the array and the host
// staging tables are shared by every call).  The array keeps the last call's
// contents; every instance grid is cleared by the clear kernel before the
// darts are thrown, so no result depends on an earlier call.
#include "voromesh_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "vm_engine.h"

using vmrt::VmError;

namespace {
constexpr size_t kScratchBytes = size_t(16) << 20; // per device
}

// The device array (see above): a stand-in name on the host.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int    kMaxDevices    = 64;
constexpr int    kMaxInstances  = 16384;
constexpr int    kWorkThreads   = 256; // cells per clear / phase work entry (the kernels' block size)
// Work entries per table: sum over instances of ceil(cells / 256) is at most
// (grid cells) / 256 + instances, and the grids fit the array.
constexpr size_t kMaxWork = kScratchBytes / 8 / kWorkThreads + kMaxInstances;

struct DeviceSlot {
    unsigned char *ptr    = nullptr; // device address of `scratch` on this device
    hipStream_t    stream = nullptr;
    hipEvent_t     e0 = nullptr, e1 = nullptr;
};

std::mutex   gMu;
DeviceSlot   gSlots[kMaxDevices];
vmk::PoiInst gInst[kMaxInstances];
vmk::PoiWork gClear[kMaxWork], gPhase[kMaxWork];

// The current device's slot, set up on its first use (symbol address, stream,
// events).  Called with gMu held.
DeviceSlot &currentSlot()
{
    int dev = 0;
    VM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw VmError(VM_EUNSUPPORTED, "device index beyond the scratch table");
    DeviceSlot &d = gSlots[dev];
    if (d.ptr) return d;
    void *p = nullptr;
    VM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!d.stream) VM_HIP(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
    if (!d.e0) VM_HIP(hipEventCreate(&d.e0));
    if (!d.e1) VM_HIP(hipEventCreate(&d.e1));
    d.ptr = static_cast<unsigned char *>(p);
    return d;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

} // namespace

extern "C" int vm_poisson_batch_noalloc(int n, const float *w, const float *h, const float *r, const unsigned *seed,
                                        int trials, const int64_t *offsets, vm_vec2 *cells, float *gpuMs)
{
    return vmrt::guardedCall([&] {
        if (!(n >= 0 && (n == 0 || (w && h && r && seed && offsets && cells)) && trials >= 0))
            throw VmError(VM_EINVAL, "invalid arguments");
        if (gpuMs) *gpuMs = 0.0f;
        if (n == 0) return;
        if (n > kMaxInstances)
            throw VmError(VM_EINVAL, "vm_poisson_batch_noalloc takes at most " + std::to_string(kMaxInstances) +
                                         " instances (got " + std::to_string(n) + ")");
        std::lock_guard<std::mutex> lock(gMu);
        // Instance records and the size of every table, checked against the
        // array before anything is enqueued.
        int64_t total = 0;
        size_t  nClear = 0, nPhase = 0;
        for (int i = 0; i < n; ++i) {
            if (!(w[i] > 0 && h[i] > 0 && r[i] > 0) || offsets[i] < 0) throw VmError(VM_EINVAL, "invalid instance");
            vmk::PoiInst &p = gInst[i];
            std::memset(&p, 0, sizeof(p));
            // poisson_disk.cu:119-124: cell r / sqrt(2), dims ceil(W / cell), in FP32.
            p.cellSize = r[i] / sqrtf(2.0f);
            p.gw = int(ceilf(w[i] / p.cellSize)), p.gh = int(ceilf(h[i] / p.cellSize));
            p.W = w[i], p.H = h[i], p.radius = r[i], p.seed = seed[i], p.roundBase = 0;
            p.x0 = 0, p.y0 = 0, p.x1 = p.gw, p.y1 = p.gh;
            p.gridOff = offsets[i];
            const int64_t c = int64_t(p.gw) * p.gh;
            total = std::max(total, offsets[i] + c);
            nClear += size_t((c + kWorkThreads - 1) / kWorkThreads);
            const int64_t np = int64_t((p.gw + 2) / 3) * ((p.gh + 2) / 3);
            nPhase += size_t((np + kWorkThreads - 1) / kWorkThreads);
        }
        const size_t offClear = align256(size_t(n) * sizeof(vmk::PoiInst));
        const size_t offPhase = offClear + align256(nClear * sizeof(vmk::PoiWork));
        const size_t offGrid  = offPhase + align256(nPhase * sizeof(vmk::PoiWork));
        const size_t need     = offGrid + size_t(total) * sizeof(vm_vec2);
        if (need > kScratchBytes || nClear > kMaxWork || nPhase > kMaxWork)
            throw VmError(VM_EINVAL, "vm_poisson_batch_noalloc needs " + std::to_string(need) +
                                         " bytes of device scratch; its fixed array holds " +
                                         std::to_string(kScratchBytes) + " (about " +
                                         std::to_string(kScratchBytes / sizeof(vm_vec2) / 1000000) +
                                         " M grid cells with the tables)");
        size_t kc = 0, kp = 0;
        for (int i = 0; i < n; ++i) {
            const vmk::PoiInst &p  = gInst[i];
            const int64_t       c  = int64_t(p.gw) * p.gh;
            const int64_t       np = int64_t((p.gw + 2) / 3) * ((p.gh + 2) / 3);
            for (int64_t b = 0; b < c; b += kWorkThreads) gClear[kc++] = vmk::PoiWork{i, int(b)};
            for (int64_t b = 0; b < np; b += kWorkThreads) gPhase[kp++] = vmk::PoiWork{i, int(b)};
        }

        DeviceSlot   &d     = currentSlot();
        unsigned char *base = d.ptr;
        auto *dInst  = reinterpret_cast<vmk::PoiInst *>(base);
        auto *dClear = reinterpret_cast<vmk::PoiWork *>(base + offClear);
        auto *dPhase = reinterpret_cast<vmk::PoiWork *>(base + offPhase);
        auto *dGrid  = reinterpret_cast<float *>(base + offGrid);
        VM_HIP(hipMemcpyAsync(dInst, gInst, size_t(n) * sizeof(vmk::PoiInst), hipMemcpyHostToDevice, d.stream));
        VM_HIP(hipMemcpyAsync(dClear, gClear, kc * sizeof(vmk::PoiWork), hipMemcpyHostToDevice, d.stream));
        VM_HIP(hipMemcpyAsync(dPhase, gPhase, kp * sizeof(vmk::PoiWork), hipMemcpyHostToDevice, d.stream));
        const int gc = int(std::min<size_t>(kc, 8192)), gp = int(std::min<size_t>(kp, 8192));
        VM_HIP(hipEventRecord(d.e0, d.stream));
        VM_HIP(vmk::launchPoissonClear(gc, d.stream, dInst, dClear, int(kc), dGrid));
        for (int round = 0; round < trials; ++round)
            for (int py = 0; py < 3; ++py)
                for (int px = 0; px < 3; ++px)
                    VM_HIP(vmk::launchPoissonPhase(gp, d.stream, dInst, dPhase, int(kp), dGrid, px, py, round));
        VM_HIP(hipEventRecord(d.e1, d.stream));
        VM_HIP(hipMemcpyAsync(cells, dGrid, size_t(total) * sizeof(vm_vec2), hipMemcpyDeviceToHost, d.stream));
        VM_HIP(hipStreamSynchronize(d.stream));
        float ms = 0;
        VM_HIP(hipEventElapsedTime(&ms, d.e0, d.e1));
        if (gpuMs) *gpuMs = ms;
    });
}
