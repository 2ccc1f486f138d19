// vm_canvas.cpp -- the voxel canvas that produces the optimizer's target slots
// and occupancy masks (include/voromesh/canvas.h:17-89 and
// src/voromesh/canvas.cpp of the reference).  Host code: it runs on the CPU
// next to the optimizer and hands over plain arrays.
//
// Voxel (i, j) covers [i, i+1) x [j, j+1) voxels; 8x8 voxels form a tile, and
// voxel (i, j) is bit (j % 8) * 8 + i % 8 of its tile's word.  Occupancy is a
// dense row-major array of tile words (bit set: voxel centre inside).  The
// surface is sparse: per tile that the outline crosses, a mask of surface
// voxels, each holding the mean of the outline samples that fell in it and the
// target slot that mirrors it.  Slot numbers, sample positions and the order in
// which edits touch slots follow the reference exactly: the optimizer draws its
// per-iteration target subsample from the slot number, so a different
// numbering would be a different fit.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "vm_engine.h"
#include "voromesh_hip.h"

namespace {

constexpr int   kTileLog2 = 3;
constexpr int   kTile     = 1 << kTileLog2;
constexpr float kPiF      = 3.14159265f; // the reference's rim-sampling constant

inline uint64_t pairKey(int lo, int hi) { return (uint64_t(uint32_t(hi)) << 32) | uint32_t(lo); }
inline int      keyLo(uint64_t k) { return int32_t(uint32_t(k)); }
inline int      keyHi(uint64_t k) { return int32_t(uint32_t(k >> 32)); }
inline int      tileBit(int i, int j) { return ((j & (kTile - 1)) << kTileLog2) | (i & (kTile - 1)); }

// Outline samples of one voxel, summed in double (exact for float inputs of one
// voxel's magnitude, so the mean does not depend on the order of the samples).
struct SampleSum {
    double sx = 0.0, sy = 0.0;
    int    n  = 0;
    void    add(float x, float y) { sx += x, sy += y, ++n; }
    vm_vec2 mean() const { return vm_vec2{float(sx / n), float(sy / n)}; }
};

struct SurfTile {
    uint64_t mask = 0;
    vm_vec2  pt[kTile * kTile];
    int32_t  slot[kTile * kTile];
};

} // namespace

struct vm_canvas {
    float                                  voxel;
    int                                    vx, vy, tx, ty;
    std::vector<uint64_t>                  inside;
    std::unordered_map<uint64_t, SurfTile> surf; // key (ty << 32 | tx)
    int64_t                                surfCount = 0;
    std::vector<vm_vec2>                   slots;
    std::vector<int32_t>                   freeSlots;
    std::vector<int32_t>                   dirty; // of the last stamp

    vm_canvas(float w, float h, float v)
        : voxel(v), vx(int(std::ceil(w / v))), vy(int(std::ceil(h / v))), tx((vx + kTile - 1) / kTile),
          ty((vy + kTile - 1) / kTile), inside(size_t(tx) * size_t(ty), 0)
    {
    }

    bool inDomain(int i, int j) const { return i >= 0 && j >= 0 && i < vx && j < vy; }
    bool isInside(int i, int j) const
    {
        return inDomain(i, j) && ((inside[size_t(j >> kTileLog2) * tx + (i >> kTileLog2)] >> tileBit(i, j)) & 1u);
    }
    bool isSurface(int i, int j) const
    {
        auto it = surf.find(pairKey(i >> kTileLog2, j >> kTileLog2));
        return it != surf.end() && ((it->second.mask >> tileBit(i, j)) & 1u);
    }

    // Free slots are reused last-freed first.
    int32_t alloc(vm_vec2 p)
    {
        int32_t s;
        if (!freeSlots.empty()) {
            s = freeSlots.back();
            freeSlots.pop_back();
        } else {
            s = int32_t(slots.size());
            slots.push_back(vm_vec2{0.0f, 0.0f});
        }
        slots[size_t(s)] = p;
        return s;
    }
    void release(int32_t s)
    {
        slots[size_t(s)] = vm_vec2{-1.0f, -1.0f};
        freeSlots.push_back(s);
    }

    void rasterize(const vm_vec2 *poly, int n)
    {
        surf.clear();
        surfCount = 0;
        slots.clear();
        freeSlots.clear();
        dirty.clear();
        std::fill(inside.begin(), inside.end(), uint64_t(0));
        if (n <= 0) return;

        // Outline: every edge sampled at quarter-voxel steps (the edge's end is
        // the next edge's first sample).
        const float                             inv = 1.0f / voxel, step = 0.25f * voxel;
        std::unordered_map<uint64_t, SampleSum> sums;
        for (int s = 0; s < n; ++s) {
            const vm_vec2 a = poly[s], b = poly[(s + 1) % n];
            const float   len = std::hypot(b.x - a.x, b.y - a.y);
            const int     m   = std::max(1, int(std::ceil(len / step)));
            for (int k = 0; k < m; ++k) {
                const float t = float(k) / float(m);
                const float x = a.x + t * (b.x - a.x), y = a.y + t * (b.y - a.y);
                const int   i = int(std::floor(x * inv)), j = int(std::floor(y * inv));
                if (inDomain(i, j)) sums[pairKey(i, j)].add(x, y);
            }
        }
        // Slots in order of (tile row, tile column, bit inside the tile).
        std::vector<uint64_t> keys;
        keys.reserve(sums.size());
        for (const auto &kv : sums) keys.push_back(kv.first);
        std::sort(keys.begin(), keys.end(), [](uint64_t p, uint64_t q) {
            const int pi = keyLo(p), pj = keyHi(p), qi = keyLo(q), qj = keyHi(q);
            if ((pj >> kTileLog2) != (qj >> kTileLog2)) return (pj >> kTileLog2) < (qj >> kTileLog2);
            if ((pi >> kTileLog2) != (qi >> kTileLog2)) return (pi >> kTileLog2) < (qi >> kTileLog2);
            return tileBit(pi, pj) < tileBit(qi, qj);
        });
        for (const uint64_t key : keys) {
            const int i = keyLo(key), j = keyHi(key), bit = tileBit(i, j);
            SurfTile &t = surf[pairKey(i >> kTileLog2, j >> kTileLog2)];
            t.mask |= uint64_t(1) << bit;
            t.pt[bit]   = sums[key].mean();
            t.slot[bit] = alloc(t.pt[bit]);
            ++surfCount;
        }
        // Occupancy: even-odd scanlines through voxel centres.
        std::vector<float> xs;
        for (int j = 0; j < vy; ++j) {
            const float y = (float(j) + 0.5f) * voxel;
            xs.clear();
            for (int s = 0, t = n - 1; s < n; t = s++) {
                const vm_vec2 a = poly[s], b = poly[t];
                if ((a.y > y) != (b.y > y)) xs.push_back(a.x + (y - a.y) * (b.x - a.x) / (b.y - a.y));
            }
            std::sort(xs.begin(), xs.end());
            for (size_t k = 0; k + 1 < xs.size(); k += 2) {
                const int i0 = std::max(int(std::ceil(xs[k] * inv - 0.5f)), 0);
                const int i1 = std::min(int(std::floor(xs[k + 1] * inv - 0.5f)), vx - 1);
                for (int i = i0; i <= i1; ++i)
                    inside[size_t(j >> kTileLog2) * tx + (i >> kTileLog2)] |= uint64_t(1) << tileBit(i, j);
            }
        }
    }

    void stamp(vm_vec2 c, float radius, bool in, vm_canvas_edit &e)
    {
        std::memset(&e, 0, sizeof(e));
        dirty.clear();
        const float inv = 1.0f / voxel, r2 = radius * radius;
        const int   i0 = std::max(int(std::floor((c.x - radius) * inv)), 0);
        const int   j0 = std::max(int(std::floor((c.y - radius) * inv)), 0);
        const int   i1 = std::min(int(std::floor((c.x + radius) * inv)), vx - 1);
        const int   j1 = std::min(int(std::floor((c.y + radius) * inv)), vy - 1);
        if (i0 > i1 || j0 > j1) return;

        // 1. Rim samples on the other side of the occupancy as it is now.  The
        //    map's iteration order (same keys inserted in the same order) sets
        //    the order in which new rim voxels take slots in step 4.
        std::unordered_map<uint64_t, SampleSum> rim;
        const int n = std::max(16, int(std::ceil(2.0f * kPiF * radius / (0.25f * voxel))));
        for (int k = 0; k < n; ++k) {
            const float a = 2.0f * kPiF * float(k) / float(n);
            const float x = c.x + radius * std::cos(a), y = c.y + radius * std::sin(a);
            const int   i = int(std::floor(x * inv)), j = int(std::floor(y * inv));
            if (inDomain(i, j) && isInside(i, j) != in) rim[pairKey(i, j)].add(x, y);
        }
        // 2. Surface samples strictly inside the disc are deleted.
        for (int ty0 = j0 >> kTileLog2; ty0 <= (j1 >> kTileLog2); ++ty0) {
            for (int tx0 = i0 >> kTileLog2; tx0 <= (i1 >> kTileLog2); ++tx0) {
                auto it = surf.find(pairKey(tx0, ty0));
                if (it == surf.end()) continue;
                SurfTile &t = it->second;
                for (uint64_t m = t.mask; m; m &= m - 1) {
                    const int     bit = __builtin_ctzll(m);
                    const vm_vec2 p   = t.pt[bit];
                    if ((p.x - c.x) * (p.x - c.x) + (p.y - c.y) * (p.y - c.y) >= r2) continue;
                    t.mask &= ~(uint64_t(1) << bit);
                    release(t.slot[bit]);
                    dirty.push_back(t.slot[bit]);
                    --surfCount;
                    ++e.removed;
                }
                if (!t.mask) surf.erase(it);
            }
        }
        // 3. Occupancy of the voxel centres inside the disc.
        for (int j = j0; j <= j1; ++j) {
            const float dy = (float(j) + 0.5f) * voxel - c.y;
            for (int i = i0; i <= i1; ++i) {
                const float dx = (float(i) + 0.5f) * voxel - c.x;
                if (dx * dx + dy * dy >= r2) continue;
                uint64_t      &w   = inside[size_t(j >> kTileLog2) * tx + (i >> kTileLog2)];
                const uint64_t bit = uint64_t(1) << tileBit(i, j);
                if (bool(w & bit) != in) {
                    w ^= bit;
                    ++e.flipped;
                }
            }
        }
        // 4. The kept rim becomes surface, replacing an old sample in its voxel.
        for (const auto &kv : rim) {
            const int     i = keyLo(kv.first), j = keyHi(kv.first), bit = tileBit(i, j);
            SurfTile     &t = surf[pairKey(i >> kTileLog2, j >> kTileLog2)];
            const vm_vec2 p = kv.second.mean();
            if ((t.mask >> bit) & 1u) {
                slots[size_t(t.slot[bit])] = p;
            } else {
                t.mask |= uint64_t(1) << bit;
                t.slot[bit] = alloc(p);
                ++surfCount;
                ++e.added;
            }
            t.pt[bit] = p;
            dirty.push_back(t.slot[bit]);
        }
        e.tx0       = i0 >> kTileLog2;
        e.ty0       = j0 >> kTileLog2;
        e.tx1       = (i1 >> kTileLog2) + 1;
        e.ty1       = (j1 >> kTileLog2) + 1;
        e.num_dirty = int32_t(dirty.size());
    }
};

namespace {

void require(bool ok, const char *msg)
{
    if (!ok) throw vmrt::VmError(VM_EINVAL, msg);
}

} // namespace

extern "C" {

VM_API int vm_canvas_create(float width, float height, float voxel_size, vm_canvas **out)
{
    return vmrt::guardedCall([&] {
        require(out != nullptr, "vm_canvas_create: null output");
        *out = nullptr;
        require(width > 0.0f && height > 0.0f && voxel_size > 0.0f,
                "vm_canvas_create: width, height and voxel size must be positive");
        require(double(width) / voxel_size * (double(height) / voxel_size) <= double(1u << 30),
                "vm_canvas_create: more than 2^30 voxels");
        *out = new vm_canvas(width, height, voxel_size);
    });
}

VM_API int vm_canvas_destroy(vm_canvas *c)
{
    delete c;
    return VM_OK;
}

VM_API int vm_canvas_rasterize(vm_canvas *c, const vm_vec2 *polygon, int n)
{
    return vmrt::guardedCall([&] {
        require(c && n >= 0 && (n == 0 || polygon), "vm_canvas_rasterize: bad arguments");
        c->rasterize(polygon, n);
    });
}

VM_API int vm_canvas_stamp_disc(vm_canvas *c, vm_vec2 center, float radius, int inside, vm_canvas_edit *edit)
{
    return vmrt::guardedCall([&] {
        require(c && edit, "vm_canvas_stamp_disc: bad arguments");
        c->stamp(center, radius, inside != 0, *edit);
    });
}

VM_API int vm_canvas_dirty(const vm_canvas *c, int32_t *out, int n)
{
    return vmrt::guardedCall([&] {
        require(c && n >= 0 && (n == 0 || out), "vm_canvas_dirty: bad arguments");
        require(size_t(n) >= c->dirty.size(), "vm_canvas_dirty: buffer smaller than the edit's dirty list");
        std::copy(c->dirty.begin(), c->dirty.end(), out);
    });
}

VM_API int vm_canvas_dims(const vm_canvas *c, int *voxels_x, int *voxels_y, int *tiles_x, int *tiles_y,
                          float *voxel_size)
{
    return vmrt::guardedCall([&] {
        require(c != nullptr, "vm_canvas_dims: null canvas");
        if (voxels_x) *voxels_x = c->vx;
        if (voxels_y) *voxels_y = c->vy;
        if (tiles_x) *tiles_x = c->tx;
        if (tiles_y) *tiles_y = c->ty;
        if (voxel_size) *voxel_size = c->voxel;
    });
}

VM_API int vm_canvas_counts(const vm_canvas *c, int64_t *num_slots, int64_t *surface_count, int64_t *surface_tiles)
{
    return vmrt::guardedCall([&] {
        require(c != nullptr, "vm_canvas_counts: null canvas");
        if (num_slots) *num_slots = int64_t(c->slots.size());
        if (surface_count) *surface_count = c->surfCount;
        if (surface_tiles) *surface_tiles = int64_t(c->surf.size());
    });
}

VM_API int vm_canvas_slots(const vm_canvas *c, vm_vec2 *out, int64_t n)
{
    return vmrt::guardedCall([&] {
        require(c && n >= 0 && (n == 0 || out), "vm_canvas_slots: bad arguments");
        require(size_t(n) >= c->slots.size(), "vm_canvas_slots: buffer smaller than the slot array");
        std::copy(c->slots.begin(), c->slots.end(), out);
    });
}

VM_API int vm_canvas_masks(const vm_canvas *c, uint64_t *out, int64_t n)
{
    return vmrt::guardedCall([&] {
        require(c && n >= 0 && (n == 0 || out), "vm_canvas_masks: bad arguments");
        require(size_t(n) >= c->inside.size(), "vm_canvas_masks: buffer smaller than the tile array");
        std::copy(c->inside.begin(), c->inside.end(), out);
    });
}

VM_API int vm_canvas_voxel(const vm_canvas *c, int i, int j, int *inside, int *surface)
{
    return vmrt::guardedCall([&] {
        require(c != nullptr, "vm_canvas_voxel: null canvas");
        if (inside) *inside = c->isInside(i, j) ? 1 : 0;
        if (surface) *surface = c->isSurface(i, j) ? 1 : 0;
    });
}

} // extern "C"
