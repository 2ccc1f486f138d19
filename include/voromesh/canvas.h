#pragma once
// canvas.h -- voromesh::Canvas: the 2D shape as surface samples (target slots)
// and inside/outside occupancy (8x8-voxel tile masks).
//
// Drop-in for the reference's include/voromesh/canvas.h:25-89 (same types,
// methods and results: slot numbers, sample positions, masks and the dirty
// lists of edits are bit-identical).  Header-only over the C ABI of
// libvoromesh_hip.so (vm_canvas_*), which runs the canvas on the host; a
// failing call prints the error on stderr and exits with status 1, like the
// other classes of this API.
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "voromesh/shape2d.h"
#include "voromesh_hip.h"

namespace voromesh {

class Canvas
{
public:
    static constexpr int kTileLog2 = 3;
    static constexpr int kTileSize = 1 << kTileLog2; // voxels per tile side

    /// Layout of one sparse surface tile (canvas.h:31-36 of the reference); the
    /// tiles themselves live in the library.
    struct SurfaceTile
    {
        uint64_t mask = 0;
        Vec2     points[kTileSize * kTileSize];
        int      slots[kTileSize * kTileSize];
    };

    struct Edit
    {
        std::vector<int> dirtySlots;            // target slots whose value changed
        int tx0 = 0, ty0 = 0, tx1 = 0, ty1 = 0; // occupancy tiles [tx0,tx1) x [ty0,ty1) rewritten
        int removed = 0, added = 0, flipped = 0; // surface samples, occupancy voxels

        bool Changed() const { return removed + added + flipped > 0; }
    };

    Canvas(float width, float height, float voxelSize)
    {
        check(vm_canvas_create(width, height, voxelSize, &mC), "Canvas");
        check(vm_canvas_dims(mC, &mVoxelsX, &mVoxelsY, &mTilesX, &mTilesY, &mVoxelSize), "Canvas");
        mInside.assign(size_t(mTilesX) * size_t(mTilesY), 0);
    }
    ~Canvas() { vm_canvas_destroy(mC); }
    Canvas(const Canvas &)            = delete;
    Canvas &operator=(const Canvas &) = delete;

    /// Replaces the canvas with a closed polygon: its outline becomes the
    /// surface and its interior (by voxel centre) the occupancy.
    void RasterizeShape(const std::vector<Vec2> &polygon)
    {
        check(vm_canvas_rasterize(mC, reinterpret_cast<const vm_vec2 *>(polygon.data()), int(polygon.size())),
              "RasterizeShape");
        refresh();
    }

    /// Stamps a disc: its interior becomes `inside`, surface samples strictly
    /// inside it are deleted, and its rim becomes surface where it lay on the
    /// other side.
    Edit StampDisc(Vec2 center, float radius, bool inside)
    {
        vm_canvas_edit e{};
        check(vm_canvas_stamp_disc(mC, vm_vec2{center.x, center.y}, radius, inside ? 1 : 0, &e), "StampDisc");
        Edit out;
        out.dirtySlots.resize(size_t(e.num_dirty));
        check(vm_canvas_dirty(mC, out.dirtySlots.data(), e.num_dirty), "StampDisc");
        out.tx0 = e.tx0, out.ty0 = e.ty0, out.tx1 = e.tx1, out.ty1 = e.ty1;
        out.removed = e.removed, out.added = e.added, out.flipped = e.flipped;
        if (out.Changed()) refresh();
        return out;
    }

    bool InDomain(int i, int j) const { return i >= 0 && j >= 0 && i < mVoxelsX && j < mVoxelsY; }
    bool IsInside(int i, int j) const
    {
        return InDomain(i, j) &&
               ((mInside[size_t(j >> kTileLog2) * size_t(mTilesX) + size_t(i >> kTileLog2)] >>
                 (((j & (kTileSize - 1)) << kTileLog2) | (i & (kTileSize - 1)))) & 1u);
    }
    bool IsSurface(int i, int j) const
    {
        int in = 0, surf = 0;
        check(vm_canvas_voxel(mC, i, j, &in, &surf), "IsSurface");
        return surf != 0;
    }

    /// Host mirror of the GPU target array; {-1, -1} marks a free slot.
    const std::vector<Vec2>     &TargetSlots() const { return mSlots; }
    const std::vector<uint64_t> &InsideMasks() const { return mInside; }

    int    TilesX() const { return mTilesX; }
    int    TilesY() const { return mTilesY; }
    float  VoxelSize() const { return mVoxelSize; }
    size_t SurfaceCount() const { return size_t(mSurfaceCount); }
    size_t SurfaceTileCount() const { return size_t(mSurfaceTiles); }

private:
    static void check(int status, const char *what)
    {
        if (status != VM_OK) {
            std::fprintf(stderr, "voromesh: %s failed (status %d): %s\n", what, status, vm_last_error());
            std::exit(1);
        }
    }

    void refresh()
    {
        int64_t n = 0;
        check(vm_canvas_counts(mC, &n, &mSurfaceCount, &mSurfaceTiles), "Canvas");
        mSlots.resize(size_t(n));
        check(vm_canvas_slots(mC, reinterpret_cast<vm_vec2 *>(mSlots.data()), n), "Canvas");
        check(vm_canvas_masks(mC, mInside.data(), int64_t(mInside.size())), "Canvas");
    }

    vm_canvas            *mC = nullptr;
    int                   mVoxelsX = 0, mVoxelsY = 0, mTilesX = 0, mTilesY = 0;
    float                 mVoxelSize = 0.0f;
    int64_t               mSurfaceCount = 0, mSurfaceTiles = 0;
    std::vector<Vec2>     mSlots;
    std::vector<uint64_t> mInside;
};

} // namespace voromesh
