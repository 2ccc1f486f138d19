#pragma once
// poisson_disk.h -- poisson::Sampler on MI355X (gfx950).
//
// Drop-in for the reference's C++ API (include/poisson/poisson_disk.h:13-62):
// the phase-group dart thrower re-runnable on any block of background-grid
// cells, with a host mirror of the grid.  Header-only over the C ABI of
// libvoromesh_hip.so; a failing call prints the error and exits with status 1.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "voromesh_hip.h"

namespace poisson {

/// Same layout as float2 and vm_vec2. x < 0 marks an empty cell.
struct Sample
{
    float x, y;
};
static_assert(sizeof(Sample) == sizeof(vm_vec2), "Sample must mirror vm_vec2");

class Sampler
{
public:
    Sampler(float width, float height, float radius, int trials = 30, unsigned seed = 1)
        : mWidth(width), mHeight(height), mRadius(radius), mTrials(trials)
    {
        check(vm_sampler_create(width, height, radius, trials, seed, &mS), "Sampler");
        check(vm_sampler_dims(mS, &mGridW, &mGridH, &mCellSize), "Sampler");
        pull();
    }
    ~Sampler() { vm_sampler_destroy(mS); }

    Sampler(const Sampler &) = delete;
    Sampler &operator=(const Sampler &) = delete;

    float Resample(int cx0, int cy0, int cx1, int cy1)
    {
        float ms = 0.0f;
        check(vm_sampler_resample(mS, cx0, cy0, cx1, cy1, &ms), "Resample");
        // Only the region changed: refresh that part of the host mirror.
        check(vm_sampler_cells_region(mS, cx0, cy0, cx1, cy1, reinterpret_cast<vm_vec2 *>(mCells.data()),
                                      int(mCells.size())),
              "Cells");
        return ms;
    }
    float ResampleAll() { return Resample(0, 0, mGridW, mGridH); }

    bool Verify(int cx0, int cy0, int cx1, int cy1) const
    {
        int ok = 0;
        check(vm_sampler_verify(mS, cx0, cy0, cx1, cy1, &ok), "Verify");
        return ok != 0;
    }
    bool VerifyAll() const { return Verify(0, 0, mGridW, mGridH); }

    const std::vector<Sample> &Cells() const { return mCells; }

    float Width() const { return mWidth; }
    float Height() const { return mHeight; }
    float Radius() const { return mRadius; }
    float CellSize() const { return mCellSize; }
    int   GridW() const { return mGridW; }
    int   GridH() const { return mGridH; }
    int   Trials() const { return mTrials; }

private:
    static void check(int status, const char *what)
    {
        if (status != VM_OK) {
            std::fprintf(stderr, "poisson: %s failed (status %d): %s\n", what, status, vm_last_error());
            std::exit(1);
        }
    }
    void pull()
    {
        mCells.resize(size_t(mGridW) * mGridH);
        check(vm_sampler_cells(mS, reinterpret_cast<vm_vec2 *>(mCells.data()), int(mCells.size())), "Cells");
    }

    float               mWidth, mHeight, mRadius, mCellSize = 0.0f;
    int                 mGridW = 0, mGridH = 0, mTrials;
    vm_sampler         *mS = nullptr;
    std::vector<Sample> mCells;
};

} // namespace poisson
