// vm_tune_table.h -- per-size-class choices of the tuner (written by
// bb-temp-testing_amd/vmhip/tune.py from a measured sweep; edit by re-running it).
// Row i covers fits of 2^(kTuneLog2Min + i) .. 2^(kTuneLog2Min + i + 1) - 1 sites.
#pragma once

namespace vmrt {

struct TuneRow {
    float cellScale;   // multiple of the base hash-grid cell max(1.5, (sqrt(K/pi) + 0.8) / 2) spacings
    float chunkExtent; // target chunk diagonal, in mean site spacings
    int   tpl;         // loss kernel targets per lane
    int   k;           // loss neighbours when a fit leaves k to the tuner
};

constexpr int     kTuneLog2Min = 8;
constexpr int     kTuneLog2Max = 22;
constexpr TuneRow kTuneTable[kTuneLog2Max - kTuneLog2Min + 1] = {
    {1.00f, 4.00f, 2, 8}, // 2^8: 2.013e+09 site-iters/s at k=16, 2.072e+09 at k=8
    {1.00f, 3.00f, 2, 8}, // 2^9: 2.818e+09 site-iters/s at k=16, 2.891e+09 at k=8
    {1.00f, 2.20f, 2, 12}, // 2^10: 3.497e+09 site-iters/s at k=16, 3.569e+09 at k=12
    {1.00f, 3.00f, 2, 12}, // 2^11: 3.993e+09 site-iters/s at k=16, 4.103e+09 at k=12
    {1.00f, 2.20f, 3, 12}, // 2^12: 4.246e+09 site-iters/s at k=16, 4.299e+09 at k=12
    {1.00f, 4.00f, 3, 12}, // 2^13: 4.209e+09 site-iters/s at k=16, 4.313e+09 at k=12
    {0.80f, 1.60f, 4, 12}, // 2^14: 3.875e+09 site-iters/s at k=16, 4.000e+09 at k=12
    {0.90f, 1.60f, 4, 12}, // 2^15: 4.821e+09 site-iters/s at k=16, 4.955e+09 at k=12
    {1.00f, 1.60f, 4, 12}, // 2^16: 4.790e+09 site-iters/s at k=16, 4.873e+09 at k=12
    {1.00f, 3.00f, 4, 12}, // 2^17: 4.898e+09 site-iters/s at k=16, 4.917e+09 at k=12
    {0.90f, 2.20f, 4, 16}, // 2^18: 4.210e+09 site-iters/s at k=16, 4.210e+09 at k=16
    {0.90f, 3.00f, 3, 16}, // 2^19: 4.558e+09 site-iters/s at k=16, 4.558e+09 at k=16
    {0.90f, 2.20f, 3, 16}, // 2^20: 5.406e+09 site-iters/s at k=16, 5.406e+09 at k=16
    {0.90f, 2.20f, 3, 16}, // 2^21: largest measured class
    {0.90f, 2.20f, 3, 16}, // 2^22: largest measured class
};

} // namespace vmrt
