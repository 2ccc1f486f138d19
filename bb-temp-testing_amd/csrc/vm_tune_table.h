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
    {1.00f, 2.20f, 3, 8}, // 2^8: 1.991e+09 site-iters/s at k=16, 2.035e+09 at k=8
    {1.00f, 2.20f, 3, 8}, // 2^9: 2.751e+09 site-iters/s at k=16, 2.847e+09 at k=8
    {1.00f, 2.20f, 3, 12}, // 2^10: 3.462e+09 site-iters/s at k=16, 3.578e+09 at k=12
    {1.00f, 2.20f, 3, 12}, // 2^11: 3.890e+09 site-iters/s at k=16, 4.057e+09 at k=12
    {1.00f, 1.60f, 3, 12}, // 2^12: 4.371e+09 site-iters/s at k=16, 4.505e+09 at k=12
    {1.00f, 2.20f, 3, 16}, // 2^13: 4.435e+09 site-iters/s at k=16, 4.435e+09 at k=16
    {1.00f, 1.60f, 3, 12}, // 2^14: 3.865e+09 site-iters/s at k=16, 4.076e+09 at k=12
    {1.00f, 2.20f, 3, 12}, // 2^15: 4.808e+09 site-iters/s at k=16, 5.103e+09 at k=12
    {1.00f, 2.20f, 3, 16}, // 2^16: 4.968e+09 site-iters/s at k=16, 4.968e+09 at k=16
    {0.90f, 2.20f, 3, 12}, // 2^17: 5.113e+09 site-iters/s at k=16, 5.255e+09 at k=12
    {1.00f, 2.20f, 3, 16}, // 2^18: 4.358e+09 site-iters/s at k=16, 4.358e+09 at k=16
    {0.90f, 2.20f, 3, 16}, // 2^19: 4.632e+09 site-iters/s at k=16, 4.632e+09 at k=16
    {0.90f, 2.20f, 3, 16}, // 2^20: 5.726e+09 site-iters/s at k=16, 5.726e+09 at k=16
    {0.90f, 2.20f, 3, 16}, // 2^21: largest measured class
    {0.90f, 2.20f, 3, 16}, // 2^22: largest measured class
};

} // namespace vmrt
