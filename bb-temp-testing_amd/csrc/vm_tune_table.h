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
    {0.90f, 2.20f, 3, 16}, // 2^8: 2.916e+09 site-iters/s at k=16, 2.917e+09 at k=16
    {1.00f, 2.20f, 3, 12}, // 2^9: 3.940e+09 site-iters/s at k=16, 3.981e+09 at k=12
    {1.00f, 2.20f, 3, 12}, // 2^10: 4.847e+09 site-iters/s at k=16, 4.809e+09 at k=12
    {1.00f, 2.20f, 3, 16}, // 2^11: 5.743e+09 site-iters/s at k=16, 5.728e+09 at k=16
    {1.00f, 2.20f, 3, 12}, // 2^12: 6.334e+09 site-iters/s at k=16, 6.331e+09 at k=12
    {1.00f, 1.60f, 3, 12}, // 2^13: 6.854e+09 site-iters/s at k=16, 6.974e+09 at k=12
    {0.90f, 2.20f, 3, 16}, // 2^14: 6.684e+09 site-iters/s at k=16, 6.761e+09 at k=16
    {1.30f, 1.60f, 3, 16}, // 2^15: 6.845e+09 site-iters/s at k=16, 6.860e+09 at k=16
    {0.90f, 1.60f, 4, 12}, // 2^16: 6.796e+09 site-iters/s at k=16, 6.998e+09 at k=12
    {1.30f, 1.60f, 4, 16}, // 2^17: 7.113e+09 site-iters/s at k=16, 7.088e+09 at k=16
    {1.15f, 2.20f, 3, 16}, // 2^18: 6.622e+09 site-iters/s at k=16, 6.642e+09 at k=16
    {1.15f, 2.20f, 3, 16}, // 2^19: 5.799e+09 site-iters/s at k=16, 5.811e+09 at k=16
    {1.15f, 2.20f, 3, 16}, // 2^20: 5.815e+09 site-iters/s at k=16, 5.836e+09 at k=16
    {1.15f, 2.20f, 3, 16}, // 2^21: largest measured class
    {1.15f, 2.20f, 3, 16}, // 2^22: largest measured class
};

} // namespace vmrt
