// vm_tune_table.h -- per-size-class choices of the tuner (written by
// bb-temp-testing_amd/vmhip/tune.py from a measured sweep; edit by re-running it).
// Row i covers fits of 2^(kTuneLog2Min + i) .. 2^(kTuneLog2Min + i + 1) - 1 sites.
#pragma once

namespace vmrt {

struct TuneRow {
    float cellScale;   // multiple of the base hash-grid cell max(1.5, (sqrt(K/pi) + 0.8) / 2) spacings
    float chunkExtent; // target chunk diagonal, in mean site spacings
};

constexpr int     kTuneLog2Min = 8;
constexpr int     kTuneLog2Max = 22;
constexpr TuneRow kTuneTable[kTuneLog2Max - kTuneLog2Min + 1] = {
    {1.00f, 2.20f}, // 2^8: measured 1.515e+09 site-iters/s
    {1.00f, 2.20f}, // 2^9: measured 2.062e+09 site-iters/s
    {1.00f, 2.20f}, // 2^10: measured 2.613e+09 site-iters/s
    {1.00f, 3.00f}, // 2^11: measured 3.030e+09 site-iters/s
    {1.00f, 2.20f}, // 2^12: measured 3.355e+09 site-iters/s
    {1.00f, 2.20f}, // 2^13: measured 3.588e+09 site-iters/s
    {1.00f, 3.00f}, // 2^14: measured 3.569e+09 site-iters/s
    {1.00f, 2.20f}, // 2^15: measured 3.764e+09 site-iters/s
    {1.00f, 2.20f}, // 2^16: measured 3.852e+09 site-iters/s
    {1.00f, 3.00f}, // 2^17: measured 3.902e+09 site-iters/s
    {1.00f, 1.60f}, // 2^18: measured 3.817e+09 site-iters/s
    {1.00f, 1.60f}, // 2^19: largest measured class
    {1.00f, 1.60f}, // 2^20: largest measured class
    {1.00f, 1.60f}, // 2^21: largest measured class
    {1.00f, 1.60f}, // 2^22: largest measured class
};

} // namespace vmrt
