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
    {1.00f, 2.20f}, // 2^8: measured 1.546e+09 site-iters/s
    {1.00f, 2.20f}, // 2^9: measured 2.077e+09 site-iters/s
    {1.00f, 2.20f}, // 2^10: measured 2.641e+09 site-iters/s
    {1.00f, 2.20f}, // 2^11: measured 3.070e+09 site-iters/s
    {1.00f, 2.20f}, // 2^12: measured 3.374e+09 site-iters/s
    {1.00f, 1.60f}, // 2^13: measured 3.525e+09 site-iters/s
    {1.00f, 1.60f}, // 2^14: measured 3.485e+09 site-iters/s
    {1.00f, 2.20f}, // 2^15: measured 3.611e+09 site-iters/s
    {1.00f, 1.60f}, // 2^16: measured 3.672e+09 site-iters/s
    {1.00f, 2.20f}, // 2^17: measured 3.793e+09 site-iters/s
    {0.90f, 2.20f}, // 2^18: measured 3.645e+09 site-iters/s
    {0.90f, 2.20f}, // 2^19: largest measured class
    {0.90f, 2.20f}, // 2^20: largest measured class
    {0.90f, 2.20f}, // 2^21: largest measured class
    {0.90f, 2.20f}, // 2^22: largest measured class
};

} // namespace vmrt
