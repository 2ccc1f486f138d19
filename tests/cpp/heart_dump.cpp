// heart_dump.cpp -- prints voromesh::MakeHeart(cx, cy, width, segments) of the
// drop-in header as hexadecimal float bits, one vertex per line, so the CPU
// tests can compare it with the outline the reference's MakeHeart
// (src/voromesh/shape2d.cpp) writes into tests/golden/canvas_heart.npz.
// Usage: heart_dump cx cy width segments
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "voromesh/shape2d.h"

int main(int argc, char **argv)
{
    if (argc != 5) {
        std::fprintf(stderr, "usage: heart_dump cx cy width segments\n");
        return 2;
    }
    const auto poly = voromesh::MakeHeart(float(std::atof(argv[1])), float(std::atof(argv[2])),
                                          float(std::atof(argv[3])), std::atoi(argv[4]));
    for (const auto &p : poly) {
        uint32_t bx, by;
        std::memcpy(&bx, &p.x, 4);
        std::memcpy(&by, &p.y, 4);
        std::printf("%08x %08x\n", bx, by);
    }
    return 0;
}
