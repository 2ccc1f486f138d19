#pragma once
// shape2d.h -- 2D point type of the VoroMesh API (include/voromesh/shape2d.h:11-14 of
// the reference) and the parametric heart outline its demo fits.
#include <cmath>
#include <vector>

#include "voromesh_hip.h"

namespace voromesh {

/// Same layout as float2 and vm_vec2.
struct Vec2
{
    float x, y;
};
static_assert(sizeof(Vec2) == sizeof(vm_vec2), "Vec2 must mirror vm_vec2");

/// Closed polygon (last vertex not repeated) on the heart curve
/// x = 16 sin^3 t, y = 13 cos t - 5 cos 2t - 2 cos 3t - cos 4t, scaled to
/// `width` and centred on (cx, cy); y points down (screen convention).
inline std::vector<Vec2> MakeHeart(float cx, float cy, float width, int segments)
{
    const float pi = 3.14159265358979f, s = width / 32.0f, yMid = -2.5f;
    std::vector<Vec2> out;
    out.reserve(segments > 0 ? segments : 0);
    for (int i = 0; i < segments; ++i) {
        float t = 2.0f * pi * float(i) / float(segments);
        float st = std::sin(t);
        float hx = 16.0f * st * st * st;
        float hy = 13.0f * std::cos(t) - 5.0f * std::cos(2.0f * t) - 2.0f * std::cos(3.0f * t) - std::cos(4.0f * t);
        out.push_back(Vec2{cx + s * hx, cy - s * (hy - yMid)});
    }
    return out;
}

} // namespace voromesh
