// ref_canvas_shim.cpp -- test infrastructure only: a C interface over the
// reference's own canvas (src/voromesh/canvas.cpp) and heart outline
// (src/voromesh/shape2d.cpp), compiled in place from the reference sources by
// oracle/Makefile into oracle/_ref/libref_canvas.so.  The golden-fixture script
// (tests/golden/make_canvas_golden.py) and the canvas parity tests call it to
// check the product's canvas; the product never links it.
#include <cstdint>
#include <vector>

#include "voromesh/canvas.h"
#include "voromesh/shape2d.h"

using voromesh::Canvas;
using voromesh::Vec2;

extern "C" {

void *ref_canvas_create(float w, float h, float voxel) { return new Canvas(w, h, voxel); }
void  ref_canvas_destroy(void *c) { delete static_cast<Canvas *>(c); }

void ref_canvas_rasterize(void *c, const float *xy, int n)
{
    std::vector<Vec2> poly(size_t(n > 0 ? n : 0));
    for (int i = 0; i < n; ++i) poly[size_t(i)] = Vec2{xy[2 * i], xy[2 * i + 1]};
    static_cast<Canvas *>(c)->RasterizeShape(poly);
}

// edit[7] = {tx0, ty0, tx1, ty1, removed, added, flipped}; returns the dirty
// count (all of them copied when it fits `cap`).
int ref_canvas_stamp(void *c, float cx, float cy, float r, int inside, int32_t *edit, int32_t *dirty, int cap)
{
    Canvas::Edit e = static_cast<Canvas *>(c)->StampDisc(Vec2{cx, cy}, r, inside != 0);
    const int32_t f[7] = {e.tx0, e.ty0, e.tx1, e.ty1, e.removed, e.added, e.flipped};
    for (int i = 0; i < 7; ++i) edit[i] = f[i];
    const int n = int(e.dirtySlots.size());
    for (int i = 0; i < n && i < cap; ++i) dirty[i] = e.dirtySlots[size_t(i)];
    return n;
}

int64_t ref_canvas_slots(void *c, float *xy, int64_t cap)
{
    const auto &s = static_cast<Canvas *>(c)->TargetSlots();
    for (size_t i = 0; i < s.size() && int64_t(i) < cap; ++i) xy[2 * i] = s[i].x, xy[2 * i + 1] = s[i].y;
    return int64_t(s.size());
}

int64_t ref_canvas_masks(void *c, uint64_t *m, int64_t cap)
{
    const auto &s = static_cast<Canvas *>(c)->InsideMasks();
    for (size_t i = 0; i < s.size() && int64_t(i) < cap; ++i) m[i] = s[i];
    return int64_t(s.size());
}

void ref_canvas_counts(void *c, int64_t *out)
{
    const Canvas *k = static_cast<Canvas *>(c);
    out[0] = int64_t(k->TargetSlots().size());
    out[1] = int64_t(k->SurfaceCount());
    out[2] = int64_t(k->SurfaceTileCount());
    out[3] = k->TilesX(), out[4] = k->TilesY();
}

void ref_make_heart(float cx, float cy, float width, int segments, float *xy)
{
    const std::vector<Vec2> p = voromesh::MakeHeart(cx, cy, width, segments);
    for (size_t i = 0; i < p.size(); ++i) xy[2 * i] = p[i].x, xy[2 * i + 1] = p[i].y;
}

} // extern "C"
