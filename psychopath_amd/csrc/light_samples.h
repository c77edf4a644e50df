// Time-sampled light parameters: the light records of layout.h resolved at a
// ray's time.  The reference lerps each list of a light on its own with
// lerp_slice (src/lerp.rs:20-35) wherever it uses it: dimensions, radius,
// direction and colour (rectangle_light.rs:64,131-132,275,318;
// sphere_light.rs:66,102-103,224,324; distant_disk_light.rs:67-69).  Colours
// lerp component-wise within one representation (color.rs:284-333).
#pragma once

#include "../../include/psy_rt.h"
#include "layout.h"
#include "psymath.h"

namespace psy {

PSY_HD float lerp_list(const float* s, uint32_t n, float time) {
    uint32_t i1;
    float a2;
    if (!lerp_slice_pos(n, time, i1, a2)) return s[i1];
    return lerpf(s[i1], s[i1 + 1], a2);
}

// A list of n colour samples (four floats each: kind, a, b, c) at `time`.
PSY_HD GpuColor color_list(const float* s, uint32_t n, float time) {
    uint32_t i1;
    float a2;
    const bool mix = lerp_slice_pos(n, time, i1, a2);
    const float* x = s + 4 * i1;
    GpuColor c{(uint32_t)x[0], x[1], x[2], x[3]};
    if (!mix) return c;
    const float* y = x + 4;
    const float inv = 1.0f - a2;
    c.a = (x[1] * inv) + (y[1] * a2);
    c.b = (x[2] * inv) + (y[2] * a2);
    if (c.kind == PSY_COLOR_XYZ) c.c = (x[3] * inv) + (y[3] * a2);
    return c;
}

PSY_HD GpuRectLight rect_at(GpuRectLight L, const float* ls, float time) {
    if (L.counts == 0) return L;
    const uint32_t nd = L.counts & 0xffu, nc = (L.counts >> 8) & 0xffu;
    const float* s = ls + L.samples;
    if (nd > 0) {
        // (w, h) pairs: Lerp for (T, T) lerps each element (lerp.rs:70-75).
        uint32_t i1;
        float a2;
        if (!lerp_slice_pos(nd, time, i1, a2)) {
            L.w = s[2 * i1];
            L.h = s[2 * i1 + 1];
        } else {
            L.w = lerpf(s[2 * i1], s[2 * i1 + 2], a2);
            L.h = lerpf(s[2 * i1 + 1], s[2 * i1 + 3], a2);
        }
    }
    if (nc > 0) L.color = color_list(s + 2 * nd, nc, time);
    return L;
}

PSY_HD GpuSphereLight sphere_at(GpuSphereLight L, const float* ls, float time) {
    if (L.counts == 0) return L;
    const uint32_t nr = L.counts & 0xffu, nc = (L.counts >> 8) & 0xffu;
    const float* s = ls + L.samples;
    if (nr > 0) L.radius = lerp_list(s, nr, time);
    if (nc > 0) L.color = color_list(s + nr, nc, time);
    return L;
}

PSY_HD GpuDistantLight distant_at(GpuDistantLight L, const float* ls, float time) {
    if (L.counts == 0) return L;
    const uint32_t nr = L.counts & 0xffu, nd = (L.counts >> 8) & 0xffu, nc = (L.counts >> 16) & 0xffu;
    const float* s = ls + L.samples;
    if (nr > 0) L.radius = lerp_list(s, nr, time);
    if (nd > 0) {
        // Lerp for Vector: (a * (1 - alpha)) + (b * alpha) per component.
        const float* d = s + nr;
        uint32_t i1;
        float a2;
        if (!lerp_slice_pos(nd, time, i1, a2)) {
            L.dir[0] = d[3 * i1];
            L.dir[1] = d[3 * i1 + 1];
            L.dir[2] = d[3 * i1 + 2];
        } else {
            const V3 v = lerp3(v3(d[3 * i1], d[3 * i1 + 1], d[3 * i1 + 2]),
                               v3(d[3 * i1 + 3], d[3 * i1 + 4], d[3 * i1 + 5]), a2);
            L.dir[0] = v.x;
            L.dir[1] = v.y;
            L.dir[2] = v.z;
        }
    }
    if (nc > 0) L.color = color_list(s + nr + 3 * nd, nc, time);
    return L;
}

}  // namespace psy
