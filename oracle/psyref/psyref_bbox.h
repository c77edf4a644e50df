// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
//
// Axis-aligned boxes: restates src/bbox.rs and src/bbox4.rs.
#pragma once

#include <limits>
#include <vector>

#include "psyref_math.h"

namespace psyref {

static const float BBOX_MAXT_ADJUST = 1.00000024f;  // bbox.rs:14, bbox4.rs:14
static const float F_INF = std::numeric_limits<float>::infinity();

// bbox.rs:9-29
struct BBox {
    Vec3 min, max;
    BBox() : min(F_INF, F_INF, F_INF), max(-F_INF, -F_INF, -F_INF) {}
    BBox(Vec3 mn, Vec3 mx) : min(mn), max(mx) {}

    // bbox.rs:35-49 (single-box slab test; used by the light tree)
    bool intersect_ray(Vec3 orig, Vec3 dir_inv, float max_t) const {
        Vec3 t1 = (min - orig) * dir_inv;
        Vec3 t2 = (max - orig) * dir_inv;
        Vec3 far4 = vmax(t1, t2);
        Vec3 near4 = vmin(t1, t2);
        // Vec4 reductions with the padding lane set to +inf / 0.
        float far_min = sse_min(sse_min(far4.x, far4.z), sse_min(far4.y, F_INF));
        float near_max = sse_max(sse_max(near4.x, near4.z), sse_max(near4.y, 0.0f));
        float far_hit_t = sse_min(far_min * BBOX_MAXT_ADJUST, max_t);
        return near_max <= far_hit_t;
    }

    // bbox.rs:52-75
    BBox transformed(const Transform& xform) const {
        Vec3 vs[8] = {
            Vec3(min.x, min.y, min.z), Vec3(min.x, min.y, max.z), Vec3(min.x, max.y, min.z),
            Vec3(min.x, max.y, max.z), Vec3(max.x, min.y, min.z), Vec3(max.x, min.y, max.z),
            Vec3(max.x, max.y, min.z), Vec3(max.x, max.y, max.z),
        };
        BBox b;
        for (int i = 0; i < 8; i++) {
            Vec3 v = xform_point(vs[i], xform);
            b.min = vmin(v, b.min);
            b.max = vmax(v, b.max);
        }
        return b;
    }

    // bbox.rs:77-80
    float surface_area() const {
        Vec3 d = max - min;
        return (((d.x * d.y) + (d.y * d.z)) + (d.z * d.x)) * 2.0f;
    }
    // bbox.rs:82-84: min.lerp(max, 0.5)
    Vec3 center() const { return lerp(min, max, 0.5f); }
    // bbox.rs:86-88
    float diagonal() const { return length(max - min); }
};

// BBox | BBox and BBox | Point (bbox.rs:96-137)
static inline BBox bunion(const BBox& a, const BBox& b) { return BBox(vmin(a.min, b.min), vmax(a.max, b.max)); }
static inline BBox bunion(const BBox& a, Vec3 p) { return BBox(vmin(a.min, p), vmax(a.max, p)); }
// Lerp for BBox (bbox.rs:139-146)
static inline BBox lerp(const BBox& a, const BBox& b, float alpha) {
    return BBox(lerp(a.min, b.min, alpha), lerp(a.max, b.max, alpha));
}

// transform_bbox_slice_from (bbox.rs:148-169)
static inline void transform_bbox_slice_from(const BBox* in, size_t n_in, const Transform* xf, size_t n_xf,
                                             std::vector<BBox>& out) {
    out.clear();
    if (n_xf == 0) {
        out.insert(out.end(), in, in + n_in);
    } else if (n_in == n_xf) {
        for (size_t i = 0; i < n_in; i++) out.push_back(in[i].transformed(xform_inverse(xf[i])));
    } else if (n_in > n_xf) {
        float s = (float)(n_in - 1);
        for (size_t i = 0; i < n_in; i++)
            out.push_back(in[i].transformed(xform_inverse(lerp_slice(xf, n_xf, (float)i / s))));
    } else {
        float s = (float)(n_xf - 1);
        for (size_t i = 0; i < n_xf; i++)
            out.push_back(lerp_slice(in, n_in, (float)i / s).transformed(xform_inverse(xf[i])));
    }
}

// bbox4.rs:16-21: four boxes, SoA per axis.
struct BBox4 {
    float xmin[4], xmax[4], ymin[4], ymax[4], zmin[4], zmax[4];

    BBox4() {
        for (int i = 0; i < 4; i++) {
            xmin[i] = ymin[i] = zmin[i] = F_INF;
            xmax[i] = ymax[i] = zmax[i] = -F_INF;
        }
    }
    // bbox4.rs:44-59
    static BBox4 from_bboxes(const BBox& b1, const BBox& b2, const BBox& b3, const BBox& b4) {
        BBox4 r;
        const BBox* bs[4] = {&b1, &b2, &b3, &b4};
        for (int i = 0; i < 4; i++) {
            r.xmin[i] = bs[i]->min.x; r.xmax[i] = bs[i]->max.x;
            r.ymin[i] = bs[i]->min.y; r.ymax[i] = bs[i]->max.y;
            r.zmin[i] = bs[i]->min.z; r.zmax[i] = bs[i]->max.z;
        }
        return r;
    }

    // bbox4.rs:62-97: returns a 4-bit hit mask.
    unsigned intersect_ray(Vec3 orig, Vec3 dir_inv, float max_t) const {
        unsigned mask = 0;
        for (int i = 0; i < 4; i++) {
            float t1x = (xmin[i] - orig.x) * dir_inv.x;
            float t1y = (ymin[i] - orig.y) * dir_inv.y;
            float t1z = (zmin[i] - orig.z) * dir_inv.z;
            float t2x = (xmax[i] - orig.x) * dir_inv.x;
            float t2y = (ymax[i] - orig.y) * dir_inv.y;
            float t2z = (zmax[i] - orig.z) * dir_inv.z;
            float tfx = sse_max(t1x, t2x), tfy = sse_max(t1y, t2y), tfz = sse_max(t1z, t2z);
            float tnx = sse_min(t1x, t2x), tny = sse_min(t1y, t2y), tnz = sse_min(t1z, t2z);
            float far_t = sse_min(sse_min(tfx, sse_min(tfy, tfz)) * BBOX_MAXT_ADJUST, max_t);
            float near_t = sse_max(sse_max(tnx, tny), sse_max(tnz, 0.0f));
            if (near_t < far_t) mask |= 1u << i;
        }
        return mask;
    }
};

// Lerp for BBox4 (bbox4.rs:119-136)
static inline BBox4 lerp(const BBox4& a, const BBox4& b, float alpha) {
    BBox4 r;
    for (int i = 0; i < 4; i++) {
        r.xmin[i] = lerp(a.xmin[i], b.xmin[i], alpha); r.xmax[i] = lerp(a.xmax[i], b.xmax[i], alpha);
        r.ymin[i] = lerp(a.ymin[i], b.ymin[i], alpha); r.ymax[i] = lerp(a.ymax[i], b.ymax[i], alpha);
        r.zmin[i] = lerp(a.zmin[i], b.zmin[i], alpha); r.zmax[i] = lerp(a.zmax[i], b.zmax[i], alpha);
    }
    return r;
}

}  // namespace psyref
