// Builder arithmetic shared by the host builder (builder.cpp) and the device
// builder (builder_device.hip): the per-object boxes and centroids, the SAH
// binning and split choice, the Hoare partition, the hash of quick_select and
// the union-vs-time-samples rule.  Both builders call these same functions,
// so they build the same trees bit for bit.
//   sah_split            src/accel/objects_split.rs:153-238
//   partition            src/algorithm.rs:44-81
//   quick_select / hash  src/algorithm.rs:203-231, src/hash.rs:12-21
//   union rule           src/accel/bvh_base.rs:116-132, 194-203
#pragma once

#include <stdint.h>

#include "psymath.h"

namespace psy {

constexpr uint32_t SAH_BINS = 13;          // objects_split.rs:13
constexpr uint32_t MAX_BASE_DEPTH = 42;    // bvh_base.rs:7
constexpr float UNION_FACTOR = 1.4f;       // bvh_base.rs:12

// Per-object data derived once from the input bounds.
struct ObjInfo {
    Box mid;        // lerp_slice(bounds, 0.5)
    V3 mid_center;  // mid.center()  (BBox::center, a lerp)
    V3 centroid;    // (mid.min + mid.max) * 0.5
};

PSY_HD Box lerp_slice_box(const Box* s, uint32_t n, float alpha) {
    uint32_t i1;
    float a2;
    if (!lerp_slice_pos(n, alpha, i1, a2)) return s[n - 1];
    return box_lerp(s[i1], s[i1 + 1], a2);
}

PSY_HD ObjInfo obj_info(const Box* s, uint32_t n) {
    ObjInfo o;
    o.mid = lerp_slice_box(s, n, 0.5f);
    o.mid_center = box_center(o.mid);
    o.centroid = mul(add(o.mid.lo, o.mid.hi), 0.5f);
    return o;
}

// Without NaN or negative zero in an object's derived boxes and centroids,
// min/max unions give the same box in any order, so SAH candidates can be
// accumulated per bin and in any grouping.
PSY_HD bool obj_order_free(const ObjInfo& o) {
    const float v[12] = {o.mid.lo.x, o.mid.lo.y, o.mid.lo.z, o.mid.hi.x, o.mid.hi.y, o.mid.hi.z,
                         o.mid_center.x, o.mid_center.y, o.mid_center.z, o.centroid.x, o.centroid.y, o.centroid.z};
    bool ok = true;
    for (int i = 0; i < 12; i++) {
        const float f = v[i];
        if (f != f || (f == 0.0f && signbit(f))) ok = false;
    }
    return ok;
}

// The SAH dividers of one node from its centroid bounds (mid_center).
PSY_HD void sah_divs(const Box& cb, float (&divs)[3][SAH_BINS - 1]) {
    for (int d = 0; d < 3; d++) {
        const float lo = comp(cb.lo, d), extent = comp(cb.hi, d) - lo;
        for (uint32_t k = 0; k < SAH_BINS - 1; k++) divs[d][k] = lo + extent * ((float)(k + 1) / (float)SAH_BINS);
    }
}

// Bin of a centroid coordinate: the first k whose divider it does not exceed
// (c <= divs[k]), or the last bin.
PSY_HD uint32_t sah_bin(float c, const float (&divs)[SAH_BINS - 1]) {
    uint32_t k = SAH_BINS - 1;
    for (int j = (int)SAH_BINS - 2; j >= 0; j--)
        if (c <= divs[j]) k = (uint32_t)j;
    return k;
}

// The split choice from per-bin boxes and counts: left and right sides of
// each candidate, then the smallest (area * count * diagonal) cost in the
// reference's candidate order.
PSY_HD void sah_choose(const Box (&box)[3][SAH_BINS], const uint32_t (&count)[3][SAH_BINS],
                       const float (&divs)[3][SAH_BINS - 1], int& best_d, float& best_div) {
    best_d = 0;
    best_div = 0.0f;
    float best = __builtin_huge_valf();
    for (int d = 0; d < 3; d++) {
        Box left[SAH_BINS - 1], right[SAH_BINS - 1];
        uint32_t nl[SAH_BINS - 1], nr[SAH_BINS - 1];
        Box acc = box_empty();
        uint32_t cnt = 0;
        for (uint32_t k = 0; k < SAH_BINS - 1; k++) {
            acc = box_union(acc, box[d][k]);
            cnt += count[d][k];
            left[k] = acc;
            nl[k] = cnt;
        }
        acc = box_empty();
        cnt = 0;
        for (int k = (int)SAH_BINS - 2; k >= 0; k--) {
            acc = box_union(acc, box[d][k + 1]);
            cnt += count[d][k + 1];
            right[k] = acc;
            nr[k] = cnt;
        }
        for (uint32_t k = 0; k < SAH_BINS - 1; k++) {
            const float lc = box_area(left[k]) * (float)nl[k];
            const float rc = box_area(right[k]) * (float)nr[k];
            const float cost = (lc * box_diag(left[k])) + (rc * box_diag(right[k]));
            if (cost < best) {
                best_d = d;
                best_div = divs[d][k];
                best = cost;
            }
        }
    }
}

// The SAH predicate: an object goes left when its centroid lies below the
// chosen divider on the chosen axis.
PSY_HD bool sah_left(const ObjInfo& o, int d, float div) { return comp(o.centroid, d) < div; }

// hash_u64 (src/hash.rs:12-21)
PSY_HD uint64_t hash64(uint64_t n, uint64_t seed) {
    uint64_t h = n;
    for (int i = 0; i < 4; i++) {
        h *= 0x8d879e7eb3c956e9ull;  // 32_416_190_071 * 314_604_959
        h ^= h >> 32;
        h ^= seed;
    }
    return h;
}

template <typename T>
PSY_HD void swap_vals(T& a, T& b) {
    T t = a;
    a = b;
    b = t;
}

// Hoare-style partition of algorithm.rs:44-81 over the keys and their
// records together; returns the index of the first object for which `pred`
// is false.
template <typename Pred>
PSY_HD uint32_t hoare_partition(uint32_t* s, ObjInfo* r, uint32_t n, Pred pred) {
    uint32_t a = 0, b = n;
    for (;;) {
        while (true) {
            if (a == b) return a;
            if (!pred(r[a])) break;
            a++;
        }
        while (true) {
            b--;
            if (a == b) return a;
            if (pred(r[b])) break;
        }
        swap_vals(s[a], s[b]);
        swap_vals(r[a], r[b]);
        a++;
    }
}

// Whether a node of n objects at this depth splits by SAH (room left under
// the depth limit) or by median (bvh_base.rs:156-165).
PSY_HD bool use_sah(uint32_t n, uint32_t depth) {
    uint32_t lg = 0;
    for (uint32_t v = n; v > 1; v >>= 1) lg++;  // log2_64, src/math.rs:168-187
    return lg < MAX_BASE_DEPTH - depth;
}

// median_split (objects_split.rs:291-341) with quick_select
// (algorithm.rs:203-231), in place over keys/recs[0, n); n >= 2.
PSY_HD void median_split(uint32_t* keys, ObjInfo* recs, uint32_t n, uint32_t& split, uint8_t& axis_out) {
    Box b = box_empty();
    for (uint32_t i = 0; i < n; i++) b = box_union(b, recs[i].mid);
    int axis = 0;
    float largest = -__builtin_huge_valf();
    for (int d = 0; d < 3; d++) {
        const float e = comp(b.hi, d) - comp(b.lo, d);
        if (e > largest) {
            largest = e;
            axis = d;
        }
    }
    uint32_t place = n / 2;
    if (place == 0) place = 1;

    uint32_t left = 0, right = n;
    uint64_t seed = place;
    for (;;) {
        const uint32_t i = left + (uint32_t)(hash64(right, seed) % (uint64_t)(right - left));
        swap_vals(keys[i], keys[right - 1]);
        swap_vals(recs[i], recs[right - 1]);
        const float pivot = comp(recs[right - 1].centroid, axis);
        const uint32_t ii = left + hoare_partition(keys + left, recs + left, right - 1 - left,
                                                   [=](const ObjInfo& o) { return comp(o.centroid, axis) < pivot; });
        swap_vals(keys[ii], keys[right - 1]);
        swap_vals(recs[ii], recs[right - 1]);
        if (ii == place) break;
        if (ii > place) right = ii;
        else left = ii + 1;
        seed += 1;
    }
    split = place;
    axis_out = (uint8_t)axis;
}

// The union rule over `n` time-sample boxes: keep one union box unless it is
// more than 1.4x the mean sample area.  Returns the kept count (1 or n);
// with 1 the union is written to out[0], else the samples are copied.
PSY_HD uint32_t union_or_samples(const Box* samples, uint32_t n, Box* out) {
    Box u = box_empty();
    for (uint32_t i = 0; i < n; i++) u = box_union(u, samples[i]);
    float area = 0.0f;
    for (uint32_t i = 0; i < n; i++) area = area + box_area(samples[i]);
    const float avg = area / (float)n;
    if (box_area(u) <= (avg * UNION_FACTOR)) {
        out[0] = u;
        return 1;
    }
    for (uint32_t i = 0; i < n; i++) out[i] = samples[i];
    return n;
}

}  // namespace psy
