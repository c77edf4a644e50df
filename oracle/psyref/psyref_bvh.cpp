// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
//
// BVHBase build (bvh_base.rs), splits (objects_split.rs), algorithm helpers
// (algorithm.rs), BVH4 collapse (bvh4.rs:186-358) and the traversal-order
// table (sub_crates/bvh_order/build.rs).
#include "psyref_bvh.h"

#include <algorithm>

namespace psyref {

// ---- sub_crates/bvh_order/build.rs:5-59 -----------------------------------
namespace {
struct TableHolder {
    uint8_t t[8][48];
    TableHolder() {
        for (int raydir = 0; raydir < 8; raydir++) {
            const int ray[3] = {raydir & 1, (raydir >> 1) & 1, (raydir >> 2) & 1};
            int k = 0;
            // Full: s0 = top, s1 = left, s2 = right  (build.rs:12-27)
            for (int s2 = 0; s2 < 3; s2++)
                for (int s1 = 0; s1 < 3; s1++)
                    for (int s0 = 0; s0 < 3; s0++) {
                        int p[4] = {0, 1, 2, 3};
                        if (ray[s1]) std::swap(p[0], p[1]);
                        if (ray[s2]) std::swap(p[2], p[3]);
                        if (ray[s0]) { std::swap(p[0], p[2]); std::swap(p[1], p[3]); }
                        t[raydir][k++] = (uint8_t)(p[0] + (p[1] << 2) + (p[2] << 4) + (p[3] << 6));
                    }
            // Left: three children, left pair split  (build.rs:29-40)
            for (int s1 = 0; s1 < 3; s1++)
                for (int s0 = 0; s0 < 3; s0++) {
                    int p[3] = {0, 1, 2};
                    if (ray[s1]) std::swap(p[0], p[1]);
                    if (ray[s0]) { std::swap(p[0], p[1]); std::swap(p[0], p[2]); }
                    t[raydir][k++] = (uint8_t)(p[0] + (p[1] << 2) + (p[2] << 4));
                }
            // Right: three children, right pair split  (build.rs:42-53)
            for (int s1 = 0; s1 < 3; s1++)
                for (int s0 = 0; s0 < 3; s0++) {
                    int p[3] = {0, 1, 2};
                    if (ray[s1]) std::swap(p[1], p[2]);
                    if (ray[s0]) { std::swap(p[0], p[2]); std::swap(p[0], p[1]); }
                    t[raydir][k++] = (uint8_t)(p[0] + (p[1] << 2) + (p[2] << 4));
                }
            // TopOnly  (build.rs:55-61)
            for (int s0 = 0; s0 < 3; s0++) {
                int p[2] = {0, 1};
                if (ray[s0]) std::swap(p[0], p[1]);
                t[raydir][k++] = (uint8_t)(p[0] + (p[1] << 2));
            }
        }
    }
};
}  // namespace

const uint8_t* traversal_table() {
    static TableHolder h;
    return &h.t[0][0];
}

// ---- algorithm.rs:44-81 ----------------------------------------------------
size_t partition_u32(uint32_t* s, size_t n, const std::function<bool(uint32_t)>& pred) {
    size_t a = 0, b = n;
    for (;;) {
        for (;;) {
            if (a == b) return a;
            if (!pred(s[a])) break;
            a++;
        }
        for (;;) {
            b--;
            if (a == b) return a;
            if (pred(s[b])) break;
        }
        std::swap(s[a], s[b]);
        a++;
    }
}

// algorithm.rs:203-231
template <typename T, typename Less>
static void quick_select_impl(T* s, size_t n, size_t nth, Less&& less) {
    size_t left = 0, right = n;
    uint64_t seed = (uint64_t)nth;
    for (;;) {
        size_t i = left + (size_t)(hash_u64((uint64_t)right, seed) % (uint64_t)(right - left));
        std::swap(s[i], s[right - 1]);
        // partition(slc[left..right-1], |x| order(x, val) == Less)
        T val = s[right - 1];
        size_t a = left, b = right - 1, ii;
        for (;;) {
            bool done = false;
            for (;;) {
                if (a == b) { ii = a; done = true; break; }
                if (!less(s[a], val)) break;
                a++;
            }
            if (done) break;
            for (;;) {
                b--;
                if (a == b) { ii = a; done = true; break; }
                if (less(s[b], val)) break;
            }
            if (done) break;
            std::swap(s[a], s[b]);
            a++;
        }
        std::swap(s[ii], s[right - 1]);
        if (ii == nth) return;
        if (ii > nth) right = ii; else left = ii + 1;
        seed += 1;
    }
}

void quick_select_u32(uint32_t* s, size_t n, size_t nth, const std::function<int(uint32_t, uint32_t)>& order) {
    quick_select_impl(s, n, nth, [&](uint32_t a, uint32_t b) { return order(a, b) < 0; });
}
void quick_select_i32(int32_t* s, size_t n, size_t nth) {
    quick_select_impl(s, n, nth, [](int32_t a, int32_t b) { return a < b; });
}

// ---- objects_split.rs:153-238 ----------------------------------------------
static inline BBox mid_bounds(const Bounder& bounder, uint32_t obj) {
    BoundsSlice bs = bounder(obj);
    return lerp_slice(bs.ptr, bs.len, 0.5f);
}

std::pair<size_t, size_t> sah_split(uint32_t* objects, size_t n, const Bounder& bounder) {
    // Combined centroid extents (:160-167)
    BBox bounds;
    for (size_t i = 0; i < n; i++) bounds = bunion(bounds, mid_bounds(bounder, objects[i]).center());

    // Division points (:169-180)
    float sah_divs[3][SAH_BIN_COUNT - 1];
    for (int d = 0; d < 3; d++) {
        float extent = bounds.max[d] - bounds.min[d];
        for (size_t div = 0; div < SAH_BIN_COUNT - 1; div++) {
            float part = extent * ((float)(div + 1) / (float)SAH_BIN_COUNT);
            sah_divs[d][div] = bounds.min[d] + part;
        }
    }

    // Bins (:182-202)
    BBox bin_l[3][SAH_BIN_COUNT - 1], bin_r[3][SAH_BIN_COUNT - 1];
    size_t cnt_l[3][SAH_BIN_COUNT - 1] = {}, cnt_r[3][SAH_BIN_COUNT - 1] = {};
    for (size_t i = 0; i < n; i++) {
        BBox tb = mid_bounds(bounder, objects[i]);
        Vec3 centroid = (tb.min + tb.max) * 0.5f;
        for (int d = 0; d < 3; d++)
            for (size_t div = 0; div < SAH_BIN_COUNT - 1; div++) {
                if (centroid[d] <= sah_divs[d][div]) {
                    bin_l[d][div] = bunion(bin_l[d][div], tb);
                    cnt_l[d][div] += 1;
                } else {
                    bin_r[d][div] = bunion(bin_r[d][div], tb);
                    cnt_r[d][div] += 1;
                }
            }
    }

    // Best split (:204-226)
    int dim = 0;
    float div_n = 0.0f;
    float smallest_cost = F_INF;
    for (int d = 0; d < 3; d++)
        for (size_t div = 0; div < SAH_BIN_COUNT - 1; div++) {
            float left_cost = bin_l[d][div].surface_area() * (float)cnt_l[d][div];
            float right_cost = bin_r[d][div].surface_area() * (float)cnt_r[d][div];
            float left_diag = bin_l[d][div].diagonal();
            float right_diag = bin_r[d][div].diagonal();
            float tot_cost = (left_cost * left_diag) + (right_cost * right_diag);
            if (tot_cost < smallest_cost) {
                dim = d;
                div_n = sah_divs[d][div];
                smallest_cost = tot_cost;
            }
        }

    // Partition (:228-237)
    size_t split_i = partition_u32(objects, n, [&](uint32_t obj) {
        BBox tb = mid_bounds(bounder, obj);
        float centroid = (tb.min[dim] + tb.max[dim]) * 0.5f;
        return centroid < div_n;
    });
    if (split_i < 1) split_i = 1;
    else if (split_i >= n) split_i = n - 1;
    return {split_i, (size_t)dim};
}

// objects_split.rs:291-341
std::pair<size_t, size_t> median_split(uint32_t* objects, size_t n, const Bounder& bounder) {
    BBox bounds;
    for (size_t i = 0; i < n; i++) bounds = bunion(bounds, mid_bounds(bounder, objects[i]));
    int axis = 0;
    float largest = -F_INF;
    for (int i = 0; i < 3; i++) {
        float extent = bounds.max[i] - bounds.min[i];
        if (extent > largest) { largest = extent; axis = i; }
    }
    size_t place = n / 2;
    if (place == 0) place = 1;
    quick_select_u32(objects, n, place, [&](uint32_t a, uint32_t b) {
        BBox ta = mid_bounds(bounder, a), tb = mid_bounds(bounder, b);
        float ca = (ta.min[axis] + ta.max[axis]) * 0.5f;
        float cb = (tb.min[axis] + tb.max[axis]) * 0.5f;
        if (ca < cb) return -1;
        if (ca == cb) return 0;
        return 1;
    });
    return {place, (size_t)axis};
}

// ---- bvh_base.rs -----------------------------------------------------------
BVHBase BVHBase::from_objects(uint32_t* objects, size_t n, size_t objects_per_leaf, const Bounder& bounder) {
    BVHBase bvh;
    bvh.recursive_build(0, 0, objects_per_leaf, objects, n, bounder);
    return bvh;
}

// bvh_base.rs:64-88
void BVHBase::acc_bounds(const uint32_t* objects, size_t n, const Bounder& bounder) {
    size_t max_len = 0;
    for (size_t i = 0; i < n; i++) max_len = std::max(max_len, bounder(objects[i]).len);
    bounds_cache.assign(max_len, BBox());
    for (size_t o = 0; o < n; o++) {
        BoundsSlice bs = bounder(objects[o]);
        if (bs.len == max_len) {
            for (size_t i = 0; i < bs.len; i++) bounds_cache[i] = bunion(bounds_cache[i], bs.ptr[i]);
        } else {
            float s = (float)(max_len - 1);
            for (size_t i = 0; i < max_len; i++)
                bounds_cache[i] = bunion(bounds_cache[i], lerp_slice(bs.ptr, bs.len, (float)i / s));
        }
    }
}

// The union-vs-time-samples choice (bvh_base.rs:116-132 and :194-203).
static void push_bounds_or_union(std::vector<BBox>& out, const std::vector<BBox>& samples) {
    BBox union_bounds;
    for (const BBox& b : samples) union_bounds = bunion(union_bounds, b);
    float area = 0.0f;
    for (const BBox& b : samples) area = area + b.surface_area();
    float average_area = area / (float)samples.size();
    if (union_bounds.surface_area() <= (average_area * USE_UNION_FACTOR)) {
        out.push_back(union_bounds);
    } else {
        out.insert(out.end(), samples.begin(), samples.end());
    }
}

// algorithm.rs:233-262 with merge = BBox union
static void merge_bounds_append(const BBox* s1, size_t n1, const BBox* s2, size_t n2, std::vector<BBox>& out) {
    if (n1 == 0 || n2 == 0) return;
    if (n1 == n2) {
        for (size_t i = 0; i < n1; i++) out.push_back(bunion(s1[i], s2[i]));
    } else if (n1 > n2) {
        float s = (float)(n1 - 1);
        for (size_t i = 0; i < n1; i++) out.push_back(bunion(s1[i], lerp_slice(s2, n2, (float)i / s)));
    } else {
        float s = (float)(n2 - 1);
        for (size_t i = 0; i < n2; i++) out.push_back(bunion(lerp_slice(s1, n1, (float)i / s), s2[i]));
    }
}

// bvh_base.rs:90-215
std::pair<size_t, std::pair<size_t, size_t>> BVHBase::recursive_build(size_t offset, size_t depth_, size_t per_leaf,
                                                                     uint32_t* objects, size_t n,
                                                                     const Bounder& bounder) {
    size_t me = nodes.size();
    if (n == 0) return {0, {0, 0}};
    if (n <= per_leaf) {
        size_t bi = bounds.size();
        acc_bounds(objects, n, bounder);
        push_bounds_or_union(bounds, bounds_cache);
        BVHBaseNode node{};
        node.leaf = true;
        node.bounds_start = bi;
        node.bounds_end = bounds.size();
        node.obj_start = offset;
        node.obj_end = offset + n;
        nodes.push_back(node);
        if (depth < depth_) depth = depth_;
        return {me, {bi, bounds.size()}};
    }
    nodes.push_back(BVHBaseNode{});
    std::pair<size_t, size_t> split;
    if ((size_t)log2_64((uint64_t)n) < (BVH_MAX_DEPTH - depth_))
        split = sah_split(objects, n, bounder);
    else
        split = median_split(objects, n, bounder);
    size_t split_index = split.first;

    auto c1 = recursive_build(offset, depth_ + 1, per_leaf, objects, split_index, bounder);
    auto c2 = recursive_build(offset + split_index, depth_ + 1, per_leaf, objects + split_index, n - split_index,
                              bounder);

    size_t bi = bounds.size();
    std::vector<BBox> merged;
    merge_bounds_append(bounds.data() + c1.second.first, c1.second.second - c1.second.first,
                        bounds.data() + c2.second.first, c2.second.second - c2.second.first, merged);
    push_bounds_or_union(bounds, merged);

    BVHBaseNode node{};
    node.leaf = false;
    node.bounds_start = bi;
    node.bounds_end = bounds.size();
    node.child0 = c1.first;
    node.child1 = c2.first;
    node.split_axis = (uint8_t)split.second;
    nodes[me] = node;
    return {me, {bi, bounds.size()}};
}

// ---- bvh4.rs ---------------------------------------------------------------
BVH4 BVH4::from_objects(uint32_t* objects, size_t n, size_t objects_per_leaf, const Bounder& bounder) {
    BVH4 bvh;
    if (n == 0) return bvh;
    BVHBase base = BVHBase::from_objects(objects, n, objects_per_leaf, bounder);
    bvh.has_root = true;
    bvh.nodes.emplace_back();
    bvh.node_count = bvh.construct_from_base(base, 0, 0);
    bvh.depth = (base.depth / 2) + 1;
    const BVHBaseNode& r = base.nodes[0];
    bvh.root_bounds.assign(base.bounds.begin() + r.bounds_start, base.bounds.begin() + r.bounds_end);
    return bvh;
}

// bvh4.rs:186-358
size_t BVH4::construct_from_base(const BVHBase& base, size_t bn, size_t fill) {
    size_t node_count = 0;
    const BVHBaseNode& node = base.nodes[bn];
    if (node.leaf) {
        nodes[fill].leaf = true;
        nodes[fill].obj_start = node.obj_start;
        nodes[fill].obj_end = node.obj_end;
        return 1;
    }
    const BVHBaseNode& cl = base.nodes[node.child0];
    const BVHBaseNode& cr = base.nodes[node.child1];
    size_t child_count;
    const BVHBaseNode* children[4] = {nullptr, nullptr, nullptr, nullptr};
    uint8_t code;
    if (!cl.leaf) {
        if (!cr.leaf) {
            child_count = 4;
            children[0] = &base.nodes[cl.child0];
            children[1] = &base.nodes[cl.child1];
            children[2] = &base.nodes[cr.child0];
            children[3] = &base.nodes[cr.child1];
            code = calc_traversal_code(SPLIT_FULL, node.split_axis, cl.split_axis, cr.split_axis);
        } else {
            child_count = 3;
            children[0] = &base.nodes[cl.child0];
            children[1] = &base.nodes[cl.child1];
            children[2] = &cr;
            code = calc_traversal_code(SPLIT_LEFT, node.split_axis, cl.split_axis, 0);
        }
    } else {
        if (!cr.leaf) {
            child_count = 3;
            children[0] = &cl;
            children[1] = &base.nodes[cr.child0];
            children[2] = &base.nodes[cr.child1];
            code = calc_traversal_code(SPLIT_RIGHT, node.split_axis, cr.split_axis, 0);
        } else {
            child_count = 2;
            children[0] = &cl;
            children[1] = &cr;
            code = calc_traversal_code(SPLIT_TOP_ONLY, node.split_axis, 0, 0);
        }
    }
    node_count += child_count;

    // Bounds (:272-327)
    size_t bounds_len = 0;
    for (int i = 0; i < 4; i++)
        if (children[i]) bounds_len = std::max(bounds_len, children[i]->bounds_end - children[i]->bounds_start);
    std::vector<BBox4> b4(bounds_len);
    auto child_bound = [&](int i, float time) -> BBox {
        if (!children[i]) return BBox();
        return lerp_slice(base.bounds.data() + children[i]->bounds_start,
                          children[i]->bounds_end - children[i]->bounds_start, time);
    };
    if (bounds_len < 2) {
        BBox bb[4];
        for (int i = 0; i < 4; i++)
            bb[i] = children[i] ? base.bounds[children[i]->bounds_start] : BBox();
        b4[0] = BBox4::from_bboxes(bb[0], bb[1], bb[2], bb[3]);
    } else {
        for (size_t s = 0; s < bounds_len; s++) {
            float time = (float)s / (float)(bounds_len - 1);
            b4[s] = BBox4::from_bboxes(child_bound(0, time), child_bound(1, time), child_bound(2, time),
                                       child_bound(3, time));
        }
    }

    // Children (:329-334), then this node.
    size_t cstart = nodes.size();
    nodes.resize(nodes.size() + child_count);
    for (size_t i = 0; i < child_count; i++)
        node_count += construct_from_base(base, (size_t)(children[i] - base.nodes.data()), cstart + i);

    BVH4Node& me = nodes[fill];
    me.leaf = false;
    me.bounds = std::move(b4);
    me.children_start = cstart;
    me.children_len = child_count;
    me.traversal_code = code;
    return node_count;
}

}  // namespace psyref
