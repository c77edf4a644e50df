// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
//
// BVH construction and traversal: restates src/accel/bvh_base.rs,
// src/accel/objects_split.rs, src/accel/bvh4.rs, src/algorithm.rs
// (partition, quick_select, merge_slices_append) and
// sub_crates/bvh_order (traversal table and code).
#pragma once

#include <cstdint>
#include <functional>
#include <utility>
#include <vector>

#include "psyref_bbox.h"
#include "psyref_ray.h"

namespace psyref {

static const size_t BVH_MAX_DEPTH = 42;    // bvh_base.rs:7
static const float USE_UNION_FACTOR = 1.4f;  // bvh_base.rs:12
static const size_t SAH_BIN_COUNT = 13;    // objects_split.rs:13

// Objects are identified by a u32 key; the bounder maps a key to its
// time-sampled bounds.  The builder permutes the key array in place, exactly
// as the reference permutes its object slice.
struct BoundsSlice { const BBox* ptr; size_t len; };
using Bounder = std::function<BoundsSlice(uint32_t)>;

// sub_crates/bvh_order/build.rs:5-59 — table[octant][code] packs the child
// visit order two bits per slot, first child in the low bits.
const uint8_t* traversal_table();  // [8][48]

// sub_crates/bvh_order/src/lib.rs:20-37
enum SplitKind { SPLIT_FULL, SPLIT_LEFT, SPLIT_RIGHT, SPLIT_TOP_ONLY };
static inline uint8_t calc_traversal_code(SplitKind k, uint8_t top, uint8_t a, uint8_t b) {
    switch (k) {
        case SPLIT_FULL: return top + a * 3 + b * 9;
        case SPLIT_LEFT: return top + a * 3 + 27;
        case SPLIT_RIGHT: return top + a * 3 + 27 + 9;
        default: return top + 27 + 9 + 9;
    }
}

// bvh4.rs:29-34
static inline size_t ray_code(Vec3 dir) {
    return (size_t)(dir.x < 0.0f) + ((size_t)(dir.y < 0.0f) << 1) + ((size_t)(dir.z < 0.0f) << 2);
}

// bvh_base.rs:23-45
struct BVHBaseNode {
    bool leaf;
    size_t bounds_start, bounds_end;
    size_t child0, child1;  // internal
    uint8_t split_axis;     // internal
    size_t obj_start, obj_end;  // leaf
};

// bvh_base.rs:14-21
struct BVHBase {
    std::vector<BVHBaseNode> nodes;
    std::vector<BBox> bounds;
    size_t depth = 0;
    std::vector<BBox> bounds_cache;

    static BVHBase from_objects(uint32_t* objects, size_t n, size_t objects_per_leaf, const Bounder& bounder);

   private:
    void acc_bounds(const uint32_t* objects, size_t n, const Bounder& bounder);
    std::pair<size_t, std::pair<size_t, size_t>> recursive_build(size_t offset, size_t depth, size_t per_leaf,
                                                                 uint32_t* objects, size_t n, const Bounder& b);
};

// algorithm.rs:44-81 and :203-231 (exposed for the reference's own tests)
size_t partition_u32(uint32_t* s, size_t n, const std::function<bool(uint32_t)>& pred);
void quick_select_u32(uint32_t* s, size_t n, size_t nth, const std::function<int(uint32_t, uint32_t)>& order);
void quick_select_i32(int32_t* s, size_t n, size_t nth);

// objects_split.rs:153-238, :291-341
std::pair<size_t, size_t> sah_split(uint32_t* objects, size_t n, const Bounder& bounder);
std::pair<size_t, size_t> median_split(uint32_t* objects, size_t n, const Bounder& bounder);

// bvh4.rs:37-55.  Nodes live in one vector; an internal node's children are a
// contiguous run (the reference allocates them as one arena slice).
struct BVH4Node {
    bool leaf;
    // internal
    std::vector<BBox4> bounds;  // time samples (bounds_len)
    size_t children_start = 0, children_len = 0;
    uint8_t traversal_code = 0;
    // leaf
    size_t obj_start = 0, obj_end = 0;
};

struct BVH4 {
    bool has_root = false;
    std::vector<BVH4Node> nodes;  // nodes[0] is the root when has_root
    size_t depth = 0;
    size_t node_count = 0;
    std::vector<BBox> root_bounds;  // bvh4.rs:89-92 (Boundable::bounds)

    // bvh4.rs:58-95
    static BVH4 from_objects(uint32_t* objects, size_t n, size_t objects_per_leaf, const Bounder& bounder);

    // bvh4.rs:101-184.  `leaf_fn(obj_start, obj_end)` runs the object tests for
    // the ray task on top of the stack (and must pop it).  `node_tests`
    // accumulates ray-node tests (ACCEL_NODE_RAY_TESTS, accel/mod.rs:22-25).
    template <typename LeafFn>
    void traverse(RayBatch& rays, RayStack& ray_stack, LeafFn&& leaf_fn, uint64_t& node_tests) const;

   private:
    size_t construct_from_base(const BVHBase& base, size_t base_node, size_t fill);
};

template <typename LeafFn>
void BVH4::traverse(RayBatch& rays, RayStack& ray_stack, LeafFn&& leaf_fn, uint64_t& node_tests) const {
    if (!has_root) return;
    const uint8_t* table = traversal_table() + 48 * ray_code(rays.dir_inv_local(ray_stack.next_task_ray_idx(0)));

    size_t node_stack[BVH_MAX_DEPTH * 3 + 2];
    node_stack[0] = 0;
    node_stack[1] = 0;  // root
    size_t stack_ptr = 1;

    while (stack_ptr > 0) {
        const BVH4Node& node = nodes[node_stack[stack_ptr]];
        if (!node.leaf) {
            node_tests += ray_stack.ray_count_in_next_task();
            unsigned all_hits = 0;
            const size_t nb = node.bounds.size();
            ray_stack.pop_do_next_task_and_push_rays(node.children_len, [&](size_t ray_idx) -> unsigned {
                if (rays.is_done(ray_idx)) return 0u;
                unsigned hits;
                if (nb == 1) {
                    hits = node.bounds[0].intersect_ray(rays.orig_local(ray_idx), rays.dir_inv_local(ray_idx),
                                                        rays.max_t(ray_idx));
                } else {
                    hits = lerp_slice(node.bounds.data(), nb, rays.time(ray_idx))
                               .intersect_ray(rays.orig_local(ray_idx), rays.dir_inv_local(ray_idx),
                                              rays.max_t(ray_idx));
                }
                all_hits |= hits;
                return hits;
            });

            if (all_hits != 0) {
                uint8_t order_code = table[node.traversal_code];
                size_t lane_count = 0;
                size_t i = node.children_len;
                while (i > 0) {
                    i -= 1;
                    size_t child_i = (order_code >> (i * 2)) & 3;
                    if (ray_stack.push_lane_to_task(child_i)) {
                        node_stack[stack_ptr + lane_count] = node.children_start + child_i;
                        lane_count += 1;
                    }
                }
                stack_ptr += lane_count - 1;
            } else {
                stack_ptr -= 1;
            }
        } else {
            leaf_fn(node.obj_start, node.obj_end);
            stack_ptr -= 1;
        }
    }
}

}  // namespace psyref
