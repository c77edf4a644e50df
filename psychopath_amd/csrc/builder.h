// Host BVH4 builder: the reference's accel builder (binned SAH BVH2 with the
// union-vs-time-samples rule, median splits near the depth limit, collapse to
// BVH4 with time-sampled bounds), emitting GPU node records directly.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "layout.h"
#include "psymath.h"

namespace psy {

// Objects to build over.  Object k has time-sampled bounds
// bounds[offsets[k] .. offsets[k+1]).
struct BuildInput {
    uint32_t count;
    const Box* bounds;
    const uint32_t* offsets;
};

// A built BVH4 in host form, ready to be appended to a scene's node arrays.
struct HostBvh4 {
    // Node records; child references of internal nodes are indices into
    // `nodes`; leaf references carry positions in `order`.
    std::vector<GpuNode> nodes;
    // Time-sample records referenced by GpuNode::motion (index into this).
    std::vector<GpuBox4> motion;
    // Objects in leaf order (the permutation the builder applied).
    std::vector<uint32_t> order;
    // Root reference: index 0 of `nodes`, a leaf reference, or REF_EMPTY.
    uint32_t root = REF_EMPTY;
    uint32_t depth = 0;        // BVH4 depth (src/accel/bvh4.rs:87)
    uint64_t node_count = 0;   // as counted by BVH4::construct_from_base
    uint64_t leaf_count = 0;
    std::vector<Box> root_bounds;  // time-sampled bounds of the whole tree
};

// `per_leaf` objects per leaf at most.  `threads` bounds the worker threads
// used for independent subtrees (0 = hardware concurrency).
HostBvh4 build_bvh4(const BuildInput& in, uint32_t per_leaf, unsigned threads);

// objects_split::sah_split (src/accel/objects_split.rs:153-238) over the
// objects `keys[0..n)` of `in` (n >= 2): partitions `keys` in place and
// returns the split position.  The light tree splits with it too.
uint32_t sah_split_keys(const BuildInput& in, uint32_t* keys, uint32_t n);

// A BVH2 in array form (the intermediate tree of BVHBase, bvh_base.rs:23-35,
// as the device builder leaves it).  Node k: internal with children a[k],
// b[k] and split axis meta[k] & 3, or (meta[k] & LEAF) a leaf over leaf-order
// objects [a[k], b[k]); blen[k] time-sample boxes at bounds[k * stride].
struct ArrayBvh2 {
    static constexpr uint32_t LEAF = 0x80000000u;
    std::vector<uint32_t> a, b, meta, blen;
    std::vector<Box> bounds;
    uint32_t stride = 1;
    uint32_t root = 0;
};

// The BVH4 collapse of an array-form BVH2 whose leaves hold `order`'s
// objects, with the BVH2's deepest leaf at `base_depth`: the same records
// build_bvh4 emits for that tree.
HostBvh4 collapse_bvh2(const ArrayBvh2& tree, std::vector<uint32_t>&& order, uint32_t base_depth);

// build_bvh4 on the current HIP device (builder_device.hip): the same tree,
// bit for bit, built level by level on the GPU.  Returns false, with the
// reason in `why`, when the input needs the host builder: NaN or negative
// zero in an object's derived boxes (whose unions depend on order) or a HIP
// error.  `ms`, if set, receives the device build's wall time.
bool build_bvh4_device(const BuildInput& in, uint32_t per_leaf, HostBvh4& out, std::string& why, double* ms);

// The binary light tree's partition (LightTree::from_objects,
// src/accel/light_tree.rs:211-290: one light per leaf, SAH splits only) on
// the current HIP device: the topology as an ArrayBvh2 without bounds, and
// `order`, the lights in leaf order, as the host build leaves its keys.
// Returns false, with the reason, where build_bvh4_device would.
bool build_light_bvh2_device(const BuildInput& in, ArrayBvh2& topology, std::vector<uint32_t>& order,
                             std::string& why);

// Canonical pre-order dump (shared with the parity tests):
//   internal: 0, child_count, traversal_code, bounds_len, bounds_len*24 f32 bits, children...
//   leaf:     1, obj_start, obj_end
// preceded by (has_root, depth, node_count).
void dump_bvh4(const HostBvh4& bvh, std::vector<uint32_t>& out);

// TRAVERSAL_TABLE[octant][code] (sub_crates/bvh_order/build.rs).
const uint8_t* traversal_order_table();

}  // namespace psy
