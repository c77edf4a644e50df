// Host BVH4 builder.  Follows the reference builder step for step so the GPU
// traverses exactly the tree the reference would build:
//   BVHBase::recursive_build        src/accel/bvh_base.rs:90-215
//   acc_bounds / union rule         src/accel/bvh_base.rs:64-88, 116-132, 194-203
//   sah_split / median_split        src/accel/objects_split.rs:153-238, 291-341
//   partition / quick_select        src/algorithm.rs:44-81, 203-231
//   merge_slices_append             src/algorithm.rs:233-262
//   BVH4::construct_from_base       src/accel/bvh4.rs:186-358
// Differences from the reference are in organisation only: per-object
// mid-time boxes and centroids are computed once instead of at every level,
// independent subtrees are built on worker threads, and BVH4 nodes are
// emitted straight into 128-B GPU records.
#include "builder.h"

#include "build_common.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <future>
#include <cmath>
#include <memory>
#include <thread>

namespace psy {

namespace {

// Child visit order for one node code and ray octant, as the bvh_order table
// defines it (sub_crates/bvh_order/build.rs:5-59): the children of a split
// are visited low side first unless the ray runs toward negative coordinates
// along that split's axis.
uint8_t visit_order(int octant, int code) {
    auto neg = [&](int axis) { return ((octant >> axis) & 1) != 0; };
    auto pack = [](const int* p, int n) {
        int v = 0;
        for (int i = 0; i < n; i++) v |= p[i] << (2 * i);
        return (uint8_t)v;
    };
    if (code < 27) {
        int top = code % 3, left = (code / 3) % 3, right = code / 9;
        int p[4] = {0, 1, 2, 3};
        if (neg(left)) std::swap(p[0], p[1]);
        if (neg(right)) std::swap(p[2], p[3]);
        if (neg(top)) { std::swap(p[0], p[2]); std::swap(p[1], p[3]); }
        return pack(p, 4);
    }
    if (code < 36) {  // left pair split, right child a leaf
        int top = (code - 27) % 3, left = (code - 27) / 3;
        int p[3] = {0, 1, 2};
        if (neg(left)) std::swap(p[0], p[1]);
        if (neg(top)) { int q[3] = {p[2], p[0], p[1]}; std::memcpy(p, q, sizeof(p)); }
        return pack(p, 3);
    }
    if (code < 45) {  // right pair split, left child a leaf
        int top = (code - 36) % 3, right = (code - 36) / 3;
        int p[3] = {0, 1, 2};
        if (neg(right)) std::swap(p[1], p[2]);
        if (neg(top)) { int q[3] = {p[1], p[2], p[0]}; std::memcpy(p, q, sizeof(p)); }
        return pack(p, 3);
    }
    int top = code - 45;
    int p[2] = {0, 1};
    if (neg(top)) std::swap(p[0], p[1]);
    return pack(p, 2);
}

struct OrderTable {
    uint8_t t[8 * 48];
    OrderTable() {
        for (int o = 0; o < 8; o++)
            for (int c = 0; c < 48; c++) t[o * 48 + c] = visit_order(o, c);
    }
};

// BVH2 node of the intermediate tree (BVHBaseNode, bvh_base.rs:23-35).
struct BaseNode {
    bool leaf = false;
    std::vector<Box> bounds;  // time samples
    std::unique_ptr<BaseNode> kid[2];
    uint8_t axis = 0;
    uint32_t obj_start = 0, obj_end = 0;
};

// The union rule (bvh_base.rs:116-132, 194-203) over a sample list.
void union_or_samples(const std::vector<Box>& samples, std::vector<Box>& out) {
    out.resize(samples.size());
    out.resize(union_or_samples(samples.data(), (uint32_t)samples.size(), out.data()));
}

class Builder {
   public:
    Builder(const BuildInput& in, uint32_t per_leaf, unsigned threads)
        : in_(in), per_leaf_(per_leaf), threads_(threads) {
        info_.resize(in.count);
        for (uint32_t k = 0; k < in.count; k++) {
            const Box* s = in.bounds + in.offsets[k];
            uint32_t n = in.offsets[k + 1] - in.offsets[k];
            info_[k] = obj_info(s, n);
            if (!obj_order_free(info_[k])) order_free_ = false;
        }
    }

    // Builds over keys[0..n), keeping each object's derived data in a
    // parallel array permuted with the keys, so the split passes stream
    // through memory instead of gathering.
    std::unique_ptr<BaseNode> build_all(uint32_t* keys, uint32_t n, unsigned spawn_budget, uint32_t& max_depth) {
        std::vector<ObjInfo> recs(n);
        for (uint32_t i = 0; i < n; i++) recs[i] = info_[keys[i]];
        std::vector<ObjInfo>().swap(info_);
        return build(keys, recs.data(), n, 0, 0, spawn_budget, max_depth);
    }

    std::unique_ptr<BaseNode> build(uint32_t* keys, ObjInfo* recs, uint32_t n, uint32_t offset, uint32_t depth,
                                    unsigned spawn_budget, uint32_t& max_depth) {
        auto node = std::make_unique<BaseNode>();
        if (n <= per_leaf_) {
            leaf_bounds(keys, n, node->bounds);
            node->leaf = true;
            node->obj_start = offset;
            node->obj_end = offset + n;
            max_depth = std::max(max_depth, depth);
            return node;
        }
        uint32_t split;
        uint8_t axis;
        if (use_sah(n, depth))
            sah_split(keys, recs, n, split, axis);
        else
            median_split(keys, recs, n, split, axis);
        node->axis = axis;

        uint32_t d0 = 0, d1 = 0;
        if (spawn_budget > 1 && n > 20000) {
            unsigned half = spawn_budget / 2;
            auto fut = std::async(std::launch::async, [&] {
                return build(keys + split, recs + split, n - split, offset + split, depth + 1, spawn_budget - half, d1);
            });
            node->kid[0] = build(keys, recs, split, offset, depth + 1, half, d0);
            node->kid[1] = fut.get();
        } else {
            node->kid[0] = build(keys, recs, split, offset, depth + 1, 1, d0);
            node->kid[1] = build(keys + split, recs + split, n - split, offset + split, depth + 1, 1, d1);
        }
        max_depth = std::max(max_depth, std::max(d0, d1));

        std::vector<Box> merged;
        merge_union(node->kid[0]->bounds, node->kid[1]->bounds, merged);
        union_or_samples(merged, node->bounds);
        return node;
    }

    uint32_t split_sah(uint32_t* keys, uint32_t n) {
        std::vector<ObjInfo> recs(n);
        for (uint32_t i = 0; i < n; i++) recs[i] = info_[keys[i]];
        uint32_t split;
        uint8_t axis;
        sah_split(keys, recs.data(), n, split, axis);
        return split;
    }

   private:
    const BuildInput& in_;
    uint32_t per_leaf_;
    unsigned threads_;
    std::vector<ObjInfo> info_;
    // Without NaN or negative zero in the per-object boxes and centroids,
    // min/max unions give the same box in any order, so the SAH candidates
    // can be accumulated per bin and combined afterwards.
    bool order_free_ = true;

    // acc_bounds (bvh_base.rs:64-88) followed by the union rule.
    void leaf_bounds(const uint32_t* keys, uint32_t n, std::vector<Box>& out) {
        uint32_t max_len = 0;
        for (uint32_t i = 0; i < n; i++) max_len = std::max(max_len, in_.offsets[keys[i] + 1] - in_.offsets[keys[i]]);
        std::vector<Box> acc(max_len, box_empty());
        for (uint32_t i = 0; i < n; i++) {
            const Box* s = in_.bounds + in_.offsets[keys[i]];
            uint32_t len = in_.offsets[keys[i] + 1] - in_.offsets[keys[i]];
            if (len == max_len) {
                for (uint32_t j = 0; j < len; j++) acc[j] = box_union(acc[j], s[j]);
            } else {
                float sc = (float)(max_len - 1);
                for (uint32_t j = 0; j < max_len; j++) acc[j] = box_union(acc[j], lerp_slice_box(s, len, (float)j / sc));
            }
        }
        union_or_samples(acc, out);
    }

    // merge_slices_append with union (algorithm.rs:233-262).
    static void merge_union(const std::vector<Box>& a, const std::vector<Box>& b, std::vector<Box>& out) {
        out.clear();
        uint32_t na = (uint32_t)a.size(), nb = (uint32_t)b.size();
        if (na == 0 || nb == 0) return;
        if (na == nb) {
            for (uint32_t i = 0; i < na; i++) out.push_back(box_union(a[i], b[i]));
        } else if (na > nb) {
            float s = (float)(na - 1);
            for (uint32_t i = 0; i < na; i++) out.push_back(box_union(a[i], lerp_slice_box(b.data(), nb, (float)i / s)));
        } else {
            float s = (float)(nb - 1);
            for (uint32_t i = 0; i < nb; i++) out.push_back(box_union(lerp_slice_box(a.data(), na, (float)i / s), b[i]));
        }
    }

    // Per-bin boxes and counts of one axis: an object goes to the first bin
    // k whose divider it does not exceed (c <= divs[k]), or to the last.
    struct Bins {
        Box box[3][SAH_BINS];
        uint32_t count[3][SAH_BINS];
        Bins() {
            for (int d = 0; d < 3; d++)
                for (uint32_t k = 0; k < SAH_BINS; k++) {
                    box[d][k] = box_empty();
                    count[d][k] = 0;
                }
        }
    };
    static void bin_range(const ObjInfo* recs, uint32_t lo, uint32_t hi, const float (&divs)[3][SAH_BINS - 1],
                          Bins& b) {
        for (uint32_t i = lo; i < hi; i++) {
            const ObjInfo& o = recs[i];
            for (int d = 0; d < 3; d++) {
                const uint32_t k = sah_bin(comp(o.centroid, d), divs[d]);
                b.box[d][k] = box_union(b.box[d][k], o.mid);
                b.count[d][k]++;
            }
        }
    }

    // sah_split (objects_split.rs:153-238)
    void sah_split(uint32_t* keys, ObjInfo* recs, uint32_t n, uint32_t& split, uint8_t& axis) {
        Box cb = box_empty();
        for (uint32_t i = 0; i < n; i++) cb = box_union_pt(cb, recs[i].mid_center);

        float divs[3][SAH_BINS - 1];
        sah_divs(cb, divs);

        int best_d = 0;
        float best_div = 0.0f;
        if (order_free_) {
            // Bin once per object, split large nodes over threads, then
            // combine the bins into the candidates' left and right sides.
            Bins bins;
            const unsigned parts = n >= (1u << 18) ? std::max(1u, std::min(threads_, 16u)) : 1u;
            if (parts > 1) {
                std::vector<Bins> part(parts);
                std::vector<std::thread> pool;
                for (unsigned t = 0; t < parts; t++)
                    pool.emplace_back([&, t] {
                        bin_range(recs, (uint32_t)((uint64_t)n * t / parts), (uint32_t)((uint64_t)n * (t + 1) / parts),
                                  divs, part[t]);
                    });
                for (auto& th : pool) th.join();
                for (unsigned t = 0; t < parts; t++)
                    for (int d = 0; d < 3; d++)
                        for (uint32_t k = 0; k < SAH_BINS; k++) {
                            bins.box[d][k] = box_union(bins.box[d][k], part[t].box[d][k]);
                            bins.count[d][k] += part[t].count[d][k];
                        }
            } else {
                bin_range(recs, 0, n, divs, bins);
            }
            sah_choose(bins.box, bins.count, divs, best_d, best_div);
        } else {
            // The reference's accumulation, object by object into every
            // candidate's side, for boxes whose unions depend on order.
            Box left[3][SAH_BINS - 1], right[3][SAH_BINS - 1];
            uint32_t nl[3][SAH_BINS - 1] = {}, nr[3][SAH_BINS - 1] = {};
            for (int d = 0; d < 3; d++)
                for (uint32_t k = 0; k < SAH_BINS - 1; k++) left[d][k] = right[d][k] = box_empty();
            for (uint32_t i = 0; i < n; i++) {
                const ObjInfo& o = recs[i];
                for (int d = 0; d < 3; d++) {
                    float c = comp(o.centroid, d);
                    for (uint32_t k = 0; k < SAH_BINS - 1; k++) {
                        if (c <= divs[d][k]) {
                            left[d][k] = box_union(left[d][k], o.mid);
                            nl[d][k]++;
                        } else {
                            right[d][k] = box_union(right[d][k], o.mid);
                            nr[d][k]++;
                        }
                    }
                }
            }
            float best = __builtin_huge_valf();
            for (int d = 0; d < 3; d++)
                for (uint32_t k = 0; k < SAH_BINS - 1; k++) {
                    float lc = box_area(left[d][k]) * (float)nl[d][k];
                    float rc = box_area(right[d][k]) * (float)nr[d][k];
                    float cost = (lc * box_diag(left[d][k])) + (rc * box_diag(right[d][k]));
                    if (cost < best) {
                        best_d = d;
                        best_div = divs[d][k];
                        best = cost;
                    }
                }
        }

        uint32_t s = hoare_partition(keys, recs, n, [&](const ObjInfo& o) { return sah_left(o, best_d, best_div); });
        if (s < 1) s = 1;
        else if (s >= n) s = n - 1;
        split = s;
        axis = (uint8_t)best_d;
    }
};

// Node access of the collapse for the host builder's BVH2 (BaseNode tree).
struct BaseTreeView {
    using H = const BaseNode*;
    bool leaf(H n) const { return n->leaf; }
    H kid(H n, int i) const { return n->kid[i].get(); }
    uint32_t axis(H n) const { return n->axis; }
    uint32_t obj_start(H n) const { return n->obj_start; }
    uint32_t obj_end(H n) const { return n->obj_end; }
    uint32_t blen(H n) const { return (uint32_t)n->bounds.size(); }
    const Box* bounds(H n) const { return n->bounds.data(); }
};

// Node access of the collapse for a BVH2 in array form.
struct ArrayTreeView {
    using H = uint32_t;
    const ArrayBvh2& t;
    bool leaf(H n) const { return (t.meta[n] & ArrayBvh2::LEAF) != 0; }
    H kid(H n, int i) const { return i == 0 ? t.a[n] : t.b[n]; }
    uint32_t axis(H n) const { return t.meta[n] & 3u; }
    uint32_t obj_start(H n) const { return t.a[n]; }
    uint32_t obj_end(H n) const { return t.b[n]; }
    uint32_t blen(H n) const { return t.blen[n]; }
    const Box* bounds(H n) const { return t.bounds.data() + (uint64_t)n * t.stride; }
};

// BVH4 collapse (bvh4.rs:186-358) into GPU records.
template <typename View>
class Collapser {
   public:
    using H = typename View::H;
    Collapser(HostBvh4& out, const View& v) : out_(out), v_(v), table_(traversal_order_table()) {}

    // Returns the reference for `n`: a leaf reference or the node index.
    uint32_t emit(H n, uint32_t slot) {
        if (v_.leaf(n)) {
            out_.node_count += 1;
            out_.leaf_count += 1;
            return make_leaf_ref(v_.obj_start(n), v_.obj_end(n) - v_.obj_start(n));
        }
        const H l = v_.kid(n, 0);
        const H r = v_.kid(n, 1);
        H ch[4] = {};
        uint32_t cnt;
        uint32_t code;
        if (!v_.leaf(l) && !v_.leaf(r)) {
            ch[0] = v_.kid(l, 0); ch[1] = v_.kid(l, 1); ch[2] = v_.kid(r, 0); ch[3] = v_.kid(r, 1);
            cnt = 4;
            code = v_.axis(n) + v_.axis(l) * 3u + v_.axis(r) * 9u;
        } else if (!v_.leaf(l)) {
            ch[0] = v_.kid(l, 0); ch[1] = v_.kid(l, 1); ch[2] = r;
            cnt = 3;
            code = v_.axis(n) + v_.axis(l) * 3u + 27u;
        } else if (!v_.leaf(r)) {
            ch[0] = l; ch[1] = v_.kid(r, 0); ch[2] = v_.kid(r, 1);
            cnt = 3;
            code = v_.axis(n) + v_.axis(r) * 3u + 36u;
        } else {
            ch[0] = l; ch[1] = r;
            cnt = 2;
            code = v_.axis(n) + 45u;
        }
        out_.node_count += cnt;

        uint32_t blen = 0;
        for (uint32_t i = 0; i < cnt; i++) blen = std::max(blen, v_.blen(ch[i]));
        GpuBox4 samples[64];  // bounds keep at most 64 time samples
        for (uint32_t s = 0; s < blen; s++) {
            float time = blen < 2 ? 0.0f : (float)s / (float)(blen - 1);
            for (uint32_t i = 0; i < 4; i++) {
                Box b = box_empty();
                if (i < cnt) {
                    if (blen < 2) b = v_.bounds(ch[i])[0];
                    else b = lerp_slice_box(v_.bounds(ch[i]), v_.blen(ch[i]), time);
                }
                samples[s].lo_x[i] = b.lo.x; samples[s].hi_x[i] = b.hi.x;
                samples[s].lo_y[i] = b.lo.y; samples[s].hi_y[i] = b.hi.y;
                samples[s].lo_z[i] = b.lo.z; samples[s].hi_z[i] = b.hi.z;
            }
        }

        // Children get a contiguous block of slots, like the reference's
        // per-node child array; leaves occupy no slot.
        uint32_t internal = 0;
        for (uint32_t i = 0; i < cnt; i++) internal += v_.leaf(ch[i]) ? 0 : 1;
        uint32_t base = (uint32_t)out_.nodes.size();
        out_.nodes.resize(out_.nodes.size() + internal);
        uint32_t refs[4] = {REF_EMPTY, REF_EMPTY, REF_EMPTY, REF_EMPTY};
        uint32_t next = base;
        for (uint32_t i = 0; i < cnt; i++) {
            if (v_.leaf(ch[i])) refs[i] = emit(ch[i], 0);
            else { refs[i] = emit(ch[i], next); next++; }
        }

        GpuNode& g = out_.nodes[slot];
        std::memset(&g, 0, sizeof(g));
        std::memcpy(g.lo_x, samples[0].lo_x, sizeof(GpuBox4));
        for (int i = 0; i < 4; i++) g.child[i] = refs[i];
        for (int o = 0; o < 8; o++) g.order[o] = table_[o * 48 + code];
        g.meta = cnt | (blen << 8) | (code << 16);
        g.motion = REF_EMPTY;
        if (blen > 1) {
            g.motion = (uint32_t)out_.motion.size();
            out_.motion.insert(out_.motion.end(), samples, samples + blen);
        }
        return slot;
    }

   private:
    HostBvh4& out_;
    const View& v_;
    const uint8_t* table_;
};

template <typename View>
void collapse_into(HostBvh4& out, const View& v, typename View::H root) {
    Collapser<View> c(out, v);
    if (v.leaf(root)) {
        out.root = c.emit(root, 0);
    } else {
        out.nodes.resize(1);
        out.root = c.emit(root, 0);
    }
}

void dump_ref(const HostBvh4& bvh, uint32_t ref, std::vector<uint32_t>& out) {
    if (ref & REF_LEAF) {
        uint32_t first = (ref & ~REF_LEAF) >> 2, cnt = ref & 3u;
        out.push_back(1);
        out.push_back(first);
        out.push_back(first + cnt);
        return;
    }
    const GpuNode& g = bvh.nodes[ref];
    uint32_t cnt = g.meta & 7u, blen = (g.meta >> 8) & 0xffu, code = (g.meta >> 16) & 0x3fu;
    out.push_back(0);
    out.push_back(cnt);
    out.push_back(code);
    out.push_back(blen);
    for (uint32_t s = 0; s < blen; s++) {
        const GpuBox4* b = blen > 1 ? &bvh.motion[g.motion + s] : reinterpret_cast<const GpuBox4*>(g.lo_x);
        const float* f = b->lo_x;
        for (int i = 0; i < 24; i++) {
            uint32_t u;
            std::memcpy(&u, f + i, 4);
            out.push_back(u);
        }
    }
    for (uint32_t i = 0; i < cnt; i++) dump_ref(bvh, g.child[i], out);
}

}  // namespace

const uint8_t* traversal_order_table() {
    static OrderTable t;
    return t.t;
}

HostBvh4 build_bvh4(const BuildInput& in, uint32_t per_leaf, unsigned threads) {
    HostBvh4 out;
    if (in.count == 0) return out;
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    out.order.resize(in.count);
    for (uint32_t i = 0; i < in.count; i++) out.order[i] = i;

    Builder b(in, per_leaf, threads);
    uint32_t base_depth = 0;
    std::unique_ptr<BaseNode> root = b.build_all(out.order.data(), in.count, threads, base_depth);
    out.depth = base_depth / 2 + 1;
    out.root_bounds = root->bounds;

    collapse_into(out, BaseTreeView{}, root.get());
    return out;
}

HostBvh4 collapse_bvh2(const ArrayBvh2& t, std::vector<uint32_t>&& order, uint32_t base_depth) {
    HostBvh4 out;
    out.order = std::move(order);
    if (out.order.empty()) return out;
    out.depth = base_depth / 2 + 1;
    const Box* rb = t.bounds.data() + (uint64_t)t.root * t.stride;
    out.root_bounds.assign(rb, rb + t.blen[t.root]);
    collapse_into(out, ArrayTreeView{t}, t.root);
    return out;
}

uint32_t sah_split_keys(const BuildInput& in, uint32_t* keys, uint32_t n) {
    Builder b(in, 1, 1);
    return b.split_sah(keys, n);
}

void dump_bvh4(const HostBvh4& bvh, std::vector<uint32_t>& out) {
    out.push_back(bvh.root != REF_EMPTY ? 1u : 0u);
    out.push_back(bvh.depth);
    out.push_back((uint32_t)bvh.node_count);
    if (bvh.root != REF_EMPTY) dump_ref(bvh, bvh.root, out);
}

}  // namespace psy
