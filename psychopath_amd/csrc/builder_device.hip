// Device BVH4 builder for gfx950: the host builder's tree (builder.cpp), bit
// for bit, built level by level on the GPU.
//
// The reference builds its BVH2 recursively (BVHBase::recursive_build,
// src/accel/bvh_base.rs:96-215): binned SAH over the node's objects
// (objects_split.rs:153-238), a Hoare partition (algorithm.rs:44-81), median
// splits near the depth limit, then bounds bottom-up with the union rule.
// Here every node of one depth is split at once:
//   * large nodes (more than SMALL objects) are cut into chunks of CHUNK
//     objects, one workgroup each: centroid bounds and SAH bins are reduced
//     per chunk and combined per node (unions of boxes without NaN or -0 and
//     integer counts do not depend on order, so the bins equal the
//     reference's sequential ones), the split is chosen by one lane per node
//     with the shared sah_choose, and the partition is done in parallel:
//     the Hoare partition swaps the k-th object that belongs right among the
//     first `trues` positions with the k-th object that belongs left among
//     the rest counted from the end, so ranks from a prefix count of the
//     predicate give every swap pair at once;
//   * nodes near the depth limit run the reference's median split with
//     quick_select, one lane per node (sequential, as the reference);
//   * nodes of at most SMALL objects are built to the end by one lane each,
//     with the same functions the host builder calls.
// Bounds then go bottom-up, one lane per node, depth by depth, and the BVH2
// collapses into BVH4 records on the device with the host collapse's
// numbering.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "build_common.h"
#include "builder.h"

namespace psy {

namespace {

constexpr uint32_t CHUNK = 2048;   // objects per workgroup in the chunked passes
constexpr uint32_t SMALL = 64;     // nodes of at most this many objects: one lane builds the subtree
constexpr int BLOCK = 256;
constexpr uint32_t NBINS = 3 * SAH_BINS;
constexpr uint32_t MAX_SAMPLES = 64;  // time samples of an object's bounds the device path takes
constexpr uint32_t UNUSED = 0x0000ff00u;  // meta of a node id no node took (depth field 255)
constexpr uint32_t LEAF = ArrayBvh2::LEAF;

struct Seg {
    uint32_t start, count, node, depth;
};
struct Chunk {
    uint32_t seg, begin, end;
};
// Per-node results of one level's split.
struct SegWork {
    Box cb;
    float divs[3][SAH_BINS - 1];
    int best_d;
    float best_div;
    uint32_t median;  // 1: median split (split/axis already set)
    uint32_t trues, m, split, axis;
};
struct ChunkBins {
    Box box[NBINS];
    uint32_t count[NBINS];
};

// Order-preserving map of floats without NaN onto ints, for LDS min/max.
__device__ __forceinline__ int f2o(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float o2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__device__ __forceinline__ void write_node(uint32_t* na, uint32_t* nb, uint32_t* meta, uint32_t id, uint32_t a,
                                           uint32_t b, uint32_t m) {
    na[id] = a;
    nb[id] = b;
    meta[id] = m;
}

// Per object: its key and derived boxes; flags boxes whose unions depend
// on order.
__global__ void k_init(const Box* bounds, const uint32_t* offs, uint32_t n, uint32_t* keys, ObjInfo* recs,
                       uint32_t* flag) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    keys[i] = i;
    const ObjInfo o = obj_info(bounds + offs[i], offs[i + 1] - offs[i]);
    recs[i] = o;
    if (!obj_order_free(o)) atomicOr(flag, 1u);
}

__global__ void k_fill(uint32_t* p, uint32_t v, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}

// Exclusive scan of in[0, n) into out[0, n), total in out[n]; one workgroup
// of 1024 lanes.
__global__ __launch_bounds__(1024) void k_scan(const uint32_t* in, uint32_t* out, uint32_t n) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    if (t == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += 1024) {
        const uint32_t i = base + t;
        const uint32_t v = i < n ? in[i] : 0u;
        // inclusive wave scan
        uint32_t x = v;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if ((int)lane >= off) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        if (t == 0) {
            uint32_t acc = 0;
            for (int k = 0; k < 16; k++) {
                const uint32_t s = wsum[k];
                wsum[k] = acc;
                acc += s;
            }
        }
        __syncthreads();
        const uint32_t c = carry;
        if (i < n) out[i] = c + wsum[w] + x - v;
        __syncthreads();
        if (t == 1023) carry = c + wsum[w] + x;
        __syncthreads();
    }
    if (t == 0) out[n] = carry;
}

// Chunks of each large node: count, then fill after the scan.
__global__ void k_seg_nchunks(const Seg* segs, const uint32_t* nseg, uint32_t* nch) {
    const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= *nseg) return;
    nch[s] = (segs[s].count + CHUNK - 1) / CHUNK;
}
__global__ void k_chunk_fill(const Seg* segs, const uint32_t* nseg, const uint32_t* choff, Chunk* chunks) {
    const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= *nseg) return;
    const Seg g = segs[s];
    const uint32_t end = g.start + g.count;
    uint32_t c = choff[s];
    for (uint32_t b = g.start; b < end; b += CHUNK, c++) chunks[c] = Chunk{s, b, min(end, b + CHUNK)};
}

// Centroid (mid_center) bounds of each chunk.
__global__ __launch_bounds__(BLOCK) void k_chunk_cb(const Chunk* chunks, const uint32_t* nchunks, const ObjInfo* recs,
                                                    Box* ch_cb) {
    __shared__ float red[6][BLOCK];
    const uint32_t c = blockIdx.x;
    if (c >= *nchunks) return;
    const Chunk ch = chunks[c];
    Box b = box_empty();
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += BLOCK) b = box_union_pt(b, recs[i].mid_center);
    red[0][threadIdx.x] = b.lo.x; red[1][threadIdx.x] = b.lo.y; red[2][threadIdx.x] = b.lo.z;
    red[3][threadIdx.x] = b.hi.x; red[4][threadIdx.x] = b.hi.y; red[5][threadIdx.x] = b.hi.z;
    __syncthreads();
    for (int off = BLOCK / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            for (int k = 0; k < 3; k++) red[k][threadIdx.x] = fmin_sse(red[k][threadIdx.x], red[k][threadIdx.x + off]);
            for (int k = 3; k < 6; k++) red[k][threadIdx.x] = fmax_sse(red[k][threadIdx.x], red[k][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        ch_cb[c] = Box{V3{red[0][0], red[1][0], red[2][0]}, V3{red[3][0], red[4][0], red[5][0]}};
}

// Per node: centroid bounds from its chunks, SAH dividers, or the median
// split (one lane, in place) near the depth limit.
__global__ void k_seg_prep(const Seg* segs, const uint32_t* nseg, const uint32_t* choff, const Box* ch_cb,
                           SegWork* work, uint32_t* keys, ObjInfo* recs, uint32_t sah_only) {
    const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= *nseg) return;
    const Seg g = segs[s];
    SegWork& w = work[s];
    w.m = 0;
    if (!sah_only && !use_sah(g.count, g.depth)) {
        uint32_t split;
        uint8_t axis;
        median_split(keys + g.start, recs + g.start, g.count, split, axis);
        w.median = 1;
        w.split = split;
        w.axis = axis;
        return;
    }
    w.median = 0;
    Box cb = box_empty();
    for (uint32_t c = choff[s]; c < choff[s + 1]; c++) cb = box_union(cb, ch_cb[c]);
    w.cb = cb;
    sah_divs(cb, w.divs);
}

// SAH bins of each chunk (LDS min/max and counts).
__global__ __launch_bounds__(BLOCK) void k_chunk_bins(const Chunk* chunks, const uint32_t* nchunks,
                                                      const SegWork* work, const ObjInfo* recs, ChunkBins* ch_bins) {
    __shared__ int lo[NBINS][3], hi[NBINS][3];
    __shared__ uint32_t cnt[NBINS];
    __shared__ float divs[3][SAH_BINS - 1];
    const uint32_t c = blockIdx.x;
    if (c >= *nchunks) return;
    const Chunk ch = chunks[c];
    const SegWork& w = work[ch.seg];
    if (w.median) return;
    for (uint32_t k = threadIdx.x; k < NBINS; k += BLOCK) {
        for (int j = 0; j < 3; j++) {
            lo[k][j] = f2o(__builtin_huge_valf());
            hi[k][j] = f2o(-__builtin_huge_valf());
        }
        cnt[k] = 0;
    }
    if (threadIdx.x < 3 * (SAH_BINS - 1)) divs[threadIdx.x / (SAH_BINS - 1)][threadIdx.x % (SAH_BINS - 1)] =
        w.divs[threadIdx.x / (SAH_BINS - 1)][threadIdx.x % (SAH_BINS - 1)];
    __syncthreads();
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += BLOCK) {
        const ObjInfo o = recs[i];
        const int l[3] = {f2o(o.mid.lo.x), f2o(o.mid.lo.y), f2o(o.mid.lo.z)};
        const int h[3] = {f2o(o.mid.hi.x), f2o(o.mid.hi.y), f2o(o.mid.hi.z)};
        for (int d = 0; d < 3; d++) {
            const uint32_t k = d * SAH_BINS + sah_bin(comp(o.centroid, d), divs[d]);
            for (int j = 0; j < 3; j++) {
                atomicMin(&lo[k][j], l[j]);
                atomicMax(&hi[k][j], h[j]);
            }
            atomicAdd(&cnt[k], 1u);
        }
    }
    __syncthreads();
    ChunkBins& out = ch_bins[c];
    for (uint32_t k = threadIdx.x; k < NBINS; k += BLOCK) {
        out.box[k] = Box{V3{o2f(lo[k][0]), o2f(lo[k][1]), o2f(lo[k][2])}, V3{o2f(hi[k][0]), o2f(hi[k][1]), o2f(hi[k][2])}};
        out.count[k] = cnt[k];
    }
}

// Per node (one wave): the bins over its chunks, then the split choice.
__global__ __launch_bounds__(64) void k_seg_split(const uint32_t* nseg, const uint32_t* choff, const ChunkBins* ch_bins,
                                                  SegWork* work) {
    __shared__ Box box[3][SAH_BINS];
    __shared__ uint32_t count[3][SAH_BINS];
    const uint32_t s = blockIdx.x;
    if (s >= *nseg) return;
    SegWork& w = work[s];
    if (w.median) return;
    const uint32_t k = threadIdx.x;
    if (k < NBINS) {
        Box b = box_empty();
        uint32_t n = 0;
        for (uint32_t c = choff[s]; c < choff[s + 1]; c++) {
            b = box_union(b, ch_bins[c].box[k]);
            n += ch_bins[c].count[k];
        }
        box[k / SAH_BINS][k % SAH_BINS] = b;
        count[k / SAH_BINS][k % SAH_BINS] = n;
    }
    __syncthreads();
    if (k == 0) {
        float divs[3][SAH_BINS - 1];
        for (int d = 0; d < 3; d++)
            for (uint32_t j = 0; j < SAH_BINS - 1; j++) divs[d][j] = w.divs[d][j];
        int best_d;
        float best_div;
        sah_choose(box, count, divs, best_d, best_div);
        w.best_d = best_d;
        w.best_div = best_div;
        w.axis = (uint32_t)best_d;
    }
}

// Objects of each chunk that go left.
__global__ __launch_bounds__(BLOCK) void k_chunk_count(const Chunk* chunks, const uint32_t* nchunks,
                                                       const SegWork* work, const ObjInfo* recs, uint32_t* ch_true) {
    __shared__ uint32_t tot;
    const uint32_t c = blockIdx.x;
    if (c >= *nchunks) return;
    const Chunk ch = chunks[c];
    const SegWork& w = work[ch.seg];
    if (w.median) return;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    uint32_t mine = 0;
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += BLOCK) mine += sah_left(recs[i], w.best_d, w.best_div);
    atomicAdd(&tot, mine);
    __syncthreads();
    if (threadIdx.x == 0) ch_true[c] = tot;
}

// Per node: offsets of its chunks' left counts, and the total.
__global__ void k_seg_trues(const uint32_t* nseg, const uint32_t* choff, const uint32_t* ch_true, uint32_t* ch_toff,
                            SegWork* work) {
    const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= *nseg) return;
    SegWork& w = work[s];
    if (w.median) return;
    uint32_t acc = 0;
    for (uint32_t c = choff[s]; c < choff[s + 1]; c++) {
        ch_toff[c] = acc;
        acc += ch_true[c];
    }
    w.trues = acc;
}

// The swap pairs of the partition: with T(r) the objects going left before
// position r and `trues` going left in all, an object going right at r <
// trues is the (r - T(r))-th such from the start, and one going left at
// r >= trues the (trues - T(r) - 1)-th such from the end.
__global__ __launch_bounds__(BLOCK) void k_chunk_rank(const Chunk* chunks, const uint32_t* nchunks, const Seg* segs,
                                                      SegWork* work, const ObjInfo* recs, const uint32_t* ch_toff,
                                                      uint32_t* posL, uint32_t* posR) {
    __shared__ uint32_t wcount[BLOCK / 64];
    __shared__ uint32_t misplaced;
    const uint32_t c = blockIdx.x;
    if (c >= *nchunks) return;
    const Chunk ch = chunks[c];
    SegWork& w = work[ch.seg];
    if (w.median) return;
    const uint32_t start = segs[ch.seg].start, trues = w.trues;
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    if (threadIdx.x == 0) misplaced = 0;
    uint32_t carry = ch_toff[c];
    uint32_t mine = 0;
    for (uint32_t base = ch.begin; base < ch.end; base += BLOCK) {
        const uint32_t i = base + threadIdx.x;
        const bool in = i < ch.end;
        const bool p = in && sah_left(recs[i], w.best_d, w.best_div);
        const uint64_t bal = __ballot(p);
        const uint32_t below = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        __syncthreads();
        if (lane == 0) wcount[wv] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t before = carry + below;
        for (uint32_t k = 0; k < wv; k++) before += wcount[k];
        if (in) {
            const uint32_t r = i - start;
            if (r < trues && !p) {
                posL[start + (r - before)] = i;
                mine++;
            } else if (r >= trues && p) {
                posR[start + (trues - before - 1)] = i;
            }
        }
        for (uint32_t k = 0; k < BLOCK / 64; k++) carry += wcount[k];
    }
    atomicAdd(&misplaced, mine);
    __syncthreads();
    if (threadIdx.x == 0 && misplaced) atomicAdd(&w.m, misplaced);
}

// The swaps: pair k exchanges the objects at posL[k] and posR[k].
__global__ __launch_bounds__(BLOCK) void k_chunk_swap(const Chunk* chunks, const uint32_t* nchunks, const Seg* segs,
                                                      const SegWork* work, const uint32_t* posL, const uint32_t* posR,
                                                      uint32_t* keys, ObjInfo* recs) {
    const uint32_t c = blockIdx.x;
    if (c >= *nchunks) return;
    const Chunk ch = chunks[c];
    const SegWork& w = work[ch.seg];
    if (w.median) return;
    const uint32_t start = segs[ch.seg].start;
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += BLOCK) {
        const uint32_t k = i - start;
        if (k >= w.m) break;
        const uint32_t a = posL[start + k], b = posR[start + k];
        const uint32_t ka = keys[a];
        keys[a] = keys[b];
        keys[b] = ka;
        const ObjInfo ra = recs[a];
        recs[a] = recs[b];
        recs[b] = ra;
    }
}

// Children of each node of the level: node ids node_base + 2s and + 2s + 1;
// leaves are written, larger children become candidates for the next level
// or for the one-lane subtree builds.
__global__ void k_children(const Seg* segs, const uint32_t* nseg, const SegWork* work, uint32_t node_base,
                           uint32_t per_leaf, uint32_t* na, uint32_t* nb, uint32_t* meta, Seg* cand, uint32_t* big,
                           uint32_t* small) {
    const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= *nseg) return;
    const Seg g = segs[s];
    const SegWork& w = work[s];
    uint32_t split = w.split;
    if (!w.median) {
        split = w.trues;
        if (split < 1) split = 1;
        else if (split >= g.count) split = g.count - 1;
    }
    const uint32_t c0 = node_base + 2 * s;
    write_node(na, nb, meta, g.node, c0, c0 + 1, (w.axis & 3u) | (g.depth << 8));
    for (uint32_t j = 0; j < 2; j++) {
        const uint32_t st = j == 0 ? g.start : g.start + split;
        const uint32_t cnt = j == 0 ? split : g.count - split;
        const Seg child{st, cnt, c0 + j, g.depth + 1};
        cand[2 * s + j] = child;
        big[2 * s + j] = cnt > SMALL ? 1u : 0u;
        small[2 * s + j] = cnt > per_leaf && cnt <= SMALL ? 1u : 0u;
        if (cnt <= per_leaf) write_node(na, nb, meta, c0 + j, st, st + cnt, LEAF | ((g.depth + 1) << 8));
    }
}

__global__ void k_compact(const Seg* cand, const uint32_t* flags, const uint32_t* off, uint32_t m, Seg* out,
                          uint32_t out_base) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < m && flags[i]) out[out_base + off[i]] = cand[i];
}

// One lane builds a small node's whole subtree (node ids from `id_off`).
__global__ void k_small(const Seg* small, uint32_t n_small, const uint32_t* id_off, uint32_t id_base,
                        uint32_t per_leaf, uint32_t sah_only, uint32_t* keys, ObjInfo* recs, uint32_t* na,
                        uint32_t* nb, uint32_t* meta) {
    const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= n_small) return;
    Seg stack[SMALL + 1];
    int sp = 0;
    stack[sp++] = small[s];
    uint32_t next = id_base + id_off[s];
    while (sp > 0) {
        const Seg g = stack[--sp];
        if (g.count <= per_leaf) {
            write_node(na, nb, meta, g.node, g.start, g.start + g.count, LEAF | (g.depth << 8));
            continue;
        }
        uint32_t* k = keys + g.start;
        ObjInfo* r = recs + g.start;
        uint32_t split;
        uint8_t axis;
        if (sah_only || use_sah(g.count, g.depth)) {
            Box cb = box_empty();
            for (uint32_t i = 0; i < g.count; i++) cb = box_union_pt(cb, r[i].mid_center);
            float divs[3][SAH_BINS - 1];
            sah_divs(cb, divs);
            Box box[3][SAH_BINS];
            uint32_t count[3][SAH_BINS];
            for (int d = 0; d < 3; d++)
                for (uint32_t j = 0; j < SAH_BINS; j++) {
                    box[d][j] = box_empty();
                    count[d][j] = 0;
                }
            for (uint32_t i = 0; i < g.count; i++) {
                const ObjInfo& o = r[i];
                for (int d = 0; d < 3; d++) {
                    const uint32_t j = sah_bin(comp(o.centroid, d), divs[d]);
                    box[d][j] = box_union(box[d][j], o.mid);
                    count[d][j]++;
                }
            }
            int best_d;
            float best_div;
            sah_choose(box, count, divs, best_d, best_div);
            uint32_t sp2 = hoare_partition(k, r, g.count, [=](const ObjInfo& o) { return sah_left(o, best_d, best_div); });
            if (sp2 < 1) sp2 = 1;
            else if (sp2 >= g.count) sp2 = g.count - 1;
            split = sp2;
            axis = (uint8_t)best_d;
        } else {
            median_split(k, r, g.count, split, axis);
        }
        const uint32_t c0 = next;
        next += 2;
        write_node(na, nb, meta, g.node, c0, c0 + 1, (uint32_t)axis | (g.depth << 8));
        stack[sp++] = Seg{g.start + split, g.count - split, c0 + 1, g.depth + 1};
        stack[sp++] = Seg{g.start, split, c0, g.depth + 1};
    }
}

__global__ void k_small_need(const Seg* small, uint32_t n_small, uint32_t* need) {
    const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
    if (s < n_small) need[s] = 2 * small[s].count - 2;
}

// Nodes by depth: histogram, then scatter.
__global__ void k_depth_hist(const uint32_t* meta, uint32_t n_nodes, uint32_t* hist) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n_nodes || meta[i] == UNUSED) return;
    atomicAdd(&hist[(meta[i] >> 8) & 0xffu], 1u);
}
__global__ void k_depth_scatter(const uint32_t* meta, uint32_t n_nodes, uint32_t* cursor, uint32_t* sorted) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n_nodes || meta[i] == UNUSED) return;
    sorted[atomicAdd(&cursor[(meta[i] >> 8) & 0xffu], 1u)] = i;
}

// Time-sampled bounds of the nodes of one depth (acc_bounds for leaves,
// bvh_base.rs:64-88; merge_slices_append of the children for internal nodes,
// bvh_base.rs:184-203; each followed by the union rule).
__global__ void k_bounds(const uint32_t* ids, uint32_t count, const uint32_t* na, const uint32_t* nb,
                         const uint32_t* meta, const uint32_t* keys, const Box* bounds, const uint32_t* offs,
                         uint32_t stride, uint32_t* blen, Box* nbounds) {
    const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= count) return;
    const uint32_t id = ids[t];
    Box acc[MAX_SAMPLES];
    uint32_t len;
    if (meta[id] & LEAF) {
        uint32_t max_len = 0;
        for (uint32_t i = na[id]; i < nb[id]; i++) max_len = max(max_len, offs[keys[i] + 1] - offs[keys[i]]);
        for (uint32_t j = 0; j < max_len; j++) acc[j] = box_empty();
        for (uint32_t i = na[id]; i < nb[id]; i++) {
            const Box* s = bounds + offs[keys[i]];
            const uint32_t l = offs[keys[i] + 1] - offs[keys[i]];
            if (l == max_len) {
                for (uint32_t j = 0; j < l; j++) acc[j] = box_union(acc[j], s[j]);
            } else {
                const float sc = (float)(max_len - 1);
                for (uint32_t j = 0; j < max_len; j++) acc[j] = box_union(acc[j], lerp_slice_box(s, l, (float)j / sc));
            }
        }
        len = max_len;
    } else {
        const uint32_t c0 = na[id], c1 = nb[id];
        const uint32_t la = blen[c0], lb = blen[c1];
        const Box* A = nbounds + (uint64_t)c0 * stride;
        const Box* B = nbounds + (uint64_t)c1 * stride;
        if (la == lb) {
            for (uint32_t i = 0; i < la; i++) acc[i] = box_union(A[i], B[i]);
            len = la;
        } else if (la > lb) {
            const float s = (float)(la - 1);
            for (uint32_t i = 0; i < la; i++) acc[i] = box_union(A[i], lerp_slice_box(B, lb, (float)i / s));
            len = la;
        } else {
            const float s = (float)(lb - 1);
            for (uint32_t i = 0; i < lb; i++) acc[i] = box_union(lerp_slice_box(A, la, (float)i / s), B[i]);
            len = lb;
        }
    }
    blen[id] = union_or_samples(acc, len, nbounds + (uint64_t)id * stride);
}

// ---- BVH4 collapse on the device (bvh4.rs:186-358), the host collapse's
// records and numbering (builder.cpp): every internal BVH2 node at an even
// depth becomes a BVH4 node whose children are its grandchildren through
// internal children and its leaf children; a node's internal children take a
// contiguous block of slots when it is emitted, in child order, and each
// child's subtree then takes the slots after it, depth first; motion records
// follow the same order after each node's subtree (post-order).

struct C4 {
    uint32_t ch[4];
    uint32_t cnt, code, blen, internal;  // children, traversal code, samples, internal children
};

__device__ __forceinline__ bool is_leaf(const uint32_t* meta, uint32_t id) { return (meta[id] & LEAF) != 0u; }

// Children, code and sample count of the BVH4 node at BVH2 node `id`.
__global__ void k_c4_prep(const uint32_t* ids, uint32_t count, const uint32_t* na, const uint32_t* nb,
                          const uint32_t* meta, const uint32_t* blen, C4* c4) {
    const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= count) return;
    const uint32_t x = ids[t];
    if (is_leaf(meta, x)) return;
    const uint32_t l = na[x], r = nb[x];
    const uint32_t ax = meta[x] & 3u;
    C4 c{};
    const bool li = !is_leaf(meta, l), ri = !is_leaf(meta, r);
    if (li && ri) {
        c.ch[0] = na[l]; c.ch[1] = nb[l]; c.ch[2] = na[r]; c.ch[3] = nb[r];
        c.cnt = 4;
        c.code = ax + (meta[l] & 3u) * 3u + (meta[r] & 3u) * 9u;
    } else if (li) {
        c.ch[0] = na[l]; c.ch[1] = nb[l]; c.ch[2] = r;
        c.cnt = 3;
        c.code = ax + (meta[l] & 3u) * 3u + 27u;
    } else if (ri) {
        c.ch[0] = l; c.ch[1] = na[r]; c.ch[2] = nb[r];
        c.cnt = 3;
        c.code = ax + (meta[r] & 3u) * 3u + 36u;
    } else {
        c.ch[0] = l; c.ch[1] = r;
        c.cnt = 2;
        c.code = ax + 45u;
    }
    c.blen = 0;
    c.internal = 0;
    for (uint32_t i = 0; i < c.cnt; i++) {
        c.blen = max(c.blen, blen[c.ch[i]]);
        c.internal += is_leaf(meta, c.ch[i]) ? 0u : 1u;
    }
    c4[x] = c;
}

// Bottom-up: BVH4 nodes and motion records below each node.
__global__ void k_c4_up(const uint32_t* ids, uint32_t count, const uint32_t* meta, const C4* c4, uint32_t* desc,
                        uint32_t* mdesc) {
    const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= count) return;
    const uint32_t x = ids[t];
    if (is_leaf(meta, x)) return;
    const C4 c = c4[x];
    uint32_t dn = 0, dm = 0;
    for (uint32_t i = 0; i < c.cnt; i++) {
        const uint32_t y = c.ch[i];
        if (is_leaf(meta, y)) continue;
        const uint32_t yb = c4[y].blen;
        dn += 1 + desc[y];
        dm += (yb > 1 ? yb : 0u) + mdesc[y];
    }
    desc[x] = dn;
    mdesc[x] = dm;
}

// Top-down: the slots and motion positions of each node's internal children.
__global__ void k_c4_down(const uint32_t* ids, uint32_t count, const uint32_t* meta, const C4* c4,
                          const uint32_t* desc, const uint32_t* mdesc, uint32_t* slot, uint32_t* S, uint32_t* M) {
    const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= count) return;
    const uint32_t x = ids[t];
    if (is_leaf(meta, x)) return;
    const C4 c = c4[x];
    uint32_t j = 0, next = S[x] + c.internal, m = M[x];
    for (uint32_t i = 0; i < c.cnt; i++) {
        const uint32_t y = c.ch[i];
        if (is_leaf(meta, y)) continue;
        slot[y] = S[x] + j++;
        S[y] = next;
        M[y] = m;
        next += desc[y];
        const uint32_t yb = c4[y].blen;
        m += mdesc[y] + (yb > 1 ? yb : 0u);
    }
}

// The records: four child boxes per time sample, child references, the
// visit orders of the node's code and its metadata.
__global__ void k_c4_emit(const uint32_t* ids, uint32_t count, const uint32_t* na, const uint32_t* nb,
                          const uint32_t* meta, const uint32_t* blen, const Box* nbounds, uint32_t stride,
                          const C4* c4, const uint32_t* slot, const uint32_t* mdesc, const uint32_t* M,
                          const uint8_t* table, GpuNode* nodes, GpuBox4* motion) {
    const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= count) return;
    const uint32_t x = ids[t];
    if (is_leaf(meta, x)) return;
    const C4 c = c4[x];
    const uint32_t mstart = c.blen > 1 ? M[x] + mdesc[x] : REF_EMPTY;
    for (uint32_t sm = 0; sm < c.blen; sm++) {
        const float time = c.blen < 2 ? 0.0f : (float)sm / (float)(c.blen - 1);
        GpuBox4 b4;
        for (uint32_t i = 0; i < 4; i++) {
            Box b = box_empty();
            if (i < c.cnt) {
                const Box* cb = nbounds + (uint64_t)c.ch[i] * stride;
                b = c.blen < 2 ? cb[0] : lerp_slice_box(cb, blen[c.ch[i]], time);
            }
            b4.lo_x[i] = b.lo.x; b4.hi_x[i] = b.hi.x;
            b4.lo_y[i] = b.lo.y; b4.hi_y[i] = b.hi.y;
            b4.lo_z[i] = b.lo.z; b4.hi_z[i] = b.hi.z;
        }
        if (sm == 0) {
            GpuNode& g = nodes[slot[x]];
            for (int i = 0; i < 4; i++) {
                g.lo_x[i] = b4.lo_x[i]; g.hi_x[i] = b4.hi_x[i];
                g.lo_y[i] = b4.lo_y[i]; g.hi_y[i] = b4.hi_y[i];
                g.lo_z[i] = b4.lo_z[i]; g.hi_z[i] = b4.hi_z[i];
            }
        }
        if (c.blen > 1) motion[mstart + sm] = b4;
    }
    GpuNode& g = nodes[slot[x]];
    for (uint32_t i = 0; i < 4; i++) {
        uint32_t ref = REF_EMPTY;
        if (i < c.cnt) {
            const uint32_t y = c.ch[i];
            ref = is_leaf(meta, y) ? (LEAF | (na[y] << 2) | (nb[y] - na[y])) : slot[y];
        }
        g.child[i] = ref;
    }
    for (int o = 0; o < 8; o++) g.order[o] = table[o * 48 + c.code];
    g.meta = c.cnt | (c.blen << 8) | (c.code << 16);
    g.motion = mstart;
}

// The deepest leaf (the BVH2 depth of BVHBase::depth).
__global__ void k_leaf_depth(const uint32_t* meta, uint32_t n_nodes, uint32_t* out) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n_nodes || meta[i] == UNUSED || !(meta[i] & LEAF)) return;
    atomicMax(out, (meta[i] >> 8) & 0xffu);
}

// BVH4::construct_from_base's tally: every child of a BVH4 node counts as
// a node, every leaf child as a leaf.
__global__ void k_c4_count(const uint32_t* meta, uint32_t n_nodes, const C4* c4, uint32_t* out) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n_nodes || meta[i] == UNUSED) return;
    const uint32_t depth = (meta[i] >> 8) & 0xffu;
    if (meta[i] & LEAF) {
        atomicAdd(&out[1], 1u);
        return;
    }
    if (depth & 1u) return;
    atomicAdd(&out[0], c4[i].cnt);
}

// Device allocations of one build, freed together.
struct Arena {
    std::vector<void*> ptrs;
    hipError_t err = hipSuccess;
    template <typename T>
    T* get(uint64_t count) {
        void* p = nullptr;
        if (err == hipSuccess) err = hipMalloc(&p, std::max<uint64_t>(count, 1) * sizeof(T));
        if (err == hipSuccess) ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    ~Arena() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

inline unsigned blocks_for(uint64_t n) { return (unsigned)std::max<uint64_t>(1, (n + BLOCK - 1) / BLOCK); }

#define DB_HIP(call)                                                         \
    do {                                                                     \
        const hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) {                                              \
            why = std::string(#call) + ": " + hipGetErrorString(e_);         \
            return false;                                                    \
        }                                                                    \
    } while (0)


// The device buffers of one build and its BVH2 (topology in na/nb/meta,
// object order in keys), freed together.
struct Bvh2Dev {
    Arena ar;
    hipStream_t st = nullptr;
    uint32_t n = 0, stride = 1, n_nodes = 0;
    Box* d_bounds = nullptr;
    uint32_t* d_offs = nullptr;
    uint32_t *keys = nullptr, *na = nullptr, *nb = nullptr, *meta = nullptr;
    ~Bvh2Dev() {
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    }
};

// The BVH2 phase (BVHBase::recursive_build, bvh_base.rs:96-215; with
// `sah_only` the light tree's, light_tree.rs:296-372, which never takes a
// median split), for in.count >= 1 objects.
bool build_bvh2_dev(const BuildInput& in, uint32_t per_leaf, bool sah_only, Bvh2Dev& b, std::string& why) {
    const uint32_t n = in.count;
    if (per_leaf < 1 || per_leaf > SMALL) {
        why = "objects per leaf outside the device builder's range";
        return false;
    }
    uint32_t stride = 1;
    for (uint32_t k = 0; k < n; k++) stride = std::max(stride, in.offsets[k + 1] - in.offsets[k]);
    if (stride > MAX_SAMPLES) {
        why = "more than 64 time samples of bounds per object";
        return false;
    }
    const uint64_t n_boxes = in.offsets[n];
    const uint64_t seg_cap = n / (SMALL + 1) + 2;
    const uint64_t chunk_cap = n / CHUNK + seg_cap + 2;
    const uint64_t small_cap = n / (per_leaf + 1) + 2;
    const uint64_t node_cap = 2ull * n + 2;

    Arena& ar = b.ar;
    Box* d_bounds = ar.get<Box>(n_boxes);
    uint32_t* d_offs = ar.get<uint32_t>(n + 1);
    uint32_t* keys = ar.get<uint32_t>(n);
    ObjInfo* recs = ar.get<ObjInfo>(n);
    uint32_t* posL = ar.get<uint32_t>(n);
    uint32_t* posR = ar.get<uint32_t>(n);
    uint32_t* flag = ar.get<uint32_t>(4);
    Seg* segs = ar.get<Seg>(seg_cap);
    Seg* segs_next = ar.get<Seg>(seg_cap);
    uint32_t* nseg_d = ar.get<uint32_t>(4);
    SegWork* work = ar.get<SegWork>(seg_cap);
    uint32_t* seg_nch = ar.get<uint32_t>(seg_cap + 1);
    uint32_t* seg_choff = ar.get<uint32_t>(seg_cap + 1);
    Chunk* chunks = ar.get<Chunk>(chunk_cap);
    Box* ch_cb = ar.get<Box>(chunk_cap);
    ChunkBins* ch_bins = ar.get<ChunkBins>(chunk_cap);
    uint32_t* ch_true = ar.get<uint32_t>(chunk_cap);
    uint32_t* ch_toff = ar.get<uint32_t>(chunk_cap);
    Seg* cand = ar.get<Seg>(2 * seg_cap);
    uint32_t* fbig = ar.get<uint32_t>(2 * seg_cap);
    uint32_t* fsmall = ar.get<uint32_t>(2 * seg_cap);
    uint32_t* obig = ar.get<uint32_t>(2 * seg_cap + 1);
    uint32_t* osmall = ar.get<uint32_t>(2 * seg_cap + 1);
    Seg* small = ar.get<Seg>(small_cap);
    uint32_t* small_need = ar.get<uint32_t>(small_cap);
    uint32_t* small_off = ar.get<uint32_t>(small_cap + 1);
    uint32_t* na = ar.get<uint32_t>(node_cap);
    uint32_t* nb = ar.get<uint32_t>(node_cap);
    uint32_t* meta = ar.get<uint32_t>(node_cap);
    if (ar.err != hipSuccess) {
        why = std::string("device builder allocation: ") + hipGetErrorString(ar.err);
        return false;
    }

    DB_HIP(hipStreamCreateWithFlags(&b.st, hipStreamNonBlocking));
    const hipStream_t st = b.st;

    DB_HIP(hipMemcpyAsync(d_bounds, in.bounds, n_boxes * sizeof(Box), hipMemcpyHostToDevice, st));
    DB_HIP(hipMemcpyAsync(d_offs, in.offsets, (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    DB_HIP(hipMemsetAsync(flag, 0, 4 * sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_init, dim3(blocks_for(n)), dim3(BLOCK), 0, st, d_bounds, d_offs, n, keys, recs, flag);
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(node_cap)), dim3(BLOCK), 0, st, meta, UNUSED, node_cap);
    DB_HIP(hipGetLastError());
    uint32_t hflag = 0;
    DB_HIP(hipMemcpyAsync(&hflag, flag, 4, hipMemcpyDeviceToHost, st));
    DB_HIP(hipStreamSynchronize(st));
    if (hflag) {
        why = "object boxes hold NaN or negative zero (their unions depend on order)";
        return false;
    }

    // Root: a leaf, a one-lane subtree, or the first level.
    uint32_t node_base = 1;  // ids handed out so far (the root is 0)
    uint32_t n_small = 0;
    uint32_t nseg = 0;
    const Seg root{0, n, 0, 0};
    if (n <= per_leaf) {
        const uint32_t m[3] = {0, n, LEAF};
        DB_HIP(hipMemcpyAsync(na, &m[0], 4, hipMemcpyHostToDevice, st));
        DB_HIP(hipMemcpyAsync(nb, &m[1], 4, hipMemcpyHostToDevice, st));
        DB_HIP(hipMemcpyAsync(meta, &m[2], 4, hipMemcpyHostToDevice, st));
        DB_HIP(hipStreamSynchronize(st));
    } else if (n <= SMALL) {
        DB_HIP(hipMemcpyAsync(small, &root, sizeof(Seg), hipMemcpyHostToDevice, st));
        DB_HIP(hipStreamSynchronize(st));
        n_small = 1;
    } else {
        DB_HIP(hipMemcpyAsync(segs, &root, sizeof(Seg), hipMemcpyHostToDevice, st));
        nseg = 1;
    }

    // Levels of large nodes.
    while (nseg > 0) {
        DB_HIP(hipMemcpyAsync(nseg_d, &nseg, 4, hipMemcpyHostToDevice, st));
        const unsigned gs = blocks_for(nseg);
        const uint64_t max_chunks = n / CHUNK + nseg + 1;
        hipLaunchKernelGGL(k_seg_nchunks, dim3(gs), dim3(BLOCK), 0, st, segs, nseg_d, seg_nch);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, seg_nch, seg_choff, nseg);
        hipLaunchKernelGGL(k_chunk_fill, dim3(gs), dim3(BLOCK), 0, st, segs, nseg_d, seg_choff, chunks);
        const uint32_t* nch_d = seg_choff + nseg;  // the scan's total
        hipLaunchKernelGGL(k_chunk_cb, dim3((unsigned)max_chunks), dim3(BLOCK), 0, st, chunks, nch_d, recs, ch_cb);
        hipLaunchKernelGGL(k_seg_prep, dim3(gs), dim3(BLOCK), 0, st, segs, nseg_d, seg_choff, ch_cb, work, keys, recs,
                           sah_only ? 1u : 0u);
        hipLaunchKernelGGL(k_chunk_bins, dim3((unsigned)max_chunks), dim3(BLOCK), 0, st, chunks, nch_d, work, recs,
                           ch_bins);
        hipLaunchKernelGGL(k_seg_split, dim3(nseg), dim3(64), 0, st, nseg_d, seg_choff, ch_bins, work);
        hipLaunchKernelGGL(k_chunk_count, dim3((unsigned)max_chunks), dim3(BLOCK), 0, st, chunks, nch_d, work, recs,
                           ch_true);
        hipLaunchKernelGGL(k_seg_trues, dim3(gs), dim3(BLOCK), 0, st, nseg_d, seg_choff, ch_true, ch_toff, work);
        hipLaunchKernelGGL(k_chunk_rank, dim3((unsigned)max_chunks), dim3(BLOCK), 0, st, chunks, nch_d, segs, work,
                           recs, ch_toff, posL, posR);
        hipLaunchKernelGGL(k_chunk_swap, dim3((unsigned)max_chunks), dim3(BLOCK), 0, st, chunks, nch_d, segs, work,
                           posL, posR, keys, recs);
        hipLaunchKernelGGL(k_children, dim3(gs), dim3(BLOCK), 0, st, segs, nseg_d, work, node_base, per_leaf, na, nb,
                           meta, cand, fbig, fsmall);
        const uint32_t m2 = 2 * nseg;
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, fbig, obig, m2);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, fsmall, osmall, m2);
        hipLaunchKernelGGL(k_compact, dim3(blocks_for(m2)), dim3(BLOCK), 0, st, cand, fbig, obig, m2, segs_next, 0u);
        hipLaunchKernelGGL(k_compact, dim3(blocks_for(m2)), dim3(BLOCK), 0, st, cand, fsmall, osmall, m2, small,
                           n_small);
        DB_HIP(hipGetLastError());
        uint32_t totals[2];
        DB_HIP(hipMemcpyAsync(&totals[0], obig + m2, 4, hipMemcpyDeviceToHost, st));
        DB_HIP(hipMemcpyAsync(&totals[1], osmall + m2, 4, hipMemcpyDeviceToHost, st));
        DB_HIP(hipStreamSynchronize(st));
        node_base += m2;
        n_small += totals[1];
        nseg = totals[0];
        if (nseg > seg_cap || n_small > small_cap) {
            why = "device builder ran out of node lists";
            return false;
        }
        std::swap(segs, segs_next);
    }

    // Subtrees of small nodes, one lane each.
    if (n_small > 0) {
        hipLaunchKernelGGL(k_small_need, dim3(blocks_for(n_small)), dim3(BLOCK), 0, st, small, n_small, small_need);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, small_need, small_off, n_small);
        uint32_t reserve = 0;
        DB_HIP(hipMemcpyAsync(&reserve, small_off + n_small, 4, hipMemcpyDeviceToHost, st));
        DB_HIP(hipStreamSynchronize(st));
        if ((uint64_t)node_base + reserve > node_cap) {
            why = "device builder ran out of node ids";
            return false;
        }
        hipLaunchKernelGGL(k_small, dim3(blocks_for(n_small)), dim3(BLOCK), 0, st, small, n_small, small_off, node_base,
                           per_leaf, sah_only ? 1u : 0u, keys, recs, na, nb, meta);
        DB_HIP(hipGetLastError());
        node_base += reserve;
    }
    b.n = n;
    b.stride = stride;
    b.n_nodes = node_base;
    b.d_bounds = d_bounds;
    b.d_offs = d_offs;
    b.keys = keys;
    b.na = na;
    b.nb = nb;
    b.meta = meta;
    return true;
}

}  // namespace

bool build_bvh4_device(const BuildInput& in, uint32_t per_leaf, HostBvh4& out, std::string& why, double* ms) {
    const auto t0 = std::chrono::steady_clock::now();
    out = HostBvh4{};
    const uint32_t n = in.count;
    if (n == 0) {
        if (ms) *ms = 0.0;
        return true;
    }
    Bvh2Dev b;
    if (!build_bvh2_dev(in, per_leaf, false, b, why)) return false;
    Arena& ar = b.ar;
    const hipStream_t st = b.st;
    const uint32_t n_nodes = b.n_nodes, stride = b.stride;
    const Box* d_bounds = b.d_bounds;
    const uint32_t* d_offs = b.d_offs;
    uint32_t *keys = b.keys, *na = b.na, *nb = b.nb, *meta = b.meta;

    // Bounds, deepest nodes first.
    uint32_t* hist = ar.get<uint32_t>(256);
    uint32_t* cursor = ar.get<uint32_t>(256);
    uint32_t* sorted = ar.get<uint32_t>(n_nodes);
    uint32_t* blen = ar.get<uint32_t>(n_nodes);
    Box* nbounds = ar.get<Box>((uint64_t)n_nodes * stride);
    if (ar.err != hipSuccess) {
        why = std::string("device builder allocation: ") + hipGetErrorString(ar.err);
        return false;
    }
    DB_HIP(hipMemsetAsync(hist, 0, 256 * 4, st));
    hipLaunchKernelGGL(k_depth_hist, dim3(blocks_for(n_nodes)), dim3(BLOCK), 0, st, meta, n_nodes, hist);
    std::vector<uint32_t> h(256), off(257, 0);
    DB_HIP(hipMemcpyAsync(h.data(), hist, 256 * 4, hipMemcpyDeviceToHost, st));
    DB_HIP(hipStreamSynchronize(st));
    uint32_t max_depth = 0;
    for (uint32_t d = 0; d < 256; d++) {
        off[d + 1] = off[d] + h[d];
        if (h[d]) max_depth = d;
    }
    DB_HIP(hipMemcpyAsync(cursor, off.data(), 256 * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_depth_scatter, dim3(blocks_for(n_nodes)), dim3(BLOCK), 0, st, meta, n_nodes, cursor, sorted);
    for (int d = (int)max_depth; d >= 0; d--) {
        if (!h[d]) continue;
        hipLaunchKernelGGL(k_bounds, dim3(blocks_for(h[d])), dim3(BLOCK), 0, st, sorted + off[d], h[d], na, nb, meta,
                           keys, d_bounds, d_offs, stride, blen, nbounds);
    }
    DB_HIP(hipGetLastError());

    // The BVH4 records on the device, then back to the host.
    uint32_t base_depth = 0;
    uint32_t rmeta = 0, rblen = 0;
    DB_HIP(hipMemcpyAsync(&rmeta, meta, 4, hipMemcpyDeviceToHost, st));
    DB_HIP(hipMemcpyAsync(&rblen, blen, 4, hipMemcpyDeviceToHost, st));
    DB_HIP(hipStreamSynchronize(st));
    out.order.resize(n);
    out.root_bounds.resize(rblen);
    DB_HIP(hipMemcpyAsync(out.root_bounds.data(), nbounds, rblen * sizeof(Box), hipMemcpyDeviceToHost, st));
    uint32_t* leafd = ar.get<uint32_t>(1);
    if (ar.err != hipSuccess) {
        why = std::string("device builder allocation: ") + hipGetErrorString(ar.err);
        return false;
    }
    DB_HIP(hipMemsetAsync(leafd, 0, 4, st));
    hipLaunchKernelGGL(k_leaf_depth, dim3(blocks_for(n_nodes)), dim3(BLOCK), 0, st, meta, n_nodes, leafd);
    DB_HIP(hipGetLastError());
    DB_HIP(hipMemcpyAsync(&base_depth, leafd, 4, hipMemcpyDeviceToHost, st));
    DB_HIP(hipMemcpyAsync(out.order.data(), keys, n * 4ull, hipMemcpyDeviceToHost, st));
    DB_HIP(hipStreamSynchronize(st));
    out.depth = base_depth / 2 + 1;
    if (rmeta & LEAF) {
        // A root leaf: no BVH4 node (the host collapse's root leaf reference).
        uint32_t ra = 0, rb = 0;
        DB_HIP(hipMemcpyAsync(&ra, na, 4, hipMemcpyDeviceToHost, st));
        DB_HIP(hipMemcpyAsync(&rb, nb, 4, hipMemcpyDeviceToHost, st));
        DB_HIP(hipStreamSynchronize(st));
        out.root = LEAF | (ra << 2) | (rb - ra);
        out.node_count = 1;
        out.leaf_count = 1;
    } else {
        C4* c4 = ar.get<C4>(n_nodes);
        uint32_t* desc = ar.get<uint32_t>(n_nodes);
        uint32_t* mdesc = ar.get<uint32_t>(n_nodes);
        uint32_t* slot = ar.get<uint32_t>(n_nodes);
        uint32_t* S = ar.get<uint32_t>(n_nodes);
        uint32_t* M = ar.get<uint32_t>(n_nodes);
        uint8_t* table = ar.get<uint8_t>(8 * 48);
        if (ar.err != hipSuccess) {
            why = std::string("device builder allocation: ") + hipGetErrorString(ar.err);
            return false;
        }
        DB_HIP(hipMemcpyAsync(table, traversal_order_table(), 8 * 48, hipMemcpyHostToDevice, st));
        for (uint32_t d = 0; d <= max_depth; d += 2)
            if (h[d])
                hipLaunchKernelGGL(k_c4_prep, dim3(blocks_for(h[d])), dim3(BLOCK), 0, st, sorted + off[d], h[d], na, nb,
                                   meta, blen, c4);
        for (int d = (int)(max_depth & ~1u); d >= 0; d -= 2)
            if (h[d])
                hipLaunchKernelGGL(k_c4_up, dim3(blocks_for(h[d])), dim3(BLOCK), 0, st, sorted + off[d], h[d], meta, c4,
                                   desc, mdesc);
        const uint32_t init[3] = {0u, 1u, 0u};  // the root's slot, first free slot, motion records before it
        DB_HIP(hipMemcpyAsync(slot, &init[0], 4, hipMemcpyHostToDevice, st));
        DB_HIP(hipMemcpyAsync(S, &init[1], 4, hipMemcpyHostToDevice, st));
        DB_HIP(hipMemcpyAsync(M, &init[2], 4, hipMemcpyHostToDevice, st));
        for (uint32_t d = 0; d <= max_depth; d += 2)
            if (h[d])
                hipLaunchKernelGGL(k_c4_down, dim3(blocks_for(h[d])), dim3(BLOCK), 0, st, sorted + off[d], h[d], meta,
                                   c4, desc, mdesc, slot, S, M);
        DB_HIP(hipGetLastError());
        // Totals: the root's subtree holds every node and motion record.
        uint32_t tot[2] = {0, 0};
        C4 rc4{};
        DB_HIP(hipMemcpyAsync(&tot[0], desc, 4, hipMemcpyDeviceToHost, st));
        DB_HIP(hipMemcpyAsync(&tot[1], mdesc, 4, hipMemcpyDeviceToHost, st));
        DB_HIP(hipMemcpyAsync(&rc4, c4, sizeof(C4), hipMemcpyDeviceToHost, st));
        DB_HIP(hipStreamSynchronize(st));
        const uint64_t n4 = 1ull + tot[0];
        const uint64_t nmot = (uint64_t)tot[1] + (rc4.blen > 1 ? rc4.blen : 0u);
        GpuNode* nodes4 = ar.get<GpuNode>(n4);
        GpuBox4* motion4 = ar.get<GpuBox4>(nmot);
        uint32_t* cnt_d = ar.get<uint32_t>(2);
        if (ar.err != hipSuccess) {
            why = std::string("device builder allocation: ") + hipGetErrorString(ar.err);
            return false;
        }
        for (uint32_t d = 0; d <= max_depth; d += 2)
            if (h[d])
                hipLaunchKernelGGL(k_c4_emit, dim3(blocks_for(h[d])), dim3(BLOCK), 0, st, sorted + off[d], h[d], na,
                                   nb, meta, blen, nbounds, stride, c4, slot, mdesc, M, table, nodes4, motion4);
        DB_HIP(hipMemsetAsync(cnt_d, 0, 8, st));
        hipLaunchKernelGGL(k_c4_count, dim3(blocks_for(n_nodes)), dim3(BLOCK), 0, st, meta, n_nodes, c4, cnt_d);
        DB_HIP(hipGetLastError());
        out.nodes.resize(n4);
        out.motion.resize(nmot);
        uint32_t cnts[2];
        DB_HIP(hipMemcpyAsync(out.nodes.data(), nodes4, n4 * sizeof(GpuNode), hipMemcpyDeviceToHost, st));
        DB_HIP(hipMemcpyAsync(out.motion.data(), motion4, nmot * sizeof(GpuBox4), hipMemcpyDeviceToHost, st));
        DB_HIP(hipMemcpyAsync(cnts, cnt_d, 8, hipMemcpyDeviceToHost, st));
        DB_HIP(hipStreamSynchronize(st));
        out.root = 0;
        out.node_count = (uint64_t)cnts[0] + cnts[1];  // the host collapse counts leaf children twice
        out.leaf_count = cnts[1];
    }
    if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

bool build_light_bvh2_device(const BuildInput& in, ArrayBvh2& t, std::vector<uint32_t>& order, std::string& why) {
    t = ArrayBvh2{};
    order.clear();
    if (in.count == 0) return true;
    Bvh2Dev b;
    if (!build_bvh2_dev(in, 1, true, b, why)) return false;
    t.stride = 0;
    t.root = 0;
    t.a.resize(b.n_nodes);
    t.b.resize(b.n_nodes);
    t.meta.resize(b.n_nodes);
    order.resize(b.n);
    DB_HIP(hipMemcpyAsync(t.a.data(), b.na, b.n_nodes * 4ull, hipMemcpyDeviceToHost, b.st));
    DB_HIP(hipMemcpyAsync(t.b.data(), b.nb, b.n_nodes * 4ull, hipMemcpyDeviceToHost, b.st));
    DB_HIP(hipMemcpyAsync(t.meta.data(), b.meta, b.n_nodes * 4ull, hipMemcpyDeviceToHost, b.st));
    DB_HIP(hipMemcpyAsync(order.data(), b.keys, b.n * 4ull, hipMemcpyDeviceToHost, b.st));
    DB_HIP(hipStreamSynchronize(b.st));
    return true;
}

}  // namespace psy
