// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
//
// Ray batch and ray-stream scheduler: restates src/ray.rs.
#pragma once

#include <cassert>
#include <cstdint>
#include <vector>

#include "psyref_math.h"

namespace psyref {

static const uint8_t OCCLUSION_FLAG = 1;  // ray.rs:9
static const uint8_t DONE_FLAG = 2;       // ray.rs:10

// ray.rs:15-22
struct Ray {
    Vec3 orig, dir;
    float time, wavelength, max_t;
};

// ray.rs:24-44: hot/cold SoA
struct RayBatch {
    struct Hot { Vec3 orig_local, dir_inv_local; float max_t, time; uint8_t flags; };
    struct Cold { Vec3 orig, dir; float wavelength; };
    std::vector<Hot> hot;
    std::vector<Cold> cold;

    // ray.rs:66-79
    void push(const Ray& r, bool is_occlusion) {
        hot.push_back(Hot{r.orig, r.dir, r.max_t, r.time, (uint8_t)(is_occlusion ? OCCLUSION_FLAG : 0)});
        cold.push_back(Cold{r.orig, r.dir, r.wavelength});
    }
    // ray.rs:86-98
    void set_from_ray(const Ray& r, bool is_occlusion, size_t idx) {
        hot[idx].orig_local = r.orig;
        hot[idx].dir_inv_local = recip(r.dir);
        hot[idx].max_t = r.max_t;
        hot[idx].time = r.time;
        hot[idx].flags = is_occlusion ? OCCLUSION_FLAG : 0;
        cold[idx].orig = r.orig;
        cold[idx].dir = r.dir;
        cold[idx].wavelength = r.wavelength;
    }
    void clear() { hot.clear(); cold.clear(); }
    size_t size() const { return hot.size(); }
    // ray.rs:122-127
    void update_local(size_t idx, const Transform& xf) {
        hot[idx].orig_local = xform_point(cold[idx].orig, xf);
        hot[idx].dir_inv_local = recip(xform_vector(cold[idx].dir, xf));
    }
    Vec3 orig(size_t i) const { return cold[i].orig; }
    Vec3 dir(size_t i) const { return cold[i].dir; }
    Vec3 orig_local(size_t i) const { return hot[i].orig_local; }
    Vec3 dir_inv_local(size_t i) const { return hot[i].dir_inv_local; }
    float time(size_t i) const { return hot[i].time; }
    float max_t(size_t i) const { return hot[i].max_t; }
    void set_max_t(size_t i, float t) { hot[i].max_t = t; }
    bool is_occlusion(size_t i) const { return (hot[i].flags & OCCLUSION_FLAG) != 0; }
    bool is_done(size_t i) const { return (hot[i].flags & DONE_FLAG) != 0; }
    void mark_done(size_t i) { hot[i].flags |= DONE_FLAG; }
};

// ray.rs:197-401.  Lanes are u16 index stacks with an `end_len` watermark;
// tasks name a lane and a start index.  The storage vector keeps entries past
// `len` alive, as the reference's Vec::set_len trick does.
struct RayStack {
    struct Lane {
        std::vector<uint16_t> idxs;  // capacity-backed storage
        size_t len = 0;              // logical length
        size_t end_len = 0;
        void push(uint16_t v) {
            if (len < idxs.size()) idxs[len] = v; else idxs.push_back(v);
            len++;
        }
    };
    struct Task { size_t lane, start_idx; };
    std::vector<Lane> lanes;
    std::vector<Task> tasks;

    bool is_empty() const { return tasks.empty(); }
    void ensure_lane_count(size_t n) { while (lanes.size() < n) lanes.emplace_back(); }
    size_t ray_count_in_next_task() const {
        const Task& t = tasks.back();
        return lanes[t.lane].end_len - t.start_idx;
    }
    size_t next_task_ray_idx(size_t i) const {
        const Task& t = tasks.back();
        return lanes[t.lane].idxs[i + t.start_idx];
    }
    void clear() {
        for (auto& l : lanes) { l.len = 0; l.end_len = 0; }
        tasks.clear();
    }
    void push_ray_index(size_t ray_idx, size_t lane) { lanes[lane].push((uint16_t)ray_idx); }
    // ray.rs:262-275
    bool push_lane_to_task(size_t l) {
        if (lanes[l].end_len < lanes[l].len) {
            tasks.push_back(Task{l, lanes[l].end_len});
            lanes[l].end_len = lanes[l].len;
            return true;
        }
        return false;
    }
    void push_lanes_to_tasks(const size_t* ls, size_t n) { for (size_t i = 0; i < n; i++) push_lane_to_task(ls[i]); }
    // ray.rs:288-312
    void duplicate_next_task() {
        const Task t = tasks.back();
        Lane& lane = lanes[t.lane];
        size_t start = t.start_idx, end = lane.end_len;
        for (size_t i = start; i < end; i++) lane.push(lane.idxs[i]);
        tasks.push_back(Task{t.lane, end});
        lane.end_len = lane.len;
    }
    // ray.rs:315-319
    void pop_task() {
        Task t = tasks.back();
        tasks.pop_back();
        lanes[t.lane].end_len = t.start_idx;
        lanes[t.lane].len = t.start_idx;
    }
    // ray.rs:322-334
    template <typename F> void do_next_task(F&& f) {
        const Task t = tasks.back();
        size_t end = lanes[t.lane].end_len;
        for (size_t i = t.start_idx; i < end; i++) f((size_t)lanes[t.lane].idxs[i]);
    }
    template <typename F> void pop_do_next_task(F&& f) { do_next_task(f); pop_task(); }
    // ray.rs:350-382: f returns a 4-bit child mask; the index is appended to
    // each lane in [0, out_lanes) whose bit is set.
    template <typename F> void pop_do_next_task_and_push_rays(size_t out_lanes, F&& f) {
        Task t = tasks.back();
        tasks.pop_back();
        size_t end = lanes[t.lane].end_len;
        lanes[t.lane].end_len = t.start_idx;
        lanes[t.lane].len = t.start_idx;
        for (size_t i = t.start_idx; i < end; i++) {
            uint16_t ray_idx = lanes[t.lane].idxs[i];
            unsigned mask = f((size_t)ray_idx);
            for (size_t l = 0; l < out_lanes; l++)
                if (mask & (1u << l)) lanes[l].push(ray_idx);
        }
    }
};

}  // namespace psyref
