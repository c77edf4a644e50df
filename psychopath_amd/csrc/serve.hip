// Render request server (include/psy_render.h, psy_server_*).
//
// A server owns a set of renderers (one device-resident scene each, with its
// own stream and path-state pool) and a worker thread.  Requests are queued
// on submission and admitted in arrival order as soon as the HBM budget has
// room for their film and their renderer has a free request slot; an
// admitted request keeps its film and parameters on the device until the
// caller collects the image.  For every renderer the worker keeps up to two
// waves in flight: each wave is filled, in admission order, with the next
// unrendered paths of the renderer's active requests (continuous batching:
// a request that arrives while others render joins the next wave), and the
// renderers' streams run concurrently on the device.  A request's paths are
// dispatched in order, so each pixel still accumulates its samples in
// sample order: the image is the same as a render of the request alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "render_internal.h"

using namespace psy;
using namespace psy::api;
using namespace psy::render;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    uint32_t ren = 0;
    ReqParams rp{};
    std::vector<GpuXform> cam;
    uint64_t total = 0, next = 0;
    uint64_t retired = 0;  // paths of this request in waves that have finished
    uint64_t film_bytes = 0;
    float* film = nullptr;
    uint32_t slot = 0xffffffffu;
    uint64_t waves = 0;
    uint32_t readers = 0;  // psy_server_read copies running from the film
    bool admitted = false, done = false;
    int error = PSY_OK;
    std::string msg;       // why the request failed, when it did
    Clock::time_point t_submit, t_admit, t_done;
};

struct InFlight {
    uint32_t buf;                    // plan buffer / event index
    std::vector<Request*> finishing; // requests whose last paths are in this wave
    std::vector<std::pair<Request*, uint64_t>> reached;  // each request's dispatched paths after this wave
};

struct RenState {
    psy_renderer* r = nullptr;
    GpuXform* cams = nullptr;        // CAM_SLOTS camera entries per request slot
    WavePlan* host_plans = nullptr;  // pinned, two buffers
    ReqParams* host_params = nullptr;  // pinned staging, one per request slot
    GpuXform* host_cams = nullptr;     // pinned staging, CAM_SLOTS per slot
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<Request*> active;    // admitted, not fully dispatched
    std::vector<uint32_t> free_slots;
    std::deque<InFlight> inflight;
    uint64_t seq = 0;
};

}  // namespace

struct psy_server {
    std::vector<RenState> rens;
    uint64_t budget = 0;      // bytes, 0 = unlimited
    uint64_t fixed_bytes = 0; // scenes and path state
    uint64_t film_bytes = 0;  // films of admitted, uncollected requests
    uint64_t pool_bytes = 0;  // released films kept for reuse
    std::multimap<uint64_t, float*> pool;
    uint32_t max_active = MAX_SEGS;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queued;
    std::map<uint64_t, Request*> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    int worker_error = PSY_OK;
    std::string worker_msg;
    std::thread worker;
    uint32_t fail_allocs = 0;  // film allocations to fail on purpose (psy_server_fail_film_allocs)

    uint64_t waves = 0, paths = 0;
};

namespace {

// Copies a request's parameters and camera into its slot through the slot's
// pinned staging: a slot is reused only after every wave that read it has
// retired, so the staging is never rewritten under a pending copy.
int upload_params(RenState& rs, Request* q) {
    const hipStream_t s = rs.r->stream;
    GpuXform* hc = rs.host_cams + (uint64_t)CAM_SLOTS * q->slot;
    std::memcpy(hc, q->cam.data(), q->cam.size() * sizeof(GpuXform));
    rs.host_params[q->slot] = q->rp;
    PSY_HIP(hipMemcpyAsync(rs.cams + (uint64_t)CAM_SLOTS * q->slot, hc, q->cam.size() * sizeof(GpuXform),
                           hipMemcpyHostToDevice, s));
    PSY_HIP(hipMemcpyAsync(rs.r->reqs + q->slot, rs.host_params + q->slot, sizeof(ReqParams), hipMemcpyHostToDevice,
                           s));
    return PSY_OK;
}

// A film of `bytes` from the pool of released films, or a new allocation;
// pooled films are freed first when the budget needs their room.
float* film_alloc(psy_server* sv, uint64_t bytes) {
    auto it = sv->pool.find(bytes);
    if (it != sv->pool.end()) {
        float* p = it->second;
        sv->pool.erase(it);
        sv->pool_bytes -= bytes;
        return p;
    }
    while (sv->budget && !sv->pool.empty() &&
           sv->fixed_bytes + sv->film_bytes + sv->pool_bytes + bytes > sv->budget) {
        auto victim = sv->pool.begin();
        sv->pool_bytes -= victim->first;
        (void)hipFree(victim->second);
        sv->pool.erase(victim);
    }
    float* p = nullptr;
    if (sv->fail_allocs > 0) {
        sv->fail_allocs--;
        return nullptr;
    }
    if (hipMalloc((void**)&p, bytes) == hipSuccess) return p;
    // Device memory is short: give the pooled films back and try once more.
    for (auto& kv : sv->pool) (void)hipFree(kv.second);
    sv->pool.clear();
    sv->pool_bytes = 0;
    if (hipMalloc((void**)&p, bytes) == hipSuccess) return p;
    return nullptr;
}

// True while some renderer has admitted requests or waves in flight, whose
// completion can still free device memory.
bool work_pending(const psy_server* sv) {
    for (const RenState& rs : sv->rens)
        if (!rs.active.empty() || !rs.inflight.empty()) return true;
    return false;
}

// Admits queued requests in order while budget and slots allow (called with
// the lock held).
int admit(psy_server* sv) {
    while (!sv->queued.empty()) {
        Request* q = sv->queued.front();
        RenState& rs = sv->rens[q->ren];
        if (rs.free_slots.empty()) break;
        if (sv->budget && sv->fixed_bytes + sv->film_bytes + q->film_bytes > sv->budget) break;
        q->film = film_alloc(sv, q->film_bytes);
        if (!q->film) {
            // Retry once running requests free memory; with nothing running
            // nothing will, so the request fails instead of waiting forever.
            if (work_pending(sv)) break;
            sv->queued.pop_front();
            q->done = true;
            q->error = PSY_E_NOMEM;
            q->msg = "no device memory for the request's film";
            q->t_admit = q->t_done = Clock::now();
            sv->cv_done.notify_all();
            continue;
        }
        sv->queued.pop_front();
        q->slot = rs.free_slots.back();
        rs.free_slots.pop_back();
        q->rp.film = q->film;
        q->rp.cam_xf = rs.cams + (uint64_t)CAM_SLOTS * q->slot;
        int rc = launch_zero(q->film, q->film_bytes / sizeof(float), rs.r->stream);
        if (!rc) rc = upload_params(rs, q);
        if (rc) return rc;
        sv->film_bytes += q->film_bytes;
        q->admitted = true;
        q->t_admit = Clock::now();
        rs.active.push_back(q);
    }
    return PSY_OK;
}

// Retires finished waves of one renderer; returns true if any finished.
bool retire(psy_server* sv, RenState& rs) {
    bool any = false;
    while (!rs.inflight.empty()) {
        InFlight& f = rs.inflight.front();
        if (hipEventQuery(rs.ev[f.buf]) != hipSuccess) break;
        for (auto& qr : f.reached) qr.first->retired = qr.second;
        for (Request* q : f.finishing) {
            q->done = true;
            q->t_done = Clock::now();
            rs.free_slots.push_back(q->slot);
        }
        rs.inflight.pop_front();
        any = true;
    }
    return any;
}

// Fills and launches the next wave of one renderer (lock held).
int launch(psy_server* sv, RenState& rs, bool& launched) {
    launched = false;
    if (rs.inflight.size() >= 2 || rs.active.empty()) return PSY_OK;
    psy_renderer* r = rs.r;
    const uint32_t buf = (uint32_t)(rs.seq % 2);
    WavePlan& plan = rs.host_plans[buf];
    std::memset(&plan, 0, sizeof(plan));
    InFlight f{buf, {}, {}};
    uint32_t filled = 0, steps = 1;
    for (Request* q : rs.active) {
        if (filled == r->capacity || plan.nseg == MAX_SEGS) break;
        const uint64_t m = std::min<uint64_t>(q->total - q->next, r->capacity - filled);
        if (m == 0) continue;
        plan.seg[plan.nseg++] = WaveSeg{q->slot, filled, (uint32_t)m, 0, q->next};
        q->next += m;
        q->waves++;
        f.reached.emplace_back(q, q->next);
        filled += (uint32_t)m;
        steps = std::max(steps, q->rp.max_bounces + 1);
        if (q->next == q->total) f.finishing.push_back(q);
    }
    plan.total = filled;
    rs.active.erase(std::remove_if(rs.active.begin(), rs.active.end(), [](Request* q) { return q->next == q->total; }),
                    rs.active.end());
    if (filled == 0) return PSY_OK;
    PSY_HIP(hipMemcpyAsync(r->plans + buf, &plan, sizeof(WavePlan), hipMemcpyHostToDevice, r->stream));
    int rc = launch_set_wave(r, buf);
    if (!rc) rc = run_wave(r, steps);
    if (rc) return rc;
    PSY_HIP(hipEventRecord(rs.ev[buf], r->stream));
    rs.inflight.push_back(std::move(f));
    rs.seq++;
    sv->waves++;
    sv->paths += filled;
    launched = true;
    return PSY_OK;
}

void worker_main(psy_server* sv) {
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        if (sv->stop) {
            // psy_server_destroy: requests not yet admitted are dropped and
            // no further waves start; the waves in flight still finish
            // before the worker returns.
            for (Request* q : sv->queued) {
                q->done = true;
                q->error = PSY_E_STATE;
                q->msg = "server destroyed";
            }
            sv->queued.clear();
            for (RenState& rs : sv->rens) rs.active.clear();
        }
        bool busy = false;
        int rc = admit(sv);
        bool finished = false;
        for (RenState& rs : sv->rens) {
            if (!rc && retire(sv, rs)) finished = true;
            bool launched = false;
            if (!rc) rc = launch(sv, rs, launched);
            busy = busy || launched || !rs.inflight.empty();
        }
        if (finished) {
            // Slots freed: admit before the next round.
            if (!rc) rc = admit(sv);
            sv->cv_done.notify_all();
        }
        if (rc) {
            sv->worker_error = rc;
            sv->worker_msg = psy_last_error();
            for (auto& kv : sv->reqs)
                if (!kv.second->done) {
                    kv.second->done = true;
                    kv.second->error = rc;
                }
            sv->cv_done.notify_all();
            return;
        }
        if (sv->stop && !busy) return;
        if (busy) {
            // Waves are running: poll their events without holding the lock
            // for long.
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(50));
            lk.lock();
        } else {
            sv->cv_work.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
}

}  // namespace

extern "C" {

int psy_server_create(psy_renderer* const* renderers, uint32_t count, const psy_server_config* cfg,
                      psy_server** out) {
    if (!renderers || count == 0 || !out) return fail(PSY_E_INVALID, "server needs renderers");
    *out = nullptr;
    psy_server* sv = new (std::nothrow) psy_server();
    if (!sv) return fail(PSY_E_NOMEM, "out of host memory");
    auto bail = [&](int code) {
        for (RenState& rs : sv->rens) {
            if (rs.host_plans) (void)hipHostFree(rs.host_plans);
            if (rs.host_params) (void)hipHostFree(rs.host_params);
            if (rs.host_cams) (void)hipHostFree(rs.host_cams);
            for (hipEvent_t e : rs.ev)
                if (e) (void)hipEventDestroy(e);
        }
        delete sv;
        return code;
    };
    if (cfg) {
        sv->budget = cfg->hbm_budget;
        if (cfg->max_active > 0) sv->max_active = std::min<uint32_t>(cfg->max_active, MAX_SEGS);
    }
    sv->rens.resize(count);
    for (uint32_t i = 0; i < count; i++) {
        psy_renderer* r = renderers[i];
        if (!r) return bail(fail(PSY_E_INVALID, "renderer is NULL"));
        for (uint32_t j = 0; j < i; j++)
            if (renderers[j] == r) return bail(fail(PSY_E_INVALID, "a renderer is listed twice"));
        RenState& rs = sv->rens[i];
        rs.r = r;
        rs.cams = r->cam;  // CAM_SLOTS camera entries for each of the MAX_SEGS request slots
        hipError_t e = hipHostMalloc((void**)&rs.host_plans, 2 * sizeof(WavePlan), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&rs.host_params, MAX_SEGS * sizeof(ReqParams), hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&rs.host_cams, (uint64_t)MAX_SEGS * CAM_SLOTS * sizeof(GpuXform),
                              hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&rs.ev[0], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&rs.ev[1], hipEventDisableTiming);
        if (e != hipSuccess) return bail(fail(PSY_E_DEVICE, std::string("server setup: ") + hipGetErrorString(e)));
        for (uint32_t s = sv->max_active; s-- > 0;) rs.free_slots.push_back(s);
        sv->fixed_bytes += r->device_bytes() + r->scene->device_bytes;
    }
    try {
        sv->worker = std::thread(worker_main, sv);
    } catch (...) {
        return bail(fail(PSY_E_NOMEM, "cannot start the server thread"));
    }
    *out = sv;
    return PSY_OK;
}

int psy_server_destroy(psy_server* sv) {
    if (!sv) return PSY_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    for (RenState& rs : sv->rens) {
        if (rs.r && rs.r->stream) (void)hipStreamSynchronize(rs.r->stream);
        if (rs.host_plans) (void)hipHostFree(rs.host_plans);
        if (rs.host_params) (void)hipHostFree(rs.host_params);
        if (rs.host_cams) (void)hipHostFree(rs.host_cams);
        for (hipEvent_t e : rs.ev)
            if (e) (void)hipEventDestroy(e);
    }
    for (auto& kv : sv->reqs) {
        (void)hipFree(kv.second->film);
        delete kv.second;
    }
    for (auto& kv : sv->pool) (void)hipFree(kv.second);
    delete sv;
    return PSY_OK;
}

int psy_server_submit(psy_server* sv, uint32_t renderer, const psy_camera_desc* cam, const psy_render_settings* st,
                      uint64_t* id) {
    if (!sv || !id) return fail(PSY_E_INVALID, "NULL argument");
    if (renderer >= sv->rens.size()) return fail(PSY_E_INVALID, "renderer index out of range");
    Request* q = new (std::nothrow) Request();
    if (!q) return fail(PSY_E_NOMEM, "out of host memory");
    int rc = make_params(cam, st, q->rp, q->cam);
    if (rc) {
        delete q;
        return rc;
    }
    q->ren = renderer;
    q->total = (uint64_t)q->rp.rw * q->rp.rh * st->spp;
    q->film_bytes = (uint64_t)st->width * st->height * 3 * sizeof(float);
    q->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->worker_error) {
            delete q;
            return fail(sv->worker_error, "server stopped: " + sv->worker_msg);
        }
        if (sv->budget && sv->fixed_bytes + q->film_bytes > sv->budget) {
            delete q;
            return fail(PSY_E_RANGE, "request does not fit the HBM budget");
        }
        q->id = sv->next_id++;
        sv->reqs[q->id] = q;
        sv->queued.push_back(q);
        *id = q->id;
    }
    sv->cv_work.notify_all();
    return PSY_OK;
}

int psy_server_wait(psy_server* sv, uint64_t id, float* xyz, psy_request_stats* stats) {
    if (!sv) return fail(PSY_E_INVALID, "server is NULL");
    Request* q;
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->reqs.find(id);
        if (it == sv->reqs.end()) return fail(PSY_E_INVALID, "unknown request id");
        q = it->second;
        sv->cv_done.wait(lk, [&] { return q->done; });
        sv->reqs.erase(it);
    }
    int rc = q->error;
    const std::string msg = q->msg.empty() ? sv->worker_msg : q->msg;
    hipError_t e = hipSuccess;
    if (!rc && xyz) e = hipMemcpy(xyz, q->film, q->film_bytes, hipMemcpyDeviceToHost);
    if (stats) {
        stats->paths = q->total;
        stats->waves = q->waves;
        stats->queued_ms = std::chrono::duration<double, std::milli>(q->t_admit - q->t_submit).count();
        stats->latency_ms = std::chrono::duration<double, std::milli>(q->t_done - q->t_submit).count();
    }
    {
        // The film goes back to the pool (hipFree would synchronise the
        // device under the other requests' waves) once no psy_server_read
        // copies from it.
        std::unique_lock<std::mutex> lk(sv->mu);
        sv->cv_done.wait(lk, [&] { return q->readers == 0; });
        if (q->admitted) {
            sv->film_bytes -= q->film_bytes;
            sv->pool.emplace(q->film_bytes, q->film);
            sv->pool_bytes += q->film_bytes;
        }
    }
    sv->cv_work.notify_all();
    delete q;
    if (rc) return fail(rc, "request failed: " + msg);
    if (e != hipSuccess) return fail(PSY_E_DEVICE, std::string("image copy: ") + hipGetErrorString(e));
    return PSY_OK;
}

int psy_server_progress(psy_server* sv, uint64_t id, uint64_t* pixels_done) {
    if (!sv || !pixels_done) return fail(PSY_E_INVALID, "NULL argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return fail(PSY_E_INVALID, "unknown request id");
    const Request* q = it->second;
    if (q->error) return fail(q->error, "request failed: " + sv->worker_msg);
    // Waves finish in dispatch order and a pixel's samples are consecutive
    // paths, so every pixel whose last sample has retired is final.
    *pixels_done = q->done ? q->total / q->rp.spp : q->retired / q->rp.spp;
    return PSY_OK;
}

int psy_server_read(psy_server* sv, uint64_t id, uint64_t pixel_begin, uint64_t pixel_end, float* xyz) {
    if (!sv || (!xyz && pixel_end > pixel_begin)) return fail(PSY_E_INVALID, "NULL argument");
    const float* film;
    ReqParams rp;
    Request* q;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->reqs.find(id);
        if (it == sv->reqs.end()) return fail(PSY_E_INVALID, "unknown request id");
        q = it->second;
        if (q->error) return fail(q->error, "request failed: " + (q->msg.empty() ? sv->worker_msg : q->msg));
        const uint64_t done = q->done ? q->total / q->rp.spp : q->retired / q->rp.spp;
        if (pixel_begin > pixel_end || pixel_end > done) return fail(PSY_E_RANGE, "pixels not finished yet");
        film = q->film;
        rp = q->rp;
        q->readers++;
    }
    // The film stays with the request while this copy runs (the reader
    // count holds psy_server_wait from pooling it); the finished pixels are
    // no longer written, so the copy can run beside the waves still
    // rendering the rest.  Region pixels are contiguous in the film within a
    // row of the region.
    hipError_t e = hipSuccess;
    uint64_t p = pixel_begin;
    while (p < pixel_end && e == hipSuccess) {
        const uint64_t row_end = std::min<uint64_t>(pixel_end, (p / rp.rw + 1) * rp.rw);
        e = hipMemcpy(xyz + 3 * (p - pixel_begin), film + 3 * region_pixel(rp, p), (row_end - p) * 3 * sizeof(float),
                      hipMemcpyDeviceToHost);
        p = row_end;
    }
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        q->readers--;
    }
    sv->cv_done.notify_all();
    if (e != hipSuccess) return fail(PSY_E_DEVICE, std::string("film copy: ") + hipGetErrorString(e));
    return PSY_OK;
}

int psy_server_fail_film_allocs(psy_server* sv, uint32_t count) {
    if (!sv) return fail(PSY_E_INVALID, "server is NULL");
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->fail_allocs = count;
    return PSY_OK;
}

int psy_server_get_stats(psy_server* sv, psy_server_stats* st) {
    if (!sv || !st) return fail(PSY_E_INVALID, "NULL argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    st->waves = sv->waves;
    st->paths = sv->paths;
    st->queued = sv->queued.size();
    st->fixed_bytes = sv->fixed_bytes;
    st->film_bytes = sv->film_bytes + sv->pool_bytes;
    return PSY_OK;
}

}  // extern "C"
