// C ABI of the device side (include/mlgpu.h): scene upload, propagation,
// probe-ray first hits.  Never throws across the ABI; errors are status
// codes plus mlgpu_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "pool.hpp"
#include "scene_build.hpp"

struct mlgpu_scene {
    mlgpu::DeviceScene dev;
    std::vector<void*> buffers;
    std::vector<int32_t> scope_of_frame;
    int64_t num_nodes = 0;
    int64_t num_leaves = 0;
    int64_t device_bytes = 0;
    int device = 0;
    int num_cus = 0;
    mlgpu_tuning tuning{MLGPU_SCHEDULE_AUTO, 32, 48, 0, 32};
    // wavefront workspace, grown on demand.  Calls that use it are ordered
    // on the device: a call on another stream than the previous user first
    // waits for the event recorded after that user's kernels.
    std::mutex ws_mutex;
    void* ws = nullptr;
    int64_t ws_capacity = 0;
    hipEvent_t ws_event = nullptr;
    hipStream_t ws_stream = nullptr;
    // bunch arena (photons, results, bunch table), guarded like ws; the
    // table goes up from a pinned staging buffer that is rewritten only
    // after the previous upload from it has finished
    void* arena = nullptr;
    int64_t arena_photons = 0;
    int32_t arena_bunches = 0;
    mlgpu::BunchDev* staging = nullptr;
    int32_t staging_bunches = 0;
    hipEvent_t staging_event = nullptr;
};

namespace {

using mlgpu::set_last_error;

int hip_fail(hipError_t e, const char* what) {
    set_last_error(std::string(what) + ": " + hipGetErrorString(e));
    return MLGPU_ERR_HIP;
}

template <typename T>
int upload(mlgpu_scene* s, const std::vector<T>& v, const T** out) {
    const size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    s->buffers.push_back(p);
    s->device_bytes += static_cast<int64_t>(bytes);
    if (!v.empty()) {
        e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpy");
    }
    *out = static_cast<const T*>(p);
    return MLGPU_OK;
}

bool photons_ok(const mlgpu_photons* p) {
    return p && p->support_x && p->support_y && p->support_z && p->direction_x && p->direction_y &&
           p->direction_z && p->wavelength;
}


int check_results(const mlgpu_results* r) {
    if (!r->final_frame || !r->sensor_id || !r->final_interaction || !r->num_interactions || !r->error ||
        !r->time_of_flight || !r->x || !r->y || !r->cx || !r->cy) {
        set_last_error("null result array");
        return MLGPU_ERR_ARGUMENT;
    }
    return MLGPU_OK;
}

int check_world(const mlgpu_scene* s, int32_t world_frame) {
    if (world_frame < 0 || world_frame >= static_cast<int32_t>(s->scope_of_frame.size()) ||
        s->scope_of_frame[world_frame] < 0) {
        set_last_error("world_frame must be frame 0 or a restriction frame");
        return MLGPU_ERR_ARGUMENT;
    }
    return MLGPU_OK;
}

int check_config(const mlgpu_config* cfg) {
    if (cfg->max_num_interactions_per_photon > 250) {
        set_last_error("max_num_interactions_per_photon must be <= 250");
        return MLGPU_ERR_ARGUMENT;
    }
    return MLGPU_OK;
}

int check_source(const mlgpu_source* src) {
    if (src->support != MLGPU_SUPPORT_POINT && src->support != MLGPU_SUPPORT_DISC) {
        set_last_error("bad support kind");
        return MLGPU_ERR_ARGUMENT;
    }
    if (src->direction < MLGPU_DIRECTION_AXIS || src->direction > MLGPU_DIRECTION_RING) {
        set_last_error("bad direction kind");
        return MLGPU_ERR_ARGUMENT;
    }
    if (src->spectrum != MLGPU_SPECTRUM_INVERSE_SQUARE && src->spectrum != MLGPU_SPECTRUM_UNIFORM) {
        set_last_error("bad spectrum kind");
        return MLGPU_ERR_ARGUMENT;
    }
    const bool bounded = src->direction == MLGPU_DIRECTION_TILT || src->direction == MLGPU_DIRECTION_CAP ||
                         src->direction == MLGPU_DIRECTION_RING;
    if (!(src->angle >= 0.0) || (bounded && !(src->angle <= 1.57079632679489661923)) || !std::isfinite(src->angle)) {
        set_last_error("angle must be finite and >= 0, and <= pi/2 for TILT, CAP and RING");
        return MLGPU_ERR_ARGUMENT;
    }
    if (src->direction == MLGPU_DIRECTION_RING &&
        (!(src->angle_width >= 0.0) || !(src->angle + 8.0 * src->angle_width <= 1.57079632679489661923))) {
        set_last_error("RING: need angle_width >= 0 and angle + 8 angle_width <= pi/2");
        return MLGPU_ERR_ARGUMENT;
    }
    if (!(src->radius >= 0.0) || !std::isfinite(src->radius) || !std::isfinite(src->start_distance)) {
        set_last_error("radius must be finite and >= 0, start_distance finite");
        return MLGPU_ERR_ARGUMENT;
    }
    if (!(src->lambda_min > 0.0) || !(src->lambda_max >= src->lambda_min) || !std::isfinite(src->lambda_max)) {
        set_last_error("need 0 < lambda_min <= lambda_max < inf");
        return MLGPU_ERR_ARGUMENT;
    }
    for (int k = 0; k < 9; k++)
        if (!std::isfinite(src->rot[k])) { set_last_error("pose must be finite"); return MLGPU_ERR_ARGUMENT; }
    for (int k = 0; k < 3; k++)
        if (!std::isfinite(src->pos[k])) { set_last_error("pose must be finite"); return MLGPU_ERR_ARGUMENT; }
    return MLGPU_OK;
}

int check_histogram(const mlgpu_histogram* h) {
    if (!h->counts) { set_last_error("null histogram counts"); return MLGPU_ERR_ARGUMENT; }
    if (h->num_sensors < 1 || h->num_bins < 1 ||
        static_cast<int64_t>(h->num_sensors) * h->num_bins > (int64_t{1} << 40)) {
        set_last_error("need num_sensors >= 1 and num_bins >= 1");
        return MLGPU_ERR_ARGUMENT;
    }
    if (!(h->bin_width > 0.0) || !std::isfinite(h->bin_width) || !std::isfinite(h->t_start) ||
        !(h->jitter_sigma >= 0.0) || !std::isfinite(h->jitter_sigma)) {
        set_last_error("need finite t_start, bin_width > 0 and jitter_sigma >= 0");
        return MLGPU_ERR_ARGUMENT;
    }
    return MLGPU_OK;
}

int schedule_for(const mlgpu_scene* s, int64_t n) {
    const int schedule = s->tuning.schedule;
    if (schedule != MLGPU_SCHEDULE_AUTO) return schedule;
    return n >= 65536 ? MLGPU_SCHEDULE_WAVEFRONT : MLGPU_SCHEDULE_FUSED;
}

// Calls that use the scene's shared device memory (wavefront workspace,
// bunch arena) hold ws_mutex on the host; on the device, a call on another
// stream than the previous user first waits for that user's last kernel.
int begin_shared_use(mlgpu_scene* ms, hipStream_t st) {
    hipError_t e;
    if (!ms->ws_event) {
        e = hipEventCreateWithFlags(&ms->ws_event, hipEventDisableTiming);
        if (e != hipSuccess) { ms->ws_event = nullptr; return hip_fail(e, "hipEventCreate"); }
        ms->ws_stream = st;
    } else if (ms->ws_stream != st) {
        e = hipStreamWaitEvent(st, ms->ws_event, 0);
        if (e != hipSuccess) return hip_fail(e, "hipStreamWaitEvent");
    }
    return MLGPU_OK;
}

void end_shared_use(mlgpu_scene* ms, hipStream_t st) {
    (void)hipEventRecord(ms->ws_event, st);
    ms->ws_stream = st;
}

// (Re)allocate a shared buffer once no earlier call can still use it.
int grow(mlgpu_scene* ms, void** buf, size_t bytes) {
    if (*buf) {
        (void)hipEventSynchronize(ms->ws_event);
        (void)hipFree(*buf);
        *buf = nullptr;
    }
    const hipError_t e = hipMalloc(buf, bytes);
    if (e != hipSuccess) { *buf = nullptr; return hip_fail(e, "hipMalloc"); }
    return MLGPU_OK;
}

// Hands out 256-byte aligned pieces of a buffer.
struct Carver {
    char* p;
    template <typename T>
    T* take(size_t count) {
        T* q = reinterpret_cast<T*>(p);
        p += (count * sizeof(T) + 255) & ~static_cast<size_t>(255);
        return q;
    }
};

int launch_fused(const mlgpu_scene* s, int32_t scope, const mlgpu_photons& ph, int64_t n, const mlgpu_config& cfg,
                 const mlgpu_results& r, hipStream_t st, const int32_t* ids = nullptr) {
    const hipError_t e = mlgpu::launch_propagate(s->dev, scope, ph, n, cfg.max_num_interactions_per_photon, cfg.seed,
                                                 cfg.photon_id_offset, r, st, ids);
    if (e != hipSuccess) return hip_fail(e, "propagate launch");
    return MLGPU_OK;
}

// The wavefront workspace of cap photons laid out in mem (workspace_bytes(cap)).
mlgpu::Workspace carve_workspace(void* mem, int64_t capacity) {
    const size_t cap = static_cast<size_t>(capacity);
    Carver c{static_cast<char*>(mem)};
    mlgpu::Workspace w;
    w.cap = capacity;
    w.ray = c.take<mlgpu::PhotonRay>(cap);
    w.sums = c.take<mlgpu::PhotonSums>(cap);
    w.hit_leaf = c.take<int32_t>(cap);
    w.hit_t = c.take<double>(cap);
    w.tr_cur = c.take<int32_t>(cap);
    w.tr_depth = c.take<int32_t>(cap);
    w.tr_stack = c.take<int32_t>(cap * mlgpu::kMaxTreeDepth);
    w.deferred = c.take<int32_t>(cap);
    w.queue[0] = c.take<int32_t>(cap);
    w.queue[1] = c.take<int32_t>(cap);
    w.counts = c.take<int32_t>(mlgpu::kNumCounts);
    return w;
}

// The schedules that use the scene's workspace (WAVEFRONT: photon state
// and queues; PERSISTENT: only its counters) on workspace w, whose capacity
// is at least n photons for WAVEFRONT.
hipError_t launch_on_workspace(const mlgpu_scene* s, int schedule, int32_t scope, const mlgpu_photons& ph,
                               int64_t n, uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                               const mlgpu_results& r, const mlgpu::Workspace& w, hipStream_t st,
                               const mlgpu::CallParams* cp, const int32_t* ids) {
    const mlgpu_tuning& t = s->tuning;
    if (schedule == MLGPU_SCHEDULE_PERSISTENT)
        return mlgpu::launch_propagate_persistent(s->dev, scope, ph, n, max_interactions, seed, id_offset, r, w.counts,
                                                  t.blocks_per_cu * s->num_cus, t.trace_budget, t.shade_lanes, st, cp,
                                                  ids);
    return mlgpu::launch_propagate_wavefront(s->dev, scope, ph, n, max_interactions, seed, id_offset, r, w,
                                             t.blocks_per_cu * s->num_cus, t.trace_budget, t.tail_level, st, cp, ids);
}

// WAVEFRONT or PERSISTENT propagation; the caller holds ws_mutex and began
// shared use.
int launch_wavefront(mlgpu_scene* ms, int32_t scope, const mlgpu_photons& ph, int64_t n, const mlgpu_config& cfg,
                     const mlgpu_results& r, hipStream_t st, const int32_t* ids = nullptr) {
    const int schedule = schedule_for(ms, n);
    const int64_t need = schedule == MLGPU_SCHEDULE_PERSISTENT ? 1 : n;
    if (ms->ws_capacity < need) {
        const int64_t cap = need + need / 8;
        const int rc = grow(ms, &ms->ws, mlgpu::workspace_bytes(cap));
        ms->ws_capacity = rc ? 0 : cap;
        if (rc) return rc;
    }
    const mlgpu::Workspace w = carve_workspace(ms->ws, ms->ws_capacity);
    const hipError_t e = launch_on_workspace(ms, schedule, scope, ph, n, cfg.max_num_interactions_per_photon,
                                             cfg.seed, cfg.photon_id_offset, r, w, st, nullptr, ids);
    if (e != hipSuccess) return hip_fail(e, "propagate launch");
    return MLGPU_OK;
}

size_t arena_bytes(int64_t photons, int32_t bunches) {
    const size_t m = static_cast<size_t>(photons);
    // 7 photon arrays, the result arrays, the bunch table, alignment slack
    return m * (7 * 8 + 4 + 4 + 1 + 1 + 1 + 8 + 4 * 4) + sizeof(mlgpu::BunchDev) * static_cast<size_t>(bunches) +
           20 * 256;
}

// Draw, propagate and bin bunches; the caller holds ws_mutex and began shared use.
int run_bunches(mlgpu_scene* ms, int32_t scope, const mlgpu_bunch* bunches, int32_t nb, int64_t total,
                const mlgpu_config& cfg, const mlgpu_histogram* hist, const mlgpu_results* results, hipStream_t st) {
    int rc;
    hipError_t e;
    if (ms->arena_photons < total || ms->arena_bunches < nb) {
        const int64_t cap = std::max(total + total / 8, ms->arena_photons);
        const int32_t cb = std::max(nb, ms->arena_bunches);
        rc = grow(ms, &ms->arena, arena_bytes(cap, cb));
        ms->arena_photons = rc ? 0 : cap;
        ms->arena_bunches = rc ? 0 : cb;
        if (rc) return rc;
    }
    // bunch table through the pinned staging buffer
    if (!ms->staging_event) {
        e = hipEventCreateWithFlags(&ms->staging_event, hipEventDisableTiming);
        if (e != hipSuccess) { ms->staging_event = nullptr; return hip_fail(e, "hipEventCreate"); }
    } else {
        (void)hipEventSynchronize(ms->staging_event);
    }
    if (ms->staging_bunches < nb) {
        if (ms->staging) (void)hipHostFree(ms->staging);
        ms->staging = nullptr;
        ms->staging_bunches = 0;
        e = hipHostMalloc(reinterpret_cast<void**>(&ms->staging), sizeof(mlgpu::BunchDev) * static_cast<size_t>(nb));
        if (e != hipSuccess) { ms->staging = nullptr; return hip_fail(e, "hipHostMalloc"); }
        ms->staging_bunches = nb;
    }
    int64_t start = 0;
    for (int32_t b = 0; b < nb; b++) {
        ms->staging[b].source = bunches[b].source;
        ms->staging[b].seed = bunches[b].seed;
        ms->staging[b].start = start;
        start += bunches[b].num_photons;
    }
    const size_t cap = static_cast<size_t>(ms->arena_photons);
    Carver c{static_cast<char*>(ms->arena)};
    auto* table = c.take<mlgpu::BunchDev>(static_cast<size_t>(ms->arena_bunches));
    mlgpu_photon_buffers gen{c.take<double>(cap), c.take<double>(cap), c.take<double>(cap), c.take<double>(cap),
                             c.take<double>(cap), c.take<double>(cap), c.take<double>(cap)};
    mlgpu_results own{c.take<int32_t>(cap), c.take<int32_t>(cap), c.take<uint8_t>(cap), c.take<uint8_t>(cap),
                      c.take<uint8_t>(cap), c.take<double>(cap), c.take<float>(cap), c.take<float>(cap),
                      c.take<float>(cap), c.take<float>(cap), nullptr};
    e = hipMemcpyAsync(table, ms->staging, sizeof(mlgpu::BunchDev) * static_cast<size_t>(nb), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(ms->staging_event, st);
    if (e == hipSuccess) e = mlgpu::launch_generate_bunches(table, nb, total, gen, st);
    if (e != hipSuccess) return hip_fail(e, "bunch upload or generation");
    const mlgpu_photons ph{gen.support_x, gen.support_y, gen.support_z, gen.direction_x, gen.direction_y,
                           gen.direction_z, gen.wavelength};
    const mlgpu_results& r = results ? *results : own;
    if (schedule_for(ms, total) == MLGPU_SCHEDULE_FUSED)
        rc = launch_fused(ms, scope, ph, total, cfg, r, st);
    else
        rc = launch_wavefront(ms, scope, ph, total, cfg, r, st);
    if (rc) return rc;
    if (hist) {
        e = mlgpu::launch_histogram(r, total, *hist, ms->num_cus, st);
        if (e != hipSuccess) return hip_fail(e, "histogram launch");
    }
    return MLGPU_OK;
}

}  // namespace

namespace mlgpu {

int propagate_gathered(const mlgpu_scene* s, int32_t world_frame, const mlgpu_photons* photons, int64_t n,
                       const mlgpu_config* cfg, mlgpu_results* r, const int32_t* ids, void* stream) {
    if (!s || !cfg || !r || n < 0) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    if (n == 0) return MLGPU_OK;
    if (!photons_ok(photons)) { set_last_error("null photon array"); return MLGPU_ERR_ARGUMENT; }
    int rc;
    if ((rc = check_results(r)) || (rc = check_world(s, world_frame)) || (rc = check_config(cfg))) return rc;
    if (n > INT32_MAX) { set_last_error("at most 2^31 - 1 photons per call"); return MLGPU_ERR_ARGUMENT; }
    auto* ms = const_cast<mlgpu_scene*>(s);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (schedule_for(s, n) == MLGPU_SCHEDULE_FUSED)
        return launch_fused(s, s->scope_of_frame[world_frame], *photons, n, *cfg, *r, st, ids);
    std::lock_guard<std::mutex> lock(ms->ws_mutex);
    if ((rc = begin_shared_use(ms, st))) return rc;
    rc = launch_wavefront(ms, s->scope_of_frame[world_frame], *photons, n, *cfg, *r, st, ids);
    end_shared_use(ms, st);
    return rc;
}

int histogram_gathered(const mlgpu_results* r, int64_t n, const mlgpu_histogram* h, const int32_t* ids,
                       void* stream) {
    if (!r || !h || n < 0) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    if (n == 0) return MLGPU_OK;
    if (!r->sensor_id || !r->error || !r->time_of_flight) {
        set_last_error("null array");
        return MLGPU_ERR_ARGUMENT;
    }
    const int rc = check_histogram(h);
    if (rc) return rc;
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    const hipError_t e = mlgpu::launch_histogram(*r, n, *h, cus, static_cast<hipStream_t>(stream), nullptr, ids);
    if (e != hipSuccess) return hip_fail(e, "histogram launch");
    return MLGPU_OK;
}

}  // namespace mlgpu

extern "C" {

int mlgpu_scene_create(const mlgpu_scene_desc* desc, mlgpu_scene** out) {
    if (!desc || !out) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
        set_last_error("no HIP device");
        return MLGPU_ERR_NO_DEVICE;
    }
    mlgpu::HostScene hs;
    try {
        mlgpu::build_host_scene(*desc, &hs);
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return MLGPU_ERR_SCENE;
    }
    auto* s = new mlgpu_scene();
    (void)hipGetDevice(&s->device);
    if (hipDeviceGetAttribute(&s->num_cus, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess ||
        s->num_cus < 1)
        s->num_cus = 256;
    int rc = MLGPU_OK;
    if ((rc = upload(s, hs.pairs, &s->dev.pairs)) || (rc = upload(s, hs.scopes, &s->dev.scopes)) ||
        (rc = upload(s, hs.geo, &s->dev.geo)) || (rc = upload(s, hs.boxes, &s->dev.boxes)) ||
        (rc = upload(s, hs.phys, &s->dev.phys)) || (rc = upload(s, hs.funcs, &s->dev.funcs)) ||
        (rc = upload(s, hs.points, &s->dev.points)) || (rc = upload(s, hs.buckets, &s->dev.buckets))) {
        mlgpu_scene_destroy(s);
        return rc;
    }
    s->dev.num_pairs = static_cast<int32_t>(hs.pairs.size());
    s->dev.num_leaves = static_cast<int32_t>(hs.geo.size());
    s->dev.root_scope = hs.root_scope;
    s->dev.default_refraction = hs.default_refraction;
    s->dev.eps = static_cast<float>(hs.eps);
    s->dev.hex_sin = hs.hex_sin;
    s->dev.hex_cos = hs.hex_cos;
    s->dev.num_funcs = static_cast<int32_t>(hs.funcs.size());
    s->dev.num_points = static_cast<int32_t>(hs.points.size());
    s->dev.num_buckets = static_cast<int32_t>(hs.buckets.size());
    s->dev.hot_count = hs.hot_count;
    s->dev.num_hot = 0;
    s->scope_of_frame = hs.scope_of_frame;
    s->num_nodes = static_cast<int64_t>(hs.pairs.size());
    s->num_leaves = static_cast<int64_t>(hs.geo.size());
    *out = s;
    return MLGPU_OK;
}

void mlgpu_scene_destroy(mlgpu_scene* s) {
    if (!s) return;
    if (s->ws_event) {
        (void)hipEventSynchronize(s->ws_event);
        (void)hipEventDestroy(s->ws_event);
    }
    if (s->staging_event) {
        (void)hipEventSynchronize(s->staging_event);
        (void)hipEventDestroy(s->staging_event);
    }
    if (s->staging) (void)hipHostFree(s->staging);
    if (s->arena) (void)hipFree(s->arena);
    for (void* p : s->buffers) (void)hipFree(p);
    if (s->ws) (void)hipFree(s->ws);
    delete s;
}

int mlgpu_scene_set_tuning(mlgpu_scene* s, const mlgpu_tuning* t) {
    if (!s || !t) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    if (t->schedule < MLGPU_SCHEDULE_AUTO || t->schedule > MLGPU_SCHEDULE_PERSISTENT) {
        set_last_error("bad schedule");
        return MLGPU_ERR_ARGUMENT;
    }
    if (t->blocks_per_cu < 1 || t->blocks_per_cu > 64) {
        set_last_error("blocks_per_cu must be in [1, 64]");
        return MLGPU_ERR_ARGUMENT;
    }
    if (t->shade_lanes < 1 || t->shade_lanes > 64) {
        set_last_error("shade_lanes must be in [1, 64]");
        return MLGPU_ERR_ARGUMENT;
    }
    s->tuning = *t;
    return MLGPU_OK;
}

int mlgpu_scene_get_tuning(const mlgpu_scene* s, mlgpu_tuning* t) {
    if (!s || !t) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    *t = s->tuning;
    return MLGPU_OK;
}

int mlgpu_scene_autotune(mlgpu_scene* s, int32_t world_frame, const mlgpu_photons* photons, int64_t n,
                         const mlgpu_config* cfg, int32_t repeats, mlgpu_tuning* best, double* best_ms,
                         void* stream) {
    if (!s || !cfg || n <= 0 || repeats < 1) { set_last_error("null argument, no photons or repeats < 1"); return MLGPU_ERR_ARGUMENT; }
    if (!photons_ok(photons)) { set_last_error("null photon array"); return MLGPU_ERR_ARGUMENT; }
    int rc;
    if ((rc = check_world(s, world_frame)) || (rc = check_config(cfg))) return rc;
    if (n > INT32_MAX) { set_last_error("at most 2^31 - 1 photons per call"); return MLGPU_ERR_ARGUMENT; }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    // temporary results
    const size_t cap = static_cast<size_t>(n);
    const size_t bytes = cap * (4 + 4 + 1 + 1 + 1 + 8 + 4 * 4) + 12 * 256;
    void* buf = nullptr;
    hipError_t e = hipMalloc(&buf, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc autotune results");
    Carver c{static_cast<char*>(buf)};
    mlgpu_results r{c.take<int32_t>(cap), c.take<int32_t>(cap), c.take<uint8_t>(cap), c.take<uint8_t>(cap),
                    c.take<uint8_t>(cap), c.take<double>(cap), c.take<float>(cap), c.take<float>(cap),
                    c.take<float>(cap), c.take<float>(cap), nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    std::vector<mlgpu_tuning> cands;
    cands.push_back(mlgpu_tuning{MLGPU_SCHEDULE_FUSED, s->tuning.trace_budget, s->tuning.blocks_per_cu,
                                  s->tuning.tail_level, s->tuning.shade_lanes});
    for (int32_t budget : {0, 12, 16, 24, 32, 48})
        for (int32_t bpc : {8, 16, 24, 32, 48})
            for (int32_t tail : {0, 1, 2})
                cands.push_back(mlgpu_tuning{MLGPU_SCHEDULE_WAVEFRONT, budget, bpc, tail, s->tuning.shade_lanes});
    for (int32_t quantum : {0, 8, 16, 32})
        for (int32_t bpc : {4, 8, 16})
            for (int32_t lanes : {1, 16, 32, 48})
                cands.push_back(mlgpu_tuning{MLGPU_SCHEDULE_PERSISTENT, quantum, bpc, s->tuning.tail_level, lanes});
    const mlgpu_tuning saved = s->tuning;
    mlgpu_tuning chosen = saved;
    double chosen_ms = -1.0;
    for (const mlgpu_tuning& t : cands) {
        if (e != hipSuccess) break;
        s->tuning = t;
        rc = mlgpu_propagate(s, world_frame, photons, n, cfg, &r, stream);  // warm-up
        if (rc) break;
        e = hipEventRecord(ev[0], st);
        for (int32_t k = 0; k < repeats && rc == MLGPU_OK && e == hipSuccess; k++)
            rc = mlgpu_propagate(s, world_frame, photons, n, cfg, &r, stream);
        if (rc) break;
        if (e == hipSuccess) e = hipEventRecord(ev[1], st);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
        if (e != hipSuccess) break;
        const double per_call = static_cast<double>(ms) / repeats;
        if (chosen_ms < 0.0 || per_call < chosen_ms) {
            chosen_ms = per_call;
            chosen = t;
        }
    }
    (void)hipStreamSynchronize(st);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    (void)hipFree(buf);
    if (rc || e != hipSuccess) {
        s->tuning = saved;
        return rc ? rc : hip_fail(e, "autotune timing");
    }
    s->tuning = chosen;
    if (best) *best = chosen;
    if (best_ms) *best_ms = chosen_ms;
    return MLGPU_OK;
}

int mlgpu_desc_frame_poses(const mlgpu_scene_desc* d, double* out) {
    if (!d || !out) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    try {
        mlgpu::frame_poses(*d, out);
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return MLGPU_ERR_SCENE;
    }
    return MLGPU_OK;
}

int mlgpu_scene_stats(const mlgpu_scene* s, int64_t* num_nodes, int64_t* num_leaves, int64_t* device_bytes) {
    if (!s) { set_last_error("null scene"); return MLGPU_ERR_ARGUMENT; }
    if (num_nodes) *num_nodes = s->num_nodes;
    if (num_leaves) *num_leaves = s->num_leaves;
    if (device_bytes) *device_bytes = s->device_bytes;
    return MLGPU_OK;
}

int mlgpu_propagate(const mlgpu_scene* s, int32_t world_frame, const mlgpu_photons* photons, int64_t n,
                    const mlgpu_config* cfg, mlgpu_results* r, void* stream) {
    return mlgpu::propagate_gathered(s, world_frame, photons, n, cfg, r, nullptr, stream);
}

int mlgpu_propagate_noalloc(const mlgpu_scene* s, int32_t world_frame, const mlgpu_photons* photons, int64_t n,
                            const mlgpu_config* cfg, mlgpu_results* r, void* stream) {
    if (!s || !cfg || !r || n < 0) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    if (n == 0) return MLGPU_OK;
    if (!photons_ok(photons)) { set_last_error("null photon array"); return MLGPU_ERR_ARGUMENT; }
    int rc;
    if ((rc = check_results(r)) || (rc = check_world(s, world_frame)) || (rc = check_config(cfg))) return rc;
    if (n > mlgpu::kNoAllocMaxPhotons) {
        set_last_error("mlgpu_propagate_noalloc: at most MLGPU_NOALLOC_MAX_PHOTONS (262144) photons per call");
        return MLGPU_ERR_ARGUMENT;
    }
    void* mem = nullptr;
    hipError_t e = mlgpu::noalloc_workspace(&mem);
    if (e != hipSuccess) return hip_fail(e, "no-alloc workspace address");
    // the wavefront schedule on the library's static workspace; the kernels
    // clear their counters themselves, so nothing needs zeroing beforehand
    const mlgpu::Workspace w = carve_workspace(mem, mlgpu::kNoAllocMaxPhotons);
    const int schedule = s->tuning.schedule == MLGPU_SCHEDULE_PERSISTENT ? MLGPU_SCHEDULE_PERSISTENT
                                                                          : MLGPU_SCHEDULE_WAVEFRONT;
    e = launch_on_workspace(s, schedule, s->scope_of_frame[world_frame], *photons, n,
                            cfg->max_num_interactions_per_photon, cfg->seed, cfg->photon_id_offset, *r, w,
                            static_cast<hipStream_t>(stream), nullptr, nullptr);
    if (e != hipSuccess) return hip_fail(e, "propagate launch");
    return MLGPU_OK;
}

int mlgpu_propagate_history(const mlgpu_scene* s, int32_t world_frame, const mlgpu_photons* photons, int64_t n,
                            const mlgpu_config* cfg, mlgpu_results* r, const mlgpu_history* h, void* stream) {
    if (!s || !cfg || !r || !h || n < 0) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    if (n == 0) return MLGPU_OK;
    if (!photons_ok(photons)) { set_last_error("null photon array"); return MLGPU_ERR_ARGUMENT; }
    int rc;
    if ((rc = check_results(r)) || (rc = check_world(s, world_frame)) || (rc = check_config(cfg))) return rc;
    if (!h->bounces || !h->num_bounces) { set_last_error("null history array"); return MLGPU_ERR_ARGUMENT; }
    if (h->capacity < 1 || static_cast<uint32_t>(h->capacity) < cfg->max_num_interactions_per_photon) {
        set_last_error("history capacity below max_num_interactions_per_photon");
        return MLGPU_ERR_ARGUMENT;
    }
    if (n > INT32_MAX) { set_last_error("at most 2^31 - 1 photons per call"); return MLGPU_ERR_ARGUMENT; }
    const hipError_t e = mlgpu::launch_propagate_history(
        s->dev, s->scope_of_frame[world_frame], *photons, n, cfg->max_num_interactions_per_photon, cfg->seed,
        cfg->photon_id_offset, *r, *h, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "propagate launch");
    return MLGPU_OK;
}

int mlgpu_propagate_bunches(const mlgpu_scene* s, int32_t world_frame, const mlgpu_bunch* bunches,
                            int32_t num_bunches, const mlgpu_config* cfg, const mlgpu_histogram* hist,
                            mlgpu_results* results, void* stream) {
    if (!s || !cfg || !bunches || num_bunches < 1) { set_last_error("null argument or no bunch"); return MLGPU_ERR_ARGUMENT; }
    int rc;
    if ((rc = check_world(s, world_frame)) || (rc = check_config(cfg))) return rc;
    if (results && (rc = check_results(results))) return rc;
    if (hist && (rc = check_histogram(hist))) return rc;
    int64_t total = 0;
    for (int32_t b = 0; b < num_bunches; b++) {
        if (bunches[b].num_photons < 0) { set_last_error("negative bunch size"); return MLGPU_ERR_ARGUMENT; }
        if ((rc = check_source(&bunches[b].source))) return rc;
        total += bunches[b].num_photons;
        if (total > INT32_MAX) { set_last_error("at most 2^31 - 1 photons per call"); return MLGPU_ERR_ARGUMENT; }
    }
    if (total == 0) return MLGPU_OK;
    auto* ms = const_cast<mlgpu_scene*>(s);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lock(ms->ws_mutex);
    if ((rc = begin_shared_use(ms, st))) return rc;
    rc = run_bunches(ms, s->scope_of_frame[world_frame], bunches, num_bunches, total, *cfg, hist, results, st);
    end_shared_use(ms, st);
    return rc;
}

int mlgpu_generate_bunches(const mlgpu_bunch* bunches, int32_t nb, const mlgpu_photon_buffers* out, int64_t capacity,
                           void* stream) {
    if (!bunches || !out || nb < 0 || capacity < 0) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    int64_t total = 0;
    std::vector<mlgpu::BunchDev> table(static_cast<size_t>(nb));
    for (int32_t b = 0; b < nb; b++) {
        if (bunches[b].num_photons < 0) { set_last_error("negative bunch size"); return MLGPU_ERR_ARGUMENT; }
        const int rc = check_source(&bunches[b].source);
        if (rc) return rc;
        table[static_cast<size_t>(b)] = mlgpu::BunchDev{bunches[b].source, bunches[b].seed, total};
        total += bunches[b].num_photons;
    }
    if (total > capacity) { set_last_error("capacity below the total of the bunches"); return MLGPU_ERR_ARGUMENT; }
    if (total == 0) return MLGPU_OK;
    if (!out->support_x || !out->support_y || !out->support_z || !out->direction_x || !out->direction_y ||
        !out->direction_z || !out->wavelength) {
        set_last_error("null photon array");
        return MLGPU_ERR_ARGUMENT;
    }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    mlgpu::BunchDev* dev = nullptr;
    const size_t bytes = sizeof(mlgpu::BunchDev) * table.size();
    hipError_t e = hipMalloc(&dev, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc bunch table");
    e = hipMemcpyAsync(dev, table.data(), bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = mlgpu::launch_generate_bunches(dev, nb, total, *out, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_fail(e, "bunch generation");
    return MLGPU_OK;
}

int mlgpu_generate_photons(const mlgpu_source* src, int64_t n, uint64_t seed, uint64_t photon_id_offset,
                           const mlgpu_photon_buffers* out, void* stream) {
    if (!src || !out || n < 0) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    if (n == 0) return MLGPU_OK;
    if (!out->support_x || !out->support_y || !out->support_z || !out->direction_x || !out->direction_y ||
        !out->direction_z || !out->wavelength) {
        set_last_error("null photon array");
        return MLGPU_ERR_ARGUMENT;
    }
    const int rc = check_source(src);
    if (rc) return rc;
    const hipError_t e = mlgpu::launch_generate(*src, n, seed, photon_id_offset, *out, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "generate launch");
    return MLGPU_OK;
}

int mlgpu_render_pinhole(const mlgpu_scene* s, int32_t world_frame, const mlgpu_pinhole_camera* cam,
                         const mlgpu_render_config* cfg, const float* frame_colors, int32_t n_frames, uint64_t seed,
                         float* image, int64_t* num_errors, void* stream) {
    if (!s || !cam || !cfg || !image || !num_errors) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    *num_errors = 0;
    if (world_frame < 0 || world_frame >= static_cast<int32_t>(s->scope_of_frame.size()) ||
        s->scope_of_frame[world_frame] < 0) {
        set_last_error("world_frame must be frame 0 or a restriction frame");
        return MLGPU_ERR_ARGUMENT;
    }
    if (n_frames != static_cast<int32_t>(s->scope_of_frame.size())) {
        set_last_error("n_frames must be the description's number of frames");
        return MLGPU_ERR_ARGUMENT;
    }
    // CameraDevice::assert_field_of_view_is_valid (CameraDevice.cpp:81-92)
    if (!(cam->field_of_view > 0.0) || !(cam->field_of_view <= M_PI)) {
        set_last_error("field_of_view must be in (0, pi]");
        return MLGPU_ERR_ARGUMENT;
    }
    if (cam->num_cols < 1 || cam->num_rows < 1 ||
        static_cast<int64_t>(cam->num_cols) * cam->num_rows > (int64_t{1} << 31)) {
        set_last_error("need 1 <= num_cols, num_rows and at most 2^31 pixels");
        return MLGPU_ERR_ARGUMENT;
    }
    if (cfg->max_interaction_depth < 1) { set_last_error("max_interaction_depth must be >= 1"); return MLGPU_ERR_ARGUMENT; }
    // CameraDevice / PinHoleCamera::set_pos_rot_fov (CameraDevice.cpp:19-33,
    // PinHoleCamera.cpp:10-50)
    const double rot[4] = {cam->rotation[0], cam->rotation[1], cam->rotation[2], 0.0};
    double T[12];
    mlgpu::pose_matrix(cam->position, MLGPU_ROT_XYZ, rot, T);
    auto orientation = [&](double x, double y, double z, double* o) {
        o[0] = x * T[0] + y * T[1] + z * T[2];
        o[1] = x * T[4] + y * T[5] + z * T[6];
        o[2] = x * T[8] + y * T[9] + z * T[10];
    };
    mlgpu::RenderCamera rc{};
    double axis[3];
    orientation(0.0, 0.0, 1.0, axis);
    const double an = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    for (int k = 0; k < 3; k++) axis[k] = axis[k] / an;  // Ray normalises the optical axis
    orientation(0.0, 1.0, 0.0, rc.scol);
    orientation(1.0, 0.0, 0.0, rc.srow);
    const double dist = (static_cast<double>(static_cast<uint32_t>(cam->num_cols)) / 2.0) /
                        std::tan(cam->field_of_view / 2.0);
    for (int k = 0; k < 3; k++) {
        rc.pos[k] = cam->position[k];
        rc.pp[k] = axis[k] * dist;
    }
    rc.cols = cam->num_cols;
    rc.rows = cam->num_rows;
    mlgpu::RenderParams prm{};
    prm.gi = cfg->global_illumination ? 1 : 0;
    prm.max_depth = cfg->max_interaction_depth;
    for (int k = 0; k < 3; k++) {
        prm.light[k] = cfg->incoming_direction[k];
        prm.sky[k] = cfg->sky_color[k];
    }
    // frame colours: merlict's default COLOR_DARK_GRAY (SurfaceEntity.cpp:42-43)
    std::vector<float> colors(static_cast<size_t>(n_frames) * 6, 25.0f);
    if (frame_colors) std::copy(frame_colors, frame_colors + colors.size(), colors.begin());
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t cbytes = (colors.size() * sizeof(float) + 255) & ~static_cast<size_t>(255);
    mlgpu::PoolBlock block;
    hipError_t e = block.get(cbytes + 256);
    if (e != hipSuccess) return hip_fail(e, "render workspace");
    char* mem = block.as();
    float* dcolors = reinterpret_cast<float*>(mem);
    auto* derr = reinterpret_cast<unsigned long long*>(mem + cbytes);
    unsigned long long herr = 0;
    e = hipMemcpyAsync(dcolors, colors.data(), colors.size() * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(derr, 0, sizeof(unsigned long long), st);
    if (e == hipSuccess)
        e = mlgpu::launch_render(s->dev, s->scope_of_frame[world_frame], rc, prm, dcolors, seed, image, derr, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&herr, derr, sizeof(herr), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipStreamSynchronize(st);  // the borrowed block is idle when returned
    block.release();
    if (e != hipSuccess) return hip_fail(e, "render");
    *num_errors = static_cast<int64_t>(herr);
    return MLGPU_OK;
}

int mlgpu_calibration_photons(const mlgpu_source* src, uint64_t seed, uint64_t photon_id_offset,
                              const mlgpu_results* results, int64_t n, int32_t num_lixel,
                              mlgpu_calibration_photon* out, void* stream) {
    if (!src || !results || !out || n < 0 || num_lixel < 0) {
        set_last_error("null argument or negative size");
        return MLGPU_ERR_ARGUMENT;
    }
    if (n == 0) return MLGPU_OK;
    if (!results->sensor_id || !results->error || !results->time_of_flight) {
        set_last_error("null result array");
        return MLGPU_ERR_ARGUMENT;
    }
    const int rc = check_source(src);
    if (rc) return rc;
    const hipError_t e = mlgpu::launch_calibration_photons(*src, seed, photon_id_offset, *results, n, num_lixel, out,
                                                           static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "calibration photons launch");
    return MLGPU_OK;
}

int mlgpu_corsika_photons(const float* bunches, int64_t n, uint64_t seed, uint64_t id_offset,
                          const mlgpu_photon_buffers* out, int32_t* truth_id, int64_t capacity, int64_t* num_photons,
                          void* stream) {
    if (!out || !num_photons || n < 0 || capacity < 0) {
        set_last_error("null argument or negative size");
        return MLGPU_ERR_ARGUMENT;
    }
    *num_photons = 0;
    if (n == 0) return MLGPU_OK;
    if (!bunches) {
        set_last_error("null bunches");
        return MLGPU_ERR_ARGUMENT;
    }
    if ((reinterpret_cast<uintptr_t>(bunches) & 15u) != 0) {
        set_last_error("bunches must be 16-byte aligned");
        return MLGPU_ERR_ARGUMENT;
    }
    if (capacity > 0 && (!out->support_x || !out->support_y || !out->support_z || !out->direction_x ||
                         !out->direction_y || !out->direction_z || !out->wavelength)) {
        set_last_error("null photon array");
        return MLGPU_ERR_ARGUMENT;
    }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t ws_bytes = mlgpu::corsika_workspace_bytes(n);
    mlgpu::PoolBlock block;
    hipError_t e = block.get(ws_bytes + 256);
    if (e != hipSuccess) return hip_fail(e, "CORSIKA workspace");
    void* ws = block.as<void>();
    int64_t host[2] = {0, 0};  // {bad, total}
    char* flags = static_cast<char*>(ws) + ws_bytes;  // int32 bad at +0, int64 total at +8
    int32_t* bad = reinterpret_cast<int32_t*>(flags);
    int64_t* total = reinterpret_cast<int64_t*>(flags + 8);
    int rc = MLGPU_OK;
    e = hipMemsetAsync(flags, 0, 16, st);
    if (e == hipSuccess) e = mlgpu::launch_corsika_count(bunches, n, seed, id_offset, ws, bad, total, st);
    if (e == hipSuccess) e = hipMemcpyAsync(host, flags, 16, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        rc = hip_fail(e, "CORSIKA count");
    } else if (static_cast<int32_t>(host[0]) != 0) {
        set_last_error("CORSIKA bunch weight outside [0, 16]");  // PhotonFactory.cpp:138-148
        rc = MLGPU_ERR_ARGUMENT;
    } else {
        *num_photons = host[1];
        if (host[1] > capacity) {
            set_last_error("capacity below the number of photons of the bunches (*num_photons)");
            rc = MLGPU_ERR_ARGUMENT;
        } else {
            e = mlgpu::launch_corsika_emit(bunches, n, seed, id_offset, ws, capacity, *out, truth_id, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) rc = hip_fail(e, "CORSIKA emit");
        }
    }
    (void)hipStreamSynchronize(st);  // the borrowed block is idle when returned
    return rc;
}

int mlgpu_histogram_arrivals(const mlgpu_results* r, int64_t n, const mlgpu_histogram* h, void* stream) {
    return mlgpu::histogram_gathered(r, n, h, nullptr, stream);
}

int mlgpu_count_errors(const mlgpu_results* r, int64_t n, void* stream, int64_t* num_errors) {
    if (!r || !r->error || !num_errors || n < 0) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    *num_errors = 0;
    if (n == 0) return MLGPU_OK;
    unsigned long long* counter = nullptr;
    hipError_t e = hipMalloc(&counter, sizeof(unsigned long long));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    e = hipMemsetAsync(counter, 0, sizeof(unsigned long long), st);
    if (e == hipSuccess) e = mlgpu::launch_count_errors(r->error, n, counter, st);
    unsigned long long host = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&host, counter, sizeof(host), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(counter);
    if (e != hipSuccess) return hip_fail(e, "count errors");
    *num_errors = static_cast<int64_t>(host);
    return MLGPU_OK;
}

int mlgpu_first_hit(const mlgpu_scene* s, int32_t root_frame, const mlgpu_photons* rays, int64_t n,
                    mlgpu_hits* hits, void* stream) {
    if (!s || !hits || n < 0) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    if (n == 0) return MLGPU_OK;
    if (!photons_ok(rays) || !hits->frame || !hits->distance || !hits->position_x || !hits->position_y ||
        !hits->position_z) {
        set_last_error("null array");
        return MLGPU_ERR_ARGUMENT;
    }
    if (root_frame < 0 || root_frame >= static_cast<int32_t>(s->scope_of_frame.size()) ||
        s->scope_of_frame[root_frame] < 0) {
        set_last_error("root_frame must be frame 0 or a restriction frame");
        return MLGPU_ERR_ARGUMENT;
    }
    hipError_t e = mlgpu::launch_first_hit(s->dev, s->scope_of_frame[root_frame], *rays, n, *hits,
                                           static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "first_hit launch");
    return MLGPU_OK;
}

}  // extern "C"

// ------------------------------------------------------------- bunch plans
// A captured launch sequence for repeated batches of bunches: bunch table
// and call parameters up from pinned staging, photon draw, wavefront
// propagation, histogram.  The kernels read the batch's photon count, seed
// and id offset from the parameter block, so one graph serves every batch
// up to the plan's capacity.
struct mlgpu_bunch_plan {
    const mlgpu_scene* scene = nullptr;
    int32_t max_bunches = 0;
    int64_t max_photons = 0;
    uint32_t max_interactions = 0;
    void* mem = nullptr;       // device: parameters + table, photons, results, workspace
    char* staging = nullptr;   // pinned: parameters + table
    size_t head_bytes = 0;
    mlgpu_results results{};
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    bool launched = false;
    std::mutex mutex;
};

namespace {

size_t round256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

void plan_free(mlgpu_bunch_plan* p) {
    if (p->done) {
        if (p->launched) (void)hipEventSynchronize(p->done);
        (void)hipEventDestroy(p->done);
    }
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->capture) (void)hipStreamDestroy(p->capture);
    if (p->staging) (void)hipHostFree(p->staging);
    if (p->mem) (void)hipFree(p->mem);
    delete p;
}

}  // namespace

int mlgpu_bunch_plan_create(const mlgpu_scene* s, int32_t world_frame, int32_t max_bunches, int64_t max_photons,
                            const mlgpu_config* cfg, const mlgpu_histogram* hist, mlgpu_bunch_plan** out) {
    if (!s || !cfg || !out || max_bunches < 1 || max_photons < 1 || max_photons > INT32_MAX) {
        set_last_error("null argument, or capacities outside [1, 2^31 - 1]");
        return MLGPU_ERR_ARGUMENT;
    }
    *out = nullptr;
    int rc;
    if ((rc = check_world(s, world_frame)) || (rc = check_config(cfg))) return rc;
    if (hist && (rc = check_histogram(hist))) return rc;
    auto* p = new mlgpu_bunch_plan();
    p->scene = s;
    p->max_bunches = max_bunches;
    p->max_photons = max_photons;
    p->max_interactions = cfg->max_num_interactions_per_photon;
    const size_t cap = static_cast<size_t>(max_photons);
    p->head_bytes = round256(sizeof(mlgpu::CallParams)) + round256(sizeof(mlgpu::BunchDev) * max_bunches);
    const size_t photon_bytes = 7 * round256(cap * 8);
    const size_t result_bytes = 2 * round256(cap * 4) + 3 * round256(cap) + round256(cap * 8) + 4 * round256(cap * 4);
    const size_t bytes = p->head_bytes + photon_bytes + result_bytes + mlgpu::workspace_bytes(max_photons);
    hipError_t e = hipMalloc(&p->mem, bytes);
    if (e != hipSuccess) { p->mem = nullptr; plan_free(p); return hip_fail(e, "hipMalloc bunch plan"); }
    e = hipHostMalloc(reinterpret_cast<void**>(&p->staging), p->head_bytes);
    if (e != hipSuccess) { p->staging = nullptr; plan_free(p); return hip_fail(e, "hipHostMalloc bunch plan"); }
    std::memset(p->staging, 0, p->head_bytes);
    Carver c{static_cast<char*>(p->mem)};
    auto* params = reinterpret_cast<mlgpu::CallParams*>(c.p);
    auto* table = reinterpret_cast<mlgpu::BunchDev*>(c.p + round256(sizeof(mlgpu::CallParams)));
    c.p += p->head_bytes;
    mlgpu_photon_buffers gen{c.take<double>(cap), c.take<double>(cap), c.take<double>(cap), c.take<double>(cap),
                             c.take<double>(cap), c.take<double>(cap), c.take<double>(cap)};
    p->results = mlgpu_results{c.take<int32_t>(cap), c.take<int32_t>(cap), c.take<uint8_t>(cap), c.take<uint8_t>(cap),
                               c.take<uint8_t>(cap), c.take<double>(cap), c.take<float>(cap), c.take<float>(cap),
                               c.take<float>(cap), c.take<float>(cap), nullptr};
    const mlgpu::Workspace ws = carve_workspace(c.p, max_photons);
    const mlgpu_photons ph{gen.support_x, gen.support_y, gen.support_z, gen.direction_x, gen.direction_y,
                           gen.direction_z, gen.wavelength};
    e = hipEventCreateWithFlags(&p->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->capture, hipStreamNonBlocking);
    if (e != hipSuccess) { plan_free(p); return hip_fail(e, "bunch plan stream"); }
    e = hipStreamBeginCapture(p->capture, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) { plan_free(p); return hip_fail(e, "hipStreamBeginCapture"); }
    hipError_t le = hipMemcpyAsync(params, p->staging, p->head_bytes, hipMemcpyHostToDevice, p->capture);
    if (le == hipSuccess) le = mlgpu::launch_generate_bunches(table, max_bunches, max_photons, gen, p->capture, params);
    if (le == hipSuccess)
        le = launch_on_workspace(s,
                                 s->tuning.schedule == MLGPU_SCHEDULE_PERSISTENT ? MLGPU_SCHEDULE_PERSISTENT
                                                                                 : MLGPU_SCHEDULE_WAVEFRONT,
                                 s->scope_of_frame[world_frame], ph, max_photons, p->max_interactions, cfg->seed,
                                 cfg->photon_id_offset, p->results, ws, p->capture, params, nullptr);
    if (le == hipSuccess && hist) le = mlgpu::launch_histogram(p->results, max_photons, *hist, s->num_cus, p->capture,
                                                                params);
    e = hipStreamEndCapture(p->capture, &p->graph);
    if (le != hipSuccess || e != hipSuccess) {
        if (e == hipSuccess) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
        plan_free(p);
        return hip_fail(le != hipSuccess ? le : e, "bunch plan capture");
    }
    e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) { p->exec = nullptr; plan_free(p); return hip_fail(e, "hipGraphInstantiate"); }
    *out = p;
    return MLGPU_OK;
}

int mlgpu_bunch_plan_run(mlgpu_bunch_plan* p, const mlgpu_bunch* bunches, int32_t nb, const mlgpu_config* cfg,
                         void* stream) {
    if (!p || !cfg || nb < 0 || (nb > 0 && !bunches)) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    if (nb > p->max_bunches) { set_last_error("more bunches than the plan holds"); return MLGPU_ERR_ARGUMENT; }
    if (cfg->max_num_interactions_per_photon != p->max_interactions) {
        set_last_error("max_num_interactions_per_photon differs from the plan's");
        return MLGPU_ERR_ARGUMENT;
    }
    int64_t total = 0;
    for (int32_t b = 0; b < nb; b++) {
        if (bunches[b].num_photons < 0) { set_last_error("negative bunch size"); return MLGPU_ERR_ARGUMENT; }
        const int rc = check_source(&bunches[b].source);
        if (rc) return rc;
        total += bunches[b].num_photons;
    }
    if (total > p->max_photons) { set_last_error("more photons than the plan holds"); return MLGPU_ERR_ARGUMENT; }
    if (total == 0) return MLGPU_OK;
    std::lock_guard<std::mutex> lock(p->mutex);
    // the previous replay has read the staging buffer once it has finished
    if (p->launched) (void)hipEventSynchronize(p->done);
    auto* cp = reinterpret_cast<mlgpu::CallParams*>(p->staging);
    cp->n = total;
    cp->seed = cfg->seed;
    cp->id_offset = cfg->photon_id_offset;
    cp->num_bunches = nb;
    auto* table = reinterpret_cast<mlgpu::BunchDev*>(p->staging + round256(sizeof(mlgpu::CallParams)));
    int64_t start = 0;
    for (int32_t b = 0; b < nb; b++) {
        table[b] = mlgpu::BunchDev{bunches[b].source, bunches[b].seed, start};
        start += bunches[b].num_photons;
    }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e = hipGraphLaunch(p->exec, st);
    if (e == hipSuccess) e = hipEventRecord(p->done, st);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    p->launched = true;
    return MLGPU_OK;
}

int mlgpu_bunch_plan_results(const mlgpu_bunch_plan* p, mlgpu_results* out) {
    if (!p || !out) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    *out = p->results;
    return MLGPU_OK;
}

void mlgpu_bunch_plan_destroy(mlgpu_bunch_plan* p) {
    if (p) plan_free(p);
}
