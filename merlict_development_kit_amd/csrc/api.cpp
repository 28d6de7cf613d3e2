// C ABI of the device side (include/mlgpu.h): scene upload, propagation,
// probe-ray first hits.  Never throws across the ABI; errors are status
// codes plus mlgpu_last_error().
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"
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
    mlgpu_tuning tuning{MLGPU_SCHEDULE_AUTO, 24, 32};
    // wavefront workspace, grown on demand.  Calls that use it are ordered
    // on the device: a call on another stream than the previous user first
    // waits for the event recorded after that user's kernels.
    std::mutex ws_mutex;
    void* ws = nullptr;
    int64_t ws_capacity = 0;
    hipEvent_t ws_event = nullptr;
    hipStream_t ws_stream = nullptr;
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

}  // namespace

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
        (rc = upload(s, hs.points, &s->dev.points))) {
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
    for (void* p : s->buffers) (void)hipFree(p);
    if (s->ws) (void)hipFree(s->ws);
    delete s;
}

int mlgpu_scene_set_tuning(mlgpu_scene* s, const mlgpu_tuning* t) {
    if (!s || !t) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    if (t->schedule < MLGPU_SCHEDULE_AUTO || t->schedule > MLGPU_SCHEDULE_WAVEFRONT) {
        set_last_error("bad schedule");
        return MLGPU_ERR_ARGUMENT;
    }
    if (t->blocks_per_cu < 1 || t->blocks_per_cu > 64) {
        set_last_error("blocks_per_cu must be in [1, 64]");
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

int mlgpu_scene_stats(const mlgpu_scene* s, int64_t* num_nodes, int64_t* num_leaves, int64_t* device_bytes) {
    if (!s) { set_last_error("null scene"); return MLGPU_ERR_ARGUMENT; }
    if (num_nodes) *num_nodes = s->num_nodes;
    if (num_leaves) *num_leaves = s->num_leaves;
    if (device_bytes) *device_bytes = s->device_bytes;
    return MLGPU_OK;
}

int mlgpu_propagate(const mlgpu_scene* s, int32_t world_frame, const mlgpu_photons* photons, int64_t n,
                    const mlgpu_config* cfg, mlgpu_results* r, void* stream) {
    if (!s || !cfg || !r || n < 0) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    if (n == 0) return MLGPU_OK;
    if (!photons_ok(photons)) { set_last_error("null photon array"); return MLGPU_ERR_ARGUMENT; }
    if (!r->final_frame || !r->sensor_id || !r->final_interaction || !r->num_interactions || !r->error ||
        !r->time_of_flight || !r->x || !r->y || !r->cx || !r->cy) {
        set_last_error("null result array");
        return MLGPU_ERR_ARGUMENT;
    }
    if (world_frame < 0 || world_frame >= static_cast<int32_t>(s->scope_of_frame.size()) ||
        s->scope_of_frame[world_frame] < 0) {
        set_last_error("world_frame must be frame 0 or a restriction frame");
        return MLGPU_ERR_ARGUMENT;
    }
    if (cfg->max_num_interactions_per_photon > 250) {
        set_last_error("max_num_interactions_per_photon must be <= 250");
        return MLGPU_ERR_ARGUMENT;
    }
    if (n > INT32_MAX) { set_last_error("at most 2^31 - 1 photons per call"); return MLGPU_ERR_ARGUMENT; }
    int schedule = s->tuning.schedule;
    if (schedule == MLGPU_SCHEDULE_AUTO) schedule = n >= 65536 ? MLGPU_SCHEDULE_WAVEFRONT : MLGPU_SCHEDULE_FUSED;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e;
    if (schedule == MLGPU_SCHEDULE_FUSED) {
        e = mlgpu::launch_propagate(s->dev, s->scope_of_frame[world_frame], *photons, n,
                                    cfg->max_num_interactions_per_photon, cfg->seed, cfg->photon_id_offset, *r, st);
    } else {
        auto* ms = const_cast<mlgpu_scene*>(s);
        std::lock_guard<std::mutex> lock(ms->ws_mutex);
        if (!ms->ws_event) {
            e = hipEventCreateWithFlags(&ms->ws_event, hipEventDisableTiming);
            if (e != hipSuccess) { ms->ws_event = nullptr; return hip_fail(e, "hipEventCreate"); }
        } else if (ms->ws_stream != st) {
            e = hipStreamWaitEvent(st, ms->ws_event, 0);
            if (e != hipSuccess) return hip_fail(e, "hipStreamWaitEvent");
        }
        if (ms->ws_capacity < n) {
            if (ms->ws) {
                (void)hipEventSynchronize(ms->ws_event);
                (void)hipFree(ms->ws);
                ms->ws = nullptr;
                ms->ws_capacity = 0;
            }
            const int64_t cap = n + n / 8;
            e = hipMalloc(&ms->ws, mlgpu::workspace_bytes(cap));
            if (e != hipSuccess) return hip_fail(e, "hipMalloc workspace");
            ms->ws_capacity = cap;
        }
        // carve the workspace into 256-byte aligned arrays
        char* p = static_cast<char*>(ms->ws);
        const size_t cap = static_cast<size_t>(ms->ws_capacity);
        auto take = [&](size_t bytes) { char* q = p; p += (bytes + 255) & ~static_cast<size_t>(255); return q; };
        mlgpu::Workspace w;
        w.cap = ms->ws_capacity;
        w.sx = reinterpret_cast<double*>(take(8 * cap));
        w.sy = reinterpret_cast<double*>(take(8 * cap));
        w.sz = reinterpret_cast<double*>(take(8 * cap));
        w.dx = reinterpret_cast<double*>(take(8 * cap));
        w.dy = reinterpret_cast<double*>(take(8 * cap));
        w.dz = reinterpret_cast<double*>(take(8 * cap));
        w.tof = reinterpret_cast<double*>(take(8 * cap));
        w.root = reinterpret_cast<int32_t*>(take(4 * cap));
        w.n_hist = reinterpret_cast<uint32_t*>(take(4 * cap));
        w.draws = reinterpret_cast<uint32_t*>(take(4 * cap));
        w.hit_leaf = reinterpret_cast<int32_t*>(take(4 * cap));
        w.hit_t = reinterpret_cast<double*>(take(8 * cap));
        w.tr_cur = reinterpret_cast<int32_t*>(take(4 * cap));
        w.tr_depth = reinterpret_cast<int32_t*>(take(4 * cap));
        w.tr_stack = reinterpret_cast<int32_t*>(take(4 * cap * mlgpu::kMaxTreeDepth));
        w.deferred = reinterpret_cast<int32_t*>(take(4 * cap));
        w.queue[0] = reinterpret_cast<int32_t*>(take(4 * cap));
        w.queue[1] = reinterpret_cast<int32_t*>(take(4 * cap));
        w.counts = reinterpret_cast<int32_t*>(take(4 * mlgpu::kNumCounts));
        e = mlgpu::launch_propagate_wavefront(s->dev, s->scope_of_frame[world_frame], *photons, n,
                                              cfg->max_num_interactions_per_photon, cfg->seed, cfg->photon_id_offset,
                                              *r, w, s->tuning.blocks_per_cu * s->num_cus,
                                              s->tuning.trace_budget, st);
        if (e == hipSuccess) e = hipEventRecord(ms->ws_event, st);
        ms->ws_stream = st;
    }
    if (e != hipSuccess) return hip_fail(e, "propagate launch");
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
    if (src->support != MLGPU_SUPPORT_POINT && src->support != MLGPU_SUPPORT_DISC) {
        set_last_error("bad support kind");
        return MLGPU_ERR_ARGUMENT;
    }
    if (src->direction < MLGPU_DIRECTION_AXIS || src->direction > MLGPU_DIRECTION_CAP) {
        set_last_error("bad direction kind");
        return MLGPU_ERR_ARGUMENT;
    }
    const bool bounded = src->direction == MLGPU_DIRECTION_TILT || src->direction == MLGPU_DIRECTION_CAP;
    if (!(src->angle >= 0.0) || (bounded && !(src->angle <= 1.57079632679489661923)) || !std::isfinite(src->angle)) {
        set_last_error("angle must be finite and >= 0, and <= pi/2 for TILT and CAP");
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
    const hipError_t e = mlgpu::launch_generate(*src, n, seed, photon_id_offset, *out, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "generate launch");
    return MLGPU_OK;
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
