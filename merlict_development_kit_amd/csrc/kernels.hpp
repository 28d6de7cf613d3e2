// Launchers of the HIP kernels (propagate.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/mlgpu.h"
#include "device_scene.hpp"

namespace mlgpu {

// Device workspace of the wavefront schedule.  Photon state is indexed by
// photon (length cap), trace records by queue slot (length cap).
// State of a live photon between bounce levels, in two records so that a
// gathered access touches one or two cache lines instead of one per field:
// the ray (what a trace reads) and the running sums (what the shading adds).
struct alignas(64) PhotonRay {
    double sx, sy, sz;
    double dx, dy, dz;
    double wavelength;    // the photon's, so the shading reads no input array
    int32_t root;         // scope the next trace runs in
    uint32_t hist_draws;  // history length so far (bits 0-7; the wavefront
                          // runs at most 255 levels) | Philox draws consumed << 8
};

struct alignas(16) PhotonSums {
    double tof;   // running time of flight
    double dist;  // running accumulated distance
};

struct Workspace {
    int64_t cap;
    PhotonRay* ray;
    PhotonSums* sums;
    int32_t* hit_leaf;   // best leaf so far per slot
    double* hit_t;       // its distance
    int32_t* tr_cur;     // traversal position of a deferred slot
    int32_t* tr_depth;   // its stack height
    int32_t* tr_stack;   // its stack, entry k at [k * cap + slot]
    int32_t* deferred;   // slots whose traversal goes on in the second pass
    int32_t* queue[2];
    int32_t* counts;     // [level]: live photons; [kDeferredCounts + level]: deferred slots
};

// Parameters a captured launch sequence (mlgpu_bunch_plan) reads from
// device memory at each replay: the photon count, the propagation seed and
// photon id offset, the number of bunches.  Kernels given a null pointer
// use their arguments.
struct CallParams {
    int64_t n;
    uint64_t seed;
    uint64_t id_offset;
    int32_t num_bunches;
    int32_t pad;
};

constexpr int kDeferredCounts = 256;
constexpr int kNumCounts = 2 * kDeferredCounts;
// a wavefront pass runs at most kDeferredCounts - 1 levels, so the history
// length of a live photon fits PhotonRay::hist_draws' 8 bits
static_assert(kDeferredCounts <= 256, "PhotonRay keeps the history length in 8 bits");

// bytes of workspace for cap photons (each array 256-byte aligned)
constexpr size_t workspace_bytes(int64_t cap) {
    const size_t m = static_cast<size_t>(cap);
    const size_t per = sizeof(PhotonRay) + sizeof(PhotonSums) + 4 + 8 + 4 + 4 + 4 * kMaxTreeDepth + 4 + 4 + 4;
    return m * per + 4 * kNumCounts + 40 * 256;
}

// mlgpu_propagate_noalloc: the most photons per call, whose wavefront
// workspace lives in a file-scope device array (propagate.hip), and the
// device address of this GPU's copy of that array (looked up once per
// device, with that device current).
constexpr int64_t kNoAllocMaxPhotons = int64_t{1} << 18;
hipError_t noalloc_workspace(void** ptr);

// Photon i of a batch draws from Philox stream id id_offset + i, or
// id_offset + ids[i] when ids is given (a batch gathered from a larger one).
hipError_t launch_propagate(const DeviceScene& sc, int32_t world_scope, const mlgpu_photons& in, int64_t n,
                            uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                            const mlgpu_results& out, hipStream_t stream, const int32_t* ids = nullptr);

// the fused schedule recording every photon's history (capacity >= max_interactions)
hipError_t launch_propagate_history(const DeviceScene& sc, int32_t world_scope, const mlgpu_photons& in, int64_t n,
                                    uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                                    const mlgpu_results& out, const mlgpu_history& hist, hipStream_t stream);

// n (and seed, id_offset) size the grids; with cp, the kernels take the
// photon count, seed and id offset from *cp (n is then the capacity)
hipError_t launch_propagate_wavefront(const DeviceScene& sc, int32_t world_scope, const mlgpu_photons& in,
                                      int64_t n, uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                                      const mlgpu_results& out, const Workspace& ws, int grid_blocks,
                                      int32_t trace_budget, int32_t tail_level, hipStream_t stream,
                                      const CallParams* cp = nullptr, const int32_t* ids = nullptr);

// The persistent schedule: counts is the workspace's counter block (its
// last 64-bit word is the photon chunk counter); quantum <= 0: traces run
// to their end before a shading round.
hipError_t launch_propagate_persistent(const DeviceScene& sc, int32_t world_scope, const mlgpu_photons& in,
                                       int64_t n, uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                                       const mlgpu_results& out, int32_t* counts, int grid_blocks, int32_t quantum,
                                       int32_t shade_lanes, hipStream_t stream, const CallParams* cp = nullptr,
                                       const int32_t* ids = nullptr);

hipError_t launch_first_hit(const DeviceScene& sc, int32_t root_scope, const mlgpu_photons& rays, int64_t n,
                            const mlgpu_hits& out, hipStream_t stream);

// A bunch on the device: its source, seed and first index in the
// concatenation of all bunches of a call.
struct BunchDev {
    mlgpu_source source;
    uint64_t seed;
    int64_t start;
};

hipError_t launch_generate_bunches(const BunchDev* table, int32_t num_bunches, int64_t total,
                                   const mlgpu_photon_buffers& out, hipStream_t stream,
                                   const CallParams* cp = nullptr);

hipError_t launch_generate(const mlgpu_source& src, int64_t n, uint64_t seed, uint64_t id_offset,
                           const mlgpu_photon_buffers& out, hipStream_t stream);

// Result i is jittered on the stream of photon id h.photon_id_offset + i, or
// h.photon_id_offset + ids[i] when ids is given.
hipError_t launch_histogram(const mlgpu_results& r, int64_t n, const mlgpu_histogram& h, int num_cus,
                            hipStream_t stream, const CallParams* cp = nullptr, const int32_t* ids = nullptr);

hipError_t launch_count_errors(const uint8_t* err, int64_t n, unsigned long long* counter, hipStream_t stream);

// CORSIKA bunches (corsika.hip): workspace of n bunches; count + scan
// (bad weights flagged in *bad, the photon total in *total), then emit.
size_t corsika_workspace_bytes(int64_t n);
hipError_t launch_corsika_count(const float* bunches, int64_t n, uint64_t seed, uint64_t id_offset, void* ws,
                                int32_t* bad, int64_t* total, hipStream_t stream);
hipError_t launch_corsika_emit(const float* bunches, int64_t n, uint64_t seed, uint64_t id_offset, const void* ws,
                               int64_t capacity, const mlgpu_photon_buffers& out, int32_t* truth,
                               hipStream_t stream);

// Pin-hole camera in the world frame (PinHoleCamera::set_pos_rot_fov):
// position, principal point, the sensor's row and column steps.
struct RenderCamera {
    double pos[3], pp[3], srow[3], scol[3];
    int32_t cols, rows;
};
struct RenderParams {
    int32_t gi;         // global illumination on
    int32_t max_depth;  // reflections + refractions before a ray ends black
    double light[3];    // incoming direction of the global light
    float sky[3];
    float pad;
};
// One pixel per lane (render.hip): image rows x cols x 3 floats.
hipError_t launch_render(const DeviceScene& sc, int32_t world_scope, const RenderCamera& cam,
                         const RenderParams& prm, const float* colors, uint64_t seed, float* image,
                         unsigned long long* errors, hipStream_t stream);

// Calibration photon records of n results of `src` photons (calibration.hip).
hipError_t launch_calibration_photons(const mlgpu_source& src, uint64_t seed, uint64_t id_offset,
                                      const mlgpu_results& r, int64_t n, int32_t num_lixel,
                                      mlgpu_calibration_photon* out, hipStream_t stream);

}  // namespace mlgpu
