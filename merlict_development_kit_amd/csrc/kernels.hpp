// Launchers of the HIP kernels (propagate.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/mlgpu.h"
#include "device_scene.hpp"

namespace mlgpu {

// Device workspace of the wavefront schedule, all arrays of length n (the
// photon index) except the per-slot hit record and the two queues.
struct Workspace {
    double *sx, *sy, *sz, *dx, *dy, *dz, *tof;
    int32_t* root;
    uint32_t* n_hist;
    uint32_t* draws;
    int32_t* hit_leaf;
    double* hit_t;
    int32_t* queue[2];
    int32_t* counts;  // live photons per bounce level
};

// bytes of workspace for n photons
inline size_t workspace_bytes(int64_t n) {
    const size_t m = static_cast<size_t>(n);
    return m * (7 * 8 + 4 + 4 + 4 + 4 + 8 + 4 + 4) + 256 * 4 + 16 * 256 + 1024;
}

hipError_t launch_propagate(const DeviceScene& sc, int32_t world_scope, const mlgpu_photons& in, int64_t n,
                            uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                            const mlgpu_results& out, hipStream_t stream);

hipError_t launch_propagate_wavefront(const DeviceScene& sc, int32_t world_scope, const mlgpu_photons& in,
                                      int64_t n, uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                                      const mlgpu_results& out, const Workspace& ws, int grid_blocks,
                                      hipStream_t stream);

hipError_t launch_first_hit(const DeviceScene& sc, int32_t root_scope, const mlgpu_photons& rays, int64_t n,
                            const mlgpu_hits& out, hipStream_t stream);

hipError_t launch_count_errors(const uint8_t* err, int64_t n, unsigned long long* counter, hipStream_t stream);

}  // namespace mlgpu
