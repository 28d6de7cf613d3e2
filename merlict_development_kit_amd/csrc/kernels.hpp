// Launchers of the HIP kernels (propagate.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/mlgpu.h"
#include "device_scene.hpp"

namespace mlgpu {

hipError_t launch_propagate(const DeviceScene& sc, int32_t world_node, const mlgpu_photons& in, int64_t n,
                            uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                            const mlgpu_results& out, hipStream_t stream);

hipError_t launch_first_hit(const DeviceScene& sc, int32_t root_node, const mlgpu_photons& rays, int64_t n,
                            const mlgpu_hits& out, hipStream_t stream);

hipError_t launch_count_errors(const uint8_t* err, int64_t n, unsigned long long* counter, hipStream_t stream);

}  // namespace mlgpu
