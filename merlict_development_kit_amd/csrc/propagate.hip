// HIP kernels of the photon propagator for CDNA4 (gfx950).
//
// One lane per photon.  Per bounce:
//   1. traversal: FP32 bounding-sphere tests over the depth-first node array
//      of the current root scope, stackless via skip indices, pruned by the
//      best hit found so far (merlict/RayAndFrame.cpp:137-149 visits every
//      candidate; the first minimum is the same);
//   2. leaf tests: FP64 in the object frame, with the ray support moved
//      along the ray to the point closest to the leaf origin so the solve
//      works on small numbers (merlict's equations, scenery/geometry/*.cpp);
//   3. physics: medium absorption, surface reflection, Fresnel, refraction
//      and root-frame switching exactly as merlict/PhotonAndFrame.cpp:23-124,
//      with one Philox4x32-10 stream per photon.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "kernels.hpp"
#include "propagate_core.hpp"

namespace mlgpu {
namespace {

__global__ __launch_bounds__(256) void propagate_kernel(const DeviceScene sc, const int32_t world_node,
                                                        const mlgpu_photons in, const int64_t n,
                                                        const uint32_t max_interactions, const uint64_t seed,
                                                        const uint64_t id_offset, const mlgpu_results out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    propagate_photon(sc, world_node, in, i, max_interactions, seed, id_offset, out);
}

__global__ __launch_bounds__(256) void first_hit_kernel(const DeviceScene sc, const int32_t root_node,
                                                        const mlgpu_photons rays, const int64_t n,
                                                        const mlgpu_hits out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const D3 s{rays.support_x[i], rays.support_y[i], rays.support_z[i]};
    const D3 d = normalized(D3{rays.direction_x[i], rays.direction_y[i], rays.direction_z[i]});
    Closest c;
    first_hit(sc, root_node, s, d, c);
    if (c.leaf < 0) {
        const D3 e = at(s, d, kVoidDistance);
        out.frame[i] = -1;
        out.distance[i] = kVoidDistance;
        out.position_x[i] = e.x;
        out.position_y[i] = e.y;
        out.position_z[i] = e.z;
        return;
    }
    const D3 w = position(sc.geo[c.leaf], c.hit.p);
    out.frame[i] = sc.phys[c.leaf].frame;
    out.distance[i] = c.hit.t;
    out.position_x[i] = w.x;
    out.position_y[i] = w.y;
    out.position_z[i] = w.z;
}

__global__ void count_errors_kernel(const uint8_t* err, int64_t n, unsigned long long* counter) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool bad = i < n && err[i] != 0;
    const unsigned long long ballot = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && ballot) atomicAdd(counter, static_cast<unsigned long long>(__popcll(ballot)));
}

}  // namespace

hipError_t launch_propagate(const DeviceScene& sc, int32_t world_node, const mlgpu_photons& in, int64_t n,
                            uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                            const mlgpu_results& out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int block = 256;
    const int64_t grid = (n + block - 1) / block;
    hipLaunchKernelGGL(propagate_kernel, dim3(static_cast<uint32_t>(grid)), dim3(block), 0, stream, sc, world_node,
                       in, n, max_interactions, seed, id_offset, out);
    return hipGetLastError();
}

hipError_t launch_first_hit(const DeviceScene& sc, int32_t root_node, const mlgpu_photons& rays, int64_t n,
                            const mlgpu_hits& out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int block = 256;
    const int64_t grid = (n + block - 1) / block;
    hipLaunchKernelGGL(first_hit_kernel, dim3(static_cast<uint32_t>(grid)), dim3(block), 0, stream, sc, root_node,
                       rays, n, out);
    return hipGetLastError();
}

hipError_t launch_count_errors(const uint8_t* err, int64_t n, unsigned long long* counter, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int block = 256;
    const int64_t grid = (n + block - 1) / block;
    hipLaunchKernelGGL(count_errors_kernel, dim3(static_cast<uint32_t>(grid)), dim3(block), 0, stream, err, n,
                       counter);
    return hipGetLastError();
}

}  // namespace mlgpu
