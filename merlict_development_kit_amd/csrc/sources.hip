// HIP kernels of the photon sources (sources_core.hpp): one lane per photon.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "sources_core.hpp"

namespace mlgpu {
namespace {

__global__ __launch_bounds__(256) void generate_kernel(const mlgpu_source src, const int64_t n, const uint64_t seed,
                                                       const uint64_t id_offset, const mlgpu_photon_buffers out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    D3 s, d;
    double lambda;
    source_photon(src, seed, id_offset + static_cast<uint64_t>(i), s, d, lambda);
    out.support_x[i] = s.x;
    out.support_y[i] = s.y;
    out.support_z[i] = s.z;
    out.direction_x[i] = d.x;
    out.direction_y[i] = d.y;
    out.direction_z[i] = d.z;
    out.wavelength[i] = lambda;
}

// Photon i of the concatenation of a call's bunches: its bunch is the last
// one starting at or before i.
__global__ __launch_bounds__(256) void generate_bunches_kernel(const BunchDev* __restrict__ table,
                                                               const int32_t num_bunches_arg,
                                                               const int64_t total_arg,
                                                               const mlgpu_photon_buffers out,
                                                               const CallParams* __restrict__ cp) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t total = cp ? cp->n : total_arg;
    const int32_t num_bunches = cp ? cp->num_bunches : num_bunches_arg;
    if (i >= total || num_bunches <= 0) return;
    int32_t lo = 0, hi = num_bunches - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (table[mid].start <= i) lo = mid; else hi = mid - 1;
    }
    const BunchDev& b = table[lo];
    D3 s, d;
    double lambda;
    source_photon(b.source, b.seed, static_cast<uint64_t>(i - b.start), s, d, lambda);
    out.support_x[i] = s.x;
    out.support_y[i] = s.y;
    out.support_z[i] = s.z;
    out.direction_x[i] = d.x;
    out.direction_y[i] = d.y;
    out.direction_z[i] = d.z;
    out.wavelength[i] = lambda;
}

}  // namespace

hipError_t launch_generate_bunches(const BunchDev* table, int32_t num_bunches, int64_t total,
                                   const mlgpu_photon_buffers& out, hipStream_t stream, const CallParams* cp) {
    if (total <= 0) return hipSuccess;
    const int64_t grid = (total + 255) / 256;
    hipLaunchKernelGGL(generate_bunches_kernel, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, stream, table,
                       num_bunches, total, out, cp);
    return hipGetLastError();
}

hipError_t launch_generate(const mlgpu_source& src, int64_t n, uint64_t seed, uint64_t id_offset,
                           const mlgpu_photon_buffers& out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t grid = (n + 255) / 256;
    hipLaunchKernelGGL(generate_kernel, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, stream, src, n, seed,
                       id_offset, out);
    return hipGetLastError();
}

}  // namespace mlgpu
