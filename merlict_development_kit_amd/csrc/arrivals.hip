// Arrival-time histograms (arrivals_core.hpp).  Small histograms (at most
// kLdsCells cells, e.g. 1440 pixels x 8 bins) are privatised per block in
// LDS and merged with one global atomic per touched cell; large ones (the
// plenoscope's 515,023 lixels) take global atomics directly, which rarely
// collide there.
#include <hip/hip_runtime.h>

#include "arrivals_core.hpp"
#include "kernels.hpp"

namespace mlgpu {
namespace {

constexpr int kLdsCells = 12288;  // 48 KB

__global__ __launch_bounds__(256) void histogram_lds_kernel(const mlgpu_results r, const int64_t n_arg,
                                                            const mlgpu_histogram h,
                                                            const CallParams* __restrict__ cp,
                                                            const int32_t* __restrict__ ids) {
    const int64_t n = cp ? cp->n : n_arg;
    __shared__ uint32_t cells[kLdsCells];
    const int32_t total = h.num_sensors * h.num_bins;
    for (int32_t c = static_cast<int32_t>(threadIdx.x); c < total; c += 256) cells[c] = 0u;
    __syncthreads();
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * 256) {
        const int64_t cell = arrival_cell(h, r.sensor_id[i], r.error[i], r.time_of_flight[i],
                                          h.photon_id_offset + static_cast<uint64_t>(ids ? ids[i] : i));
        if (cell >= 0) atomicAdd(&cells[cell], 1u);
    }
    __syncthreads();
    for (int32_t c = static_cast<int32_t>(threadIdx.x); c < total; c += 256)
        if (cells[c] != 0u) atomicAdd(&h.counts[c], cells[c]);
}

__global__ __launch_bounds__(256) void histogram_global_kernel(const mlgpu_results r, const int64_t n_arg,
                                                               const mlgpu_histogram h,
                                                               const CallParams* __restrict__ cp,
                                                               const int32_t* __restrict__ ids) {
    const int64_t n = cp ? cp->n : n_arg;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t cell = arrival_cell(h, r.sensor_id[i], r.error[i], r.time_of_flight[i],
                                      h.photon_id_offset + static_cast<uint64_t>(ids ? ids[i] : i));
    if (cell >= 0) atomicAdd(&h.counts[cell], 1u);
}

}  // namespace

hipError_t launch_histogram(const mlgpu_results& r, int64_t n, const mlgpu_histogram& h, int num_cus,
                            hipStream_t stream, const CallParams* cp, const int32_t* ids) {
    if (n <= 0) return hipSuccess;
    const int64_t cells = static_cast<int64_t>(h.num_sensors) * h.num_bins;
    if (cells <= kLdsCells) {
        const int64_t need = (n + 255) / 256;
        const int64_t cap = static_cast<int64_t>(num_cus) * 8;
        const int grid = static_cast<int>(need < cap ? need : cap);
        hipLaunchKernelGGL(histogram_lds_kernel, dim3(grid), dim3(256), 0, stream, r, n, h, cp, ids);
    } else {
        hipLaunchKernelGGL(histogram_global_kernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0,
                           stream, r, n, h, cp, ids);
    }
    return hipGetLastError();
}

}  // namespace mlgpu
