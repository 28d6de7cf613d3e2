// Plenoscope light-field calibration on the GPU: the results of the
// calibration photons (merlict_portal_plenoscope/calibration/
// Calibrator.cpp:48-102) and the per-lixel statistics of
// LixelStatisticsFiller (LixelStatisticsFiller.cpp:8-62, OnlineStatistics.cpp:
// 15-45).
//
// The filler keeps the reference's online state per lixel (a count, and n,
// sum, Welford mean and m2 of cx, cy, x, y and the time delay, in double) as
// structure-of-arrays in HBM.  A block of records is sorted stably by lixel
// (records that did not reach a lixel sort last), and one lane per lixel then
// adds its records in record order, so every sum is formed in the reference's
// order and the statistics agree bit for bit.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>
#include <string>

#include "../../include/mlgpu.h"
#include "kernels.hpp"
#include "pool.hpp"
#include "propagate_core.hpp"
#include "sources_core.hpp"

namespace mlgpu {
void set_last_error(const std::string& msg);

namespace {

constexpr uint32_t kNotReached = 0xffffffffu;
constexpr int kQuantities = 5;  // cx, cy, x, y, time delay

__global__ void calibration_photons_kernel(const mlgpu_source src, const uint64_t seed, const uint64_t id_offset,
                                           const mlgpu_results r, const int64_t n, const int32_t num_lixel,
                                           mlgpu_calibration_photon* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t s = r.sensor_id[i];
    mlgpu_calibration_photon c{};
    if (s >= 0 && s < num_lixel && r.error[i] == 0) {
        double x, y, lambda;
        D3 d;
        source_local(src, seed, id_offset + static_cast<uint64_t>(i), x, y, d, lambda);
        const double* R = src.rot;
        // the posed support (z = 0 before the pose) and the reverse of the
        // posed travel direction
        const double ax = x * R[0] + y * R[1] + 0.0 * R[2] + src.pos[0];
        const double ay = x * R[3] + y * R[4] + 0.0 * R[5] + src.pos[1];
        const double cx = -(d.x * R[0] + d.y * R[1] + d.z * R[2]);
        const double cy = -(d.x * R[3] + d.y * R[4] + d.z * R[5]);
        c.reached_sensor = 1u;
        c.lixel_id = static_cast<uint32_t>(s);
        c.x_pos_on_principal_aperture = static_cast<float>(ax);
        c.y_pos_on_principal_aperture = static_cast<float>(ay);
        c.x_tilt_vs_optical_axis = static_cast<float>(cx);
        c.y_tilt_vs_optical_axis = static_cast<float>(cy);
        c.time_of_flight = static_cast<float>(r.time_of_flight[i] - src.start_distance / kC);
    }
    out[i] = c;
}

// Online state of every lixel (OnlineLixelStatistics.h:9-17); n_samples of
// each quantity equals the count.
struct FillerState {
    uint32_t* count;
    double* sum[kQuantities];
    double* mean[kQuantities];
    double* m2[kQuantities];
};

__global__ void lixel_keys_kernel(const mlgpu_calibration_photon* __restrict__ b, const int64_t n,
                                  const int32_t num_lixel, uint32_t* __restrict__ key, int32_t* __restrict__ idx,
                                  int32_t* __restrict__ bad) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = kNotReached;
    if (b[i].reached_sensor) {
        if (b[i].lixel_id < static_cast<uint32_t>(num_lixel)) k = b[i].lixel_id; else atomicOr(bad, 1);
    }
    key[i] = k;
    idx[i] = static_cast<int32_t>(i);
}

// first position of lixel s in the sorted keys (s <= num_lixel)
__global__ void lixel_offsets_kernel(const uint32_t* __restrict__ key, const int64_t n, const int32_t num_lixel,
                                     int64_t* __restrict__ offsets) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s > num_lixel) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (static_cast<int64_t>(key[mid]) < s) lo = mid + 1; else hi = mid;
    }
    offsets[s] = lo;
}

// OnlineStatistics::add (OnlineStatistics.cpp:15-24)
__device__ __forceinline__ void online_add(double& sum, double& mean, double& m2, const double n, const double x) {
    const double delta = x - mean;
    mean += delta / n;
    m2 += delta * (x - mean);
    sum += x;
}

// LixelStatisticsFiller::fill_in_block (LixelStatisticsFiller.cpp:20-41)
// for the records of lixel s, in record order
__global__ void lixel_fill_kernel(const mlgpu_calibration_photon* __restrict__ b, const int32_t* __restrict__ idx,
                                  const int64_t* __restrict__ offsets, const int32_t num_lixel, FillerState st) {
    const int32_t s = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= num_lixel) return;
    const int64_t lo = offsets[s], hi = offsets[s + 1];
    if (lo == hi) return;
    uint32_t count = st.count[s];
    double sum[kQuantities], mean[kQuantities], m2[kQuantities];
#pragma unroll
    for (int q = 0; q < kQuantities; q++) {
        sum[q] = st.sum[q][s];
        mean[q] = st.mean[q][s];
        m2[q] = st.m2[q][s];
    }
    for (int64_t k = lo; k < hi; k++) {
        const mlgpu_calibration_photon c = b[idx[k]];
        count++;
        const double n = static_cast<double>(count);
        const double v[kQuantities] = {c.x_tilt_vs_optical_axis, c.y_tilt_vs_optical_axis,
                                       c.x_pos_on_principal_aperture, c.y_pos_on_principal_aperture,
                                       c.time_of_flight};
#pragma unroll
        for (int q = 0; q < kQuantities; q++) online_add(sum[q], mean[q], m2[q], n, v[q]);
    }
    st.count[s] = count;
#pragma unroll
    for (int q = 0; q < kQuantities; q++) {
        st.sum[q][s] = sum[q];
        st.mean[q][s] = mean[q];
        st.m2[q][s] = m2[q];
    }
}

// LixelStatisticsFiller::get_lixel_statistics (LixelStatisticsFiller.cpp:
// 43-62) with OnlineStatistics::mean / stddev (OnlineStatistics.cpp:26-45)
__global__ void lixel_statistics_kernel(const FillerState st, const int32_t num_lixel, const double emitted,
                                        mlgpu_lixel_statistic* __restrict__ out) {
    const int32_t s = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= num_lixel) return;
    const uint32_t count = st.count[s];
    const double n = static_cast<double>(count);
    float m[kQuantities], sd[kQuantities];
    // The reference's files hold the NaNs of an x86 host: 0 / 0 (the mean of
    // no sample) is the negative default NaN, quiet_NaN() (the deviation
    // below two samples) the positive one; both are written as such so that
    // lixel_statistics.bin agrees byte for byte.
    const float mean_nan = __uint_as_float(0xffc00000u), std_nan = __uint_as_float(0x7fc00000u);
#pragma unroll
    for (int q = 0; q < kQuantities; q++) {
        m[q] = count == 0u ? mean_nan : static_cast<float>(st.sum[q][s] / n);
        sd[q] = n < 2.0 ? std_nan : static_cast<float>(sqrt(st.m2[q][s] / n));
    }
    mlgpu_lixel_statistic r;
    r.efficiency = static_cast<float>(static_cast<double>(count) / emitted);
    r.efficiency_std = static_cast<float>(static_cast<double>(sqrtf(static_cast<float>(count))) / emitted);
    r.cx_mean = m[0];
    r.cx_std = sd[0];
    r.cy_mean = m[1];
    r.cy_std = sd[1];
    r.x_mean = m[2];
    r.x_std = sd[2];
    r.y_mean = m[3];
    r.y_std = sd[3];
    r.time_delay_mean = m[4];
    r.time_delay_std = sd[4];
    out[s] = r;
}

size_t align_up(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

int hip_fail(hipError_t e, const char* what) {
    set_last_error(std::string(what) + ": " + hipGetErrorString(e));
    return MLGPU_ERR_HIP;
}

}  // namespace

hipError_t launch_calibration_photons(const mlgpu_source& src, uint64_t seed, uint64_t id_offset,
                                      const mlgpu_results& r, int64_t n, int32_t num_lixel,
                                      mlgpu_calibration_photon* out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(calibration_photons_kernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0,
                       stream, src, seed, id_offset, r, n, num_lixel, out);
    return hipGetLastError();
}

}  // namespace mlgpu

struct mlgpu_lixel_filler {
    int64_t num_lixel = 0;
    double emitted = 0.0;  // photons_emitted_per_lixel
    char* mem = nullptr;
    mlgpu::FillerState st{};
};

extern "C" {

int mlgpu_lixel_filler_create(int64_t num_lixel, int64_t num_blocks, int64_t photons_per_block,
                              mlgpu_lixel_filler** out) {
    if (!out || num_lixel <= 0 || num_lixel > INT32_MAX || num_blocks < 0 || photons_per_block < 0) {
        mlgpu::set_last_error("need num_lixel in [1, 2^31 - 1], num_blocks >= 0, photons_per_block >= 0");
        return MLGPU_ERR_ARGUMENT;
    }
    *out = nullptr;
    auto* f = new mlgpu_lixel_filler();
    f->num_lixel = num_lixel;
    f->emitted = static_cast<double>(num_blocks) * static_cast<double>(photons_per_block) /
                 static_cast<double>(num_lixel);
    const size_t L = static_cast<size_t>(num_lixel);
    const size_t bytes = mlgpu::align_up(L * 4) + 3 * mlgpu::kQuantities * mlgpu::align_up(L * 8);
    hipError_t e = hipMalloc(&f->mem, bytes);
    if (e != hipSuccess) {
        delete f;
        return mlgpu::hip_fail(e, "hipMalloc lixel filler");
    }
    char* q = f->mem;
    f->st.count = reinterpret_cast<uint32_t*>(q);
    q += mlgpu::align_up(L * 4);
    for (int k = 0; k < mlgpu::kQuantities; k++) {
        f->st.sum[k] = reinterpret_cast<double*>(q);
        q += mlgpu::align_up(L * 8);
        f->st.mean[k] = reinterpret_cast<double*>(q);
        q += mlgpu::align_up(L * 8);
        f->st.m2[k] = reinterpret_cast<double*>(q);
        q += mlgpu::align_up(L * 8);
    }
    // OnlineStatistics(): every field 0.0, count 0
    e = hipMemset(f->mem, 0, bytes);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        (void)hipFree(f->mem);
        delete f;
        return mlgpu::hip_fail(e, "lixel filler init");
    }
    *out = f;
    return MLGPU_OK;
}

int mlgpu_lixel_filler_fill_block(mlgpu_lixel_filler* f, const mlgpu_calibration_photon* block, int64_t n,
                                  void* stream) {
    if (!f || n < 0 || (n > 0 && !block)) {
        mlgpu::set_last_error("null argument or negative size");
        return MLGPU_ERR_ARGUMENT;
    }
    if (n > INT32_MAX) {
        mlgpu::set_last_error("at most 2^31 - 1 records per block");
        return MLGPU_ERR_ARGUMENT;
    }
    if (n == 0) return MLGPU_OK;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t N = static_cast<size_t>(n), L = static_cast<size_t>(f->num_lixel);
    const int32_t num_lixel = static_cast<int32_t>(f->num_lixel);
    size_t temp_bytes = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, temp_bytes, static_cast<uint32_t*>(nullptr),
                                             static_cast<uint32_t*>(nullptr), static_cast<int32_t*>(nullptr),
                                             static_cast<int32_t*>(nullptr), N, 0, 32, st);
    if (e != hipSuccess) return mlgpu::hip_fail(e, "lixel sort");
    const size_t bytes = 2 * mlgpu::align_up(N * 4) + 2 * mlgpu::align_up(N * 4) + mlgpu::align_up((L + 1) * 8) +
                         mlgpu::align_up(temp_bytes) + mlgpu::align_up(4);
    mlgpu::PoolBlock pool_block;
    e = pool_block.get(bytes);
    if (e != hipSuccess) return mlgpu::hip_fail(e, "lixel block workspace");
    char* mem = pool_block.as();
    char* q = mem;
    auto take = [&](size_t b) { char* x = q; q += mlgpu::align_up(b); return x; };
    auto* key = reinterpret_cast<uint32_t*>(take(N * 4));
    auto* key2 = reinterpret_cast<uint32_t*>(take(N * 4));
    auto* idx = reinterpret_cast<int32_t*>(take(N * 4));
    auto* idx2 = reinterpret_cast<int32_t*>(take(N * 4));
    auto* off = reinterpret_cast<int64_t*>(take((L + 1) * 8));
    void* temp = take(temp_bytes);
    auto* bad = reinterpret_cast<int32_t*>(take(4));
    int32_t host_bad = 0;
    e = hipMemsetAsync(bad, 0, 4, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mlgpu::lixel_keys_kernel, dim3(static_cast<uint32_t>((N + 255) / 256)), dim3(256), 0, st,
                           block, n, num_lixel, key, idx, bad);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&host_bad, bad, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && !host_bad) {
        // stable: records of one lixel keep their order
        e = rocprim::radix_sort_pairs(temp, temp_bytes, key, key2, idx, idx2, N, 0, 32, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(mlgpu::lixel_offsets_kernel, dim3(static_cast<uint32_t>((L + 1 + 255) / 256)),
                               dim3(256), 0, st, key2, n, num_lixel, off);
            hipLaunchKernelGGL(mlgpu::lixel_fill_kernel, dim3(static_cast<uint32_t>((L + 255) / 256)), dim3(256), 0,
                               st, block, idx2, off, num_lixel, f->st);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    (void)hipStreamSynchronize(st);  // the borrowed block is idle when returned
    pool_block.release();
    if (e != hipSuccess) return mlgpu::hip_fail(e, "lixel statistics fill");
    if (host_bad) {
        mlgpu::set_last_error("a record reached a lixel_id >= num_lixel");
        return MLGPU_ERR_ARGUMENT;
    }
    return MLGPU_OK;
}

int mlgpu_lixel_filler_statistics(const mlgpu_lixel_filler* f, mlgpu_lixel_statistic* out, void* stream) {
    if (!f || !out) {
        mlgpu::set_last_error("null argument");
        return MLGPU_ERR_ARGUMENT;
    }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t grid = static_cast<uint32_t>((f->num_lixel + 255) / 256);
    hipLaunchKernelGGL(mlgpu::lixel_statistics_kernel, dim3(grid), dim3(256), 0, st, f->st,
                       static_cast<int32_t>(f->num_lixel), f->emitted, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mlgpu::hip_fail(e, "lixel statistics");
    return MLGPU_OK;
}

void mlgpu_lixel_filler_destroy(mlgpu_lixel_filler* f) {
    if (!f) return;
    (void)hipDeviceSynchronize();
    (void)hipFree(f->mem);
    delete f;
}

}  // extern "C"
