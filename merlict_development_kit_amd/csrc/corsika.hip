// CORSIKA photon bunches to photons on the GPU: merlict::EventIoPhotonFactory
// (merlict_corsika/PhotonFactory.cpp:8-148) over a whole event at once.
//
// Three passes over the bunch array (n x 8 float: x, y [cm], cx, cy,
// time [ns], emission height [cm], weight, wavelength [nm]):
//   1. corsika_count_kernel: per bunch the weight check (:138-148) and the
//      collapse of the weight to a photon count (:8-24; floor(w), one more
//      when a uniform from Philox stream 5 of (seed, id) is <= fmod(w, 1));
//      a block-local exclusive scan of the counts;
//   2. a device-wide exclusive scan (rocPRIM) of the block totals;
//   3. corsika_emit_kernel: each bunch writes its photons at its offset:
//      support, direction and wavelength computed in FP64 in the
//      reference's operation order (:30-136, Ray.cpp:16-20, Vec3.cpp:20-26).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>

#include "kernels.hpp"
#include "propagate_core.hpp"

namespace mlgpu {
namespace {

constexpr uint32_t kCorsikaStream = 5;
constexpr int kScanBlock = 256;

struct Bunch8 {
    float v[8];
};

__device__ __forceinline__ Bunch8 load_bunch(const float* __restrict__ b, const int64_t i) {
    Bunch8 r;
    const float4 lo = *reinterpret_cast<const float4*>(b + 8 * i);
    const float4 hi = *reinterpret_cast<const float4*>(b + 8 * i + 4);
    r.v[0] = lo.x; r.v[1] = lo.y; r.v[2] = lo.z; r.v[3] = lo.w;
    r.v[4] = hi.x; r.v[5] = hi.y; r.v[6] = hi.z; r.v[7] = hi.w;
    return r;
}

// PhotonFactory.cpp:138-148: weights outside [0, 16] are rejected
__device__ __forceinline__ bool weight_ok(const Bunch8& c) {
    const double weight = c.v[6];
    return !(weight < 0.0 || weight > 16.0);
}

// Photons of a bunch with an accepted weight (PhotonFactory.cpp:8-24).
__device__ __forceinline__ uint32_t bunch_count(const Bunch8& c, const uint64_t seed, const uint64_t id) {
    const double weight = c.v[6];
    uint32_t num = static_cast<uint32_t>(floor(weight));
    const double rest = fmod(weight, 1.0);
    uint32_t w0, w1, w2, w3;
    philox4x32_10(0u, kCorsikaStream, static_cast<uint32_t>(id), static_cast<uint32_t>(id >> 32),
                  static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32), w0, w1, w2, w3);
    if (static_cast<double>(w0 >> 8) * (1.0 / 16777216.0) <= rest) num += 1u;
    return num;
}

// inclusive scan of 256 values in LDS (Hillis-Steele); returns the total
__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t* s, const int t) {
    for (int off = 1; off < kScanBlock; off <<= 1) {
        const uint32_t v = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    return s[kScanBlock - 1];
}

__global__ __launch_bounds__(kScanBlock) void corsika_count_kernel(const float* __restrict__ bunches,
                                                                   const int64_t n, const uint64_t seed,
                                                                   const uint64_t id_offset,
                                                                   uint32_t* __restrict__ local,
                                                                   uint64_t* __restrict__ block_total,
                                                                   int32_t* __restrict__ bad) {
    __shared__ uint32_t s[kScanBlock];
    const int t = static_cast<int>(threadIdx.x);
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kScanBlock + t;
    uint32_t c = 0u;
    if (i < n) {
        const Bunch8 b = load_bunch(bunches, i);
        if (weight_ok(b))
            c = bunch_count(b, seed, id_offset + static_cast<uint64_t>(i));
        else
            atomicOr(bad, 1);
    }
    s[t] = c;
    __syncthreads();
    const uint32_t total = block_inclusive_scan(s, t);
    if (i < n) local[i] = s[t] - c;  // exclusive offset in the block
    if (t == 0) block_total[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void corsika_emit_kernel(const float* __restrict__ bunches, const int64_t n,
                                                           const uint64_t seed, const uint64_t id_offset,
                                                           const uint32_t* __restrict__ local,
                                                           const uint64_t* __restrict__ block_offset,
                                                           const int64_t capacity, const mlgpu_photon_buffers out,
                                                           int32_t* __restrict__ truth) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Bunch8 c = load_bunch(bunches, i);
    if (!weight_ok(c)) return;
    const uint32_t num = bunch_count(c, seed, id_offset + static_cast<uint64_t>(i));
    if (num == 0u) return;
    const int64_t at = static_cast<int64_t>(block_offset[i / kScanBlock]) + local[i];
    // direction_of_motion (PhotonFactory.cpp:72-124)
    const double cx = c.v[2], cy = c.v[3];
    const double d0 = cx, d1 = cy, d2 = -sqrt(1.0 - cx * cx - cy * cy);
    // the ray from the ground point up to the production point (:30-45):
    // direction -d, normalised by Ray (Ray.cpp:16-20)
    double u0 = d0 * -1.0, u1 = d1 * -1.0, u2 = d2 * -1.0;
    const double un = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
    u0 = u0 / un;
    u1 = u1 / un;
    u2 = u2 / un;
    const double gx = c.v[0] * 1e-2, gy = c.v[1] * 1e-2;
    const double s = c.v[4] * 1e-9 * 299792458.0 + 1e3;
    const double sx = gx + u0 * s, sy = gy + u1 * s, sz = 0.0 + u2 * s;
    // Photon normalises its direction
    const double dn = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double e0 = d0 / dn, e1 = d1 / dn, e2 = d2 / dn;
    const double lambda = fabs(c.v[7] * 1e-9);
    for (uint32_t j = 0; j < num; j++) {
        const int64_t k = at + j;
        if (k >= capacity) return;
        out.support_x[k] = sx;
        out.support_y[k] = sy;
        out.support_z[k] = sz;
        out.direction_x[k] = e0;
        out.direction_y[k] = e1;
        out.direction_z[k] = e2;
        out.wavelength[k] = lambda;
        if (truth) truth[k] = static_cast<int32_t>(i);
    }
}

// Workspace: per bunch its offset in its block (4 B), per block its total
// (blocks + 1 entries, the last 0) and its exclusive offset, the scan's
// temporary storage.
struct CorsikaWs {
    uint32_t* local;
    uint64_t* block_total;
    uint64_t* block_offset;
    void* temp;
    size_t temp_bytes;
};

size_t ws_align(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

size_t scan_temp_bytes(int64_t blocks) {
    size_t t = 0;
    (void)rocprim::exclusive_scan(nullptr, t, static_cast<uint64_t*>(nullptr), static_cast<uint64_t*>(nullptr),
                                  uint64_t{0}, static_cast<size_t>(blocks + 1), rocprim::plus<uint64_t>(), 0);
    return t;
}

CorsikaWs carve(void* ws, int64_t n) {
    const int64_t blocks = (n + kScanBlock - 1) / kScanBlock;
    char* q = static_cast<char*>(ws);
    CorsikaWs w;
    w.local = reinterpret_cast<uint32_t*>(q);
    q += ws_align(static_cast<size_t>(n) * 4);
    w.block_total = reinterpret_cast<uint64_t*>(q);
    q += ws_align(static_cast<size_t>(blocks + 1) * 8);
    w.block_offset = reinterpret_cast<uint64_t*>(q);
    q += ws_align(static_cast<size_t>(blocks + 1) * 8);
    w.temp = q;
    w.temp_bytes = scan_temp_bytes(blocks);
    return w;
}

}  // namespace

size_t corsika_workspace_bytes(int64_t n) {
    const int64_t blocks = (n + kScanBlock - 1) / kScanBlock;
    return ws_align(static_cast<size_t>(n) * 4) + 2 * ws_align(static_cast<size_t>(blocks + 1) * 8) +
           ws_align(scan_temp_bytes(blocks));
}

hipError_t launch_corsika_count(const float* bunches, int64_t n, uint64_t seed, uint64_t id_offset, void* ws,
                                int32_t* bad, int64_t* total, hipStream_t stream) {
    const int64_t blocks = (n + kScanBlock - 1) / kScanBlock;
    CorsikaWs w = carve(ws, n);
    hipLaunchKernelGGL(corsika_count_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(kScanBlock), 0, stream,
                       bunches, n, seed, id_offset, w.local, w.block_total, bad);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemsetAsync(w.block_total + blocks, 0, 8, stream);
    // exclusive offsets of the blocks; entry `blocks` is the grand total
    if (e == hipSuccess)
        e = rocprim::exclusive_scan(w.temp, w.temp_bytes, w.block_total, w.block_offset, uint64_t{0},
                                    static_cast<size_t>(blocks + 1), rocprim::plus<uint64_t>(), stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(total, w.block_offset + blocks, 8, hipMemcpyDeviceToDevice, stream);
    return e;
}

hipError_t launch_corsika_emit(const float* bunches, int64_t n, uint64_t seed, uint64_t id_offset, const void* ws,
                               int64_t capacity, const mlgpu_photon_buffers& out, int32_t* truth,
                               hipStream_t stream) {
    const CorsikaWs w = carve(const_cast<void*>(ws), n);
    hipLaunchKernelGGL(corsika_emit_kernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, stream,
                       bunches, n, seed, id_offset, w.local, w.block_offset, capacity, out, truth);
    return hipGetLastError();
}

}  // namespace mlgpu
