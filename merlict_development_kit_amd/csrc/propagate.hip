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

// ------------------------------------------------------- wavefront schedule
// Per bounce level: trace the rays of the live photons (traversal + FP64 leaf
// solves, hit record (leaf, t) per queue slot), then shade them (physics);
// photons that propagate on are appended to the next level's queue, one
// atomic per wave, so every level runs on a dense queue of live photons and
// waves stay full however many bounces individual photons make.

__device__ __forceinline__ void wave_append(int32_t* queue, int32_t* count, const bool want, const int32_t value) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return;
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(mask));
    base = __shfl(base, leader);
    if (want) queue[base + __popcll(mask & ((1ull << lane) - 1ull))] = value;
}

__global__ void wf_reset_kernel(int32_t* counts, int n) {
    const int i = static_cast<int>(threadIdx.x);
    if (i < n) counts[i] = 0;
}

__global__ __launch_bounds__(256) void wf_init_kernel(const int32_t world_node, const mlgpu_photons in,
                                                      const int64_t n, const uint32_t max_interactions,
                                                      const Workspace ws, const mlgpu_results out) {
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < n;
         base += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t i = base + threadIdx.x;
        bool live = false;
        if (i < n) {
            PhotonState st;
            photon_init(st, in, i, world_node);
            live = st.err == MLGPU_PHOTON_OK && st.n_hist <= max_interactions;
            if (live) {
                ws.sx[i] = st.s.x; ws.sy[i] = st.s.y; ws.sz[i] = st.s.z;
                ws.dx[i] = st.d.x; ws.dy[i] = st.d.y; ws.dz[i] = st.d.z;
                ws.tof[i] = 0.0;
                ws.root[i] = st.root;
                ws.n_hist[i] = st.n_hist;
                ws.draws[i] = 0u;
            } else {
                photon_finish(st, out, i);
            }
        }
        wave_append(ws.queue[0], ws.counts, live, static_cast<int32_t>(i));
    }
}

__global__ __launch_bounds__(256) void wf_trace_kernel(const DeviceScene sc, const Workspace ws, const int level) {
    const int32_t count = ws.counts[level];
    const int32_t* queue = ws.queue[level & 1];
    for (int32_t slot = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x); slot < count;
         slot += static_cast<int32_t>(gridDim.x * blockDim.x)) {
        const int32_t i = queue[slot];
        const D3 s{ws.sx[i], ws.sy[i], ws.sz[i]};
        const D3 d{ws.dx[i], ws.dy[i], ws.dz[i]};
        Closest c;
        first_hit(sc, ws.root[i], s, d, c);
        ws.hit_leaf[slot] = c.leaf;
        ws.hit_t[slot] = c.hit.t;
    }
}

__global__ __launch_bounds__(256) void wf_shade_kernel(const DeviceScene sc, const mlgpu_photons in, const Workspace ws,
                                                       const int level, const uint32_t max_interactions,
                                                       const uint64_t seed, const uint64_t id_offset,
                                                       const mlgpu_results out) {
    const int32_t count = ws.counts[level];
    const int32_t* queue = ws.queue[level & 1];
    int32_t* next = ws.queue[(level + 1) & 1];
    for (int32_t base = static_cast<int32_t>(blockIdx.x * blockDim.x); base < count;
         base += static_cast<int32_t>(gridDim.x * blockDim.x)) {
        const int32_t slot = base + static_cast<int32_t>(threadIdx.x);
        bool go = false;
        int32_t i = 0;
        if (slot < count) {
            i = queue[slot];
            PhotonState st;
            st.s = D3{ws.sx[i], ws.sy[i], ws.sz[i]};
            st.d = D3{ws.dx[i], ws.dy[i], ws.dz[i]};
            st.tof = ws.tof[i];
            st.root = ws.root[i];
            st.n_hist = ws.n_hist[i];
            st.err = MLGPU_PHOTON_OK;
            Rng rng;
            rng.init(seed, id_offset + static_cast<uint64_t>(i), ws.draws[i]);
            go = photon_bounce(sc, st, ws.hit_leaf[slot], ws.hit_t[slot], in.wavelength[i], rng, max_interactions);
            if (go) {
                ws.sx[i] = st.s.x; ws.sy[i] = st.s.y; ws.sz[i] = st.s.z;
                ws.dx[i] = st.d.x; ws.dy[i] = st.d.y; ws.dz[i] = st.d.z;
                ws.tof[i] = st.tof;
                ws.root[i] = st.root;
                ws.n_hist[i] = st.n_hist;
                ws.draws[i] = rng.next;
            } else {
                photon_finish(st, out, i);
            }
        }
        wave_append(next, ws.counts + level + 1, go, i);
    }
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

hipError_t launch_propagate_wavefront(const DeviceScene& sc, int32_t world_node, const mlgpu_photons& in,
                                      int64_t n, uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                                      const mlgpu_results& out, const Workspace& ws, int grid_blocks,
                                      hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int block = 256;
    const int levels = static_cast<int>(max_interactions) + 1;
    hipLaunchKernelGGL(wf_reset_kernel, dim3(1), dim3(256), 0, stream, ws.counts, levels + 1);
    const int64_t need = (n + block - 1) / block;
    const int grid = static_cast<int>(need < grid_blocks ? need : grid_blocks);
    hipLaunchKernelGGL(wf_init_kernel, dim3(grid), dim3(block), 0, stream, world_node, in, n, max_interactions, ws,
                       out);
    for (int level = 0; level < static_cast<int>(max_interactions); level++) {
        hipLaunchKernelGGL(wf_trace_kernel, dim3(grid), dim3(block), 0, stream, sc, ws, level);
        hipLaunchKernelGGL(wf_shade_kernel, dim3(grid), dim3(block), 0, stream, sc, in, ws, level, max_interactions,
                           seed, id_offset, out);
    }
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
