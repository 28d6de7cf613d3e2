// HIP kernels of the photon propagator for CDNA4 (gfx950).
//
// One lane per photon.  Per bounce:
//   1. traversal: FP32 slab tests over the binary box tree of the current
//      root scope, two child boxes per 64-byte node, nearer child first,
//      farther one on a per-lane stack in LDS, pruned by the best hit so far
//      (merlict/RayAndFrame.cpp:137-149 visits every candidate; the first
//      minimum is the same);
//   2. leaf tests: FP64 in the object frame in merlict's operation order
//      (scenery/geometry/*.cpp), so hits and distances are bit-identical;
//   3. physics: medium absorption, surface reflection, Fresnel, refraction
//      and root-frame switching exactly as merlict/PhotonAndFrame.cpp:23-124,
//      with one Philox4x32-10 stream per photon.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "kernels.hpp"
#include "propagate_core.hpp"

namespace mlgpu {
namespace {

constexpr int kBlock = 256;

// Philox stream id of photon i of a batch: id_offset + i, or id_offset +
// ids[i] when the batch was gathered from a larger one (ids: the photons'
// indices there), so a photon draws the same numbers wherever it is
// propagated.
__device__ __forceinline__ uint64_t photon_id(const uint64_t id_offset, const int32_t* __restrict__ ids,
                                              const int64_t i) {
    return id_offset + static_cast<uint64_t>(ids ? static_cast<int64_t>(ids[i]) : i);
}

// The state of a photon that propagates on, for the next level.
__device__ __forceinline__ void store_state(const Workspace& ws, const int64_t i, const PhotonState& st,
                                            const uint32_t draws, const double wavelength) {
    PhotonRay pr;
    pr.sx = st.s.x; pr.sy = st.s.y; pr.sz = st.s.z;
    pr.dx = st.d.x; pr.dy = st.d.y; pr.dz = st.d.z;
    pr.wavelength = wavelength;
    pr.root = st.root;
    pr.hist_draws = (st.n_hist & 0xFFu) | draws << 8;
    ws.ray[i] = pr;
    ws.sums[i] = PhotonSums{st.tof, st.dist};
}

// The state of a photon from its records, with its draw count.
__device__ __forceinline__ void load_state(const PhotonRay& pr, const PhotonSums& ps, PhotonState& st,
                                           uint32_t& draws) {
    st.s = D3{pr.sx, pr.sy, pr.sz};
    st.d = D3{pr.dx, pr.dy, pr.dz};
    st.tof = ps.tof;
    st.dist = ps.dist;
    st.root = pr.root;
    st.n_hist = pr.hist_draws & 0xFFu;
    st.err = MLGPU_PHOTON_OK;
    draws = pr.hist_draws >> 8;
}

// The per-lane traversal stacks of a block: column threadIdx.x.
#define MLGPU_LANE_STACK(name)                                          \
    __shared__ int32_t name##_mem[kMaxTreeDepth * kBlock];              \
    const LaneStack name { name##_mem + threadIdx.x, kBlock }

// The shading kernels copy the scene's function tables (heads, points,
// buckets) into LDS when they fit, so the three dependent reads of each
// Func1 lookup are LDS reads instead of round trips to the L2; the FACT
// telescope's tables take 9 KB.  Block-uniform: every thread calls it once,
// before any other barrier.
constexpr int kLdsFunctionBytes = 12 * 1024;

__device__ __forceinline__ DeviceScene stage_functions(const DeviceScene& sc, unsigned char* lds) {
    const int hb = sc.num_funcs * static_cast<int>(sizeof(FuncHead));
    const int pb = sc.num_points * static_cast<int>(sizeof(FuncPoint));
    const int bb = sc.num_buckets * static_cast<int>(sizeof(int32_t));
    if (hb + pb + bb > kLdsFunctionBytes) return sc;
    static_assert(sizeof(FuncHead) % 16 == 0 && sizeof(FuncPoint) % 16 == 0, "16-byte copies");
    const uint4* fh = reinterpret_cast<const uint4*>(sc.funcs);
    const uint4* fp = reinterpret_cast<const uint4*>(sc.points);
    uint4* lh = reinterpret_cast<uint4*>(lds);
    uint4* lp = reinterpret_cast<uint4*>(lds + hb);
    int32_t* lb = reinterpret_cast<int32_t*>(lds + hb + pb);
    for (int k = static_cast<int>(threadIdx.x); k < hb / 16; k += kBlock) lh[k] = fh[k];
    for (int k = static_cast<int>(threadIdx.x); k < pb / 16; k += kBlock) lp[k] = fp[k];
    for (int k = static_cast<int>(threadIdx.x); k < sc.num_buckets; k += kBlock) lb[k] = sc.buckets[k];
    __syncthreads();
    DeviceScene l = sc;
    l.funcs = reinterpret_cast<const FuncHead*>(lds);
    l.points = reinterpret_cast<const FuncPoint*>(lds + hb);
    l.buckets = lb;
    return l;
}

// The top of the world scope's tree (the scene's first hot_count nodes)
// copied into LDS for the level-0 kernel, whose traces (source to the first
// surface) stay in it: reading it from LDS spares the vector memory
// pipeline, which the per-lane node gathers keep busy.  Block-uniform, like
// stage_functions.
__device__ __forceinline__ DeviceScene stage_hot(const DeviceScene& sc) {
    const int m = sc.hot_count < kHotNodes ? sc.hot_count : kHotNodes;
    static_assert(sizeof(PairNode) == 32, "two 16-byte copies per node");
    const uint4* src = reinterpret_cast<const uint4*>(sc.pairs);
    uint4* dst = reinterpret_cast<uint4*>(hot_nodes);
    for (int k = static_cast<int>(threadIdx.x); k < 2 * m; k += kBlock) dst[k] = src[k];
    __syncthreads();
    DeviceScene l = sc;
    l.num_hot = m;
    return l;
}

#define MLGPU_STAGE_HOT(sc, sc_in) const DeviceScene sc = stage_hot(sc_in)

#define MLGPU_STAGE_FUNCTIONS(sc, sc_arg)                                   \
    __shared__ __align__(16) unsigned char sc##_fn_lds[kLdsFunctionBytes]; \
    const DeviceScene sc = stage_functions(sc_arg, sc##_fn_lds)

__global__ __launch_bounds__(kBlock) void propagate_kernel(const DeviceScene sc_arg, const int32_t world_scope,
                                                           const mlgpu_photons in, const int64_t n,
                                                           const uint32_t max_interactions, const uint64_t seed,
                                                           const uint64_t id_offset, const mlgpu_results out,
                                                           const int32_t* __restrict__ ids) {
    MLGPU_STAGE_FUNCTIONS(sc, sc_arg);
    MLGPU_LANE_STACK(stk);
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    propagate_photon(sc, world_scope, in, i, max_interactions, seed, photon_id(id_offset, ids, i), out, stk);
}

// The fused schedule with every photon's history recorded (one lane per
// photon; the entries of photon i from i * capacity on).
__global__ __launch_bounds__(kBlock) void propagate_history_kernel(const DeviceScene sc_arg, const int32_t world_scope,
                                                                   const mlgpu_photons in, const int64_t n,
                                                                   const uint32_t max_interactions,
                                                                   const uint64_t seed, const uint64_t id_offset,
                                                                   const mlgpu_results out,
                                                                   const mlgpu_history hist) {
    MLGPU_STAGE_FUNCTIONS(sc, sc_arg);
    MLGPU_LANE_STACK(stk);
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    propagate_photon_history(sc, world_scope, in, i, max_interactions, seed, photon_id(id_offset, nullptr, i), out,
                             stk, hist.bounces + i * hist.capacity, hist.num_bounces + i,
                             hist.final_ray ? hist.final_ray + 6 * i : nullptr);
}

// ------------------------------------------------------- wavefront schedule
// Per bounce level: trace the rays of the live photons (traversal + FP64 leaf
// solves, hit record (leaf, t) per queue slot), then shade them (physics);
// photons that propagate on are appended to the next level's queue, so every
// level runs on a dense queue of live photons and waves stay full however
// many bounces individual photons make.
//
// Appends are staged per block: a block works on tiles of kTile consecutive
// slots, its lanes add indices to an LDS buffer (one LDS atomic per wave),
// and the buffer goes to the global queue with one global atomic per tile,
// in coalesced stores.  A global atomic per wave instead serialises about
// 1.5e5 atomics on one counter for 1e7 photons, which costs milliseconds.

constexpr int kTile = 8 * kBlock;
constexpr int kShadeTile = 8 * kBlock;  // 2, 4 and 8 blocks' worth measured alike

template <int Cap>
struct TileQueue {
    int32_t buf[Cap];
    int32_t n;
    int32_t flush_n;
    int32_t flush_base;
};

template <int Cap>
__device__ __forceinline__ void tile_push(TileQueue<Cap>& q, const bool want, const int32_t value) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return;
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&q.n, __popcll(mask));
    base = __shfl(base, leader);
    if (want) q.buf[base + __popcll(mask & ((1ull << lane) - 1ull))] = value;
}

// Block-uniform: every thread of the block calls it once per tile.
template <int Cap>
__device__ __forceinline__ void tile_flush(TileQueue<Cap>& q, int32_t* queue, int32_t* count) {
    __syncthreads();  // all pushes of this tile are in q
    if (threadIdx.x == 0) {
        const int32_t m = q.n;
        q.flush_n = m;
        q.flush_base = m > 0 ? atomicAdd(count, m) : 0;
        q.n = 0;
    }
    __syncthreads();
    const int32_t m = q.flush_n, base = q.flush_base;
    for (int32_t j = static_cast<int32_t>(threadIdx.x); j < m; j += kBlock) queue[base + j] = q.buf[j];
    __syncthreads();  // q.buf is free for the next tile
}

__global__ void wf_reset_kernel(int32_t* counts, int n) {
    for (int i = static_cast<int>(threadIdx.x); i < n; i += static_cast<int>(blockDim.x)) counts[i] = 0;
}

// Level 0 in one pass: every photon is read from the input, traced from its
// production point and shaded; the photons that propagate on are written to
// the workspace and queued for level 1.  Level-0 traces are short (the
// first surface in front of the source), so they need neither the budget
// nor the hit records of the later levels.
__global__ __launch_bounds__(kBlock) void wf_first_kernel(const DeviceScene sc_arg, const int32_t world_scope,
                                                          const mlgpu_photons in, const int64_t n_arg,
                                                          const uint32_t max_interactions, const uint64_t seed_arg,
                                                          const uint64_t id_offset_arg, const Workspace ws,
                                                          const mlgpu_results out, const CallParams* __restrict__ cp,
                                                          const int32_t* __restrict__ ids) {
    const int64_t n = cp ? cp->n : n_arg;
    const uint64_t seed = cp ? cp->seed : seed_arg;
    const uint64_t id_offset = cp ? cp->id_offset : id_offset_arg;
    MLGPU_STAGE_FUNCTIONS(scf, sc_arg);
    MLGPU_STAGE_HOT(sc, scf);
    MLGPU_LANE_STACK(stk);
    __shared__ TileQueue<kTile> q;
    if (threadIdx.x == 0) q.n = 0;
    __syncthreads();
    for (int64_t tile = static_cast<int64_t>(blockIdx.x) * kTile; tile < n;
         tile += static_cast<int64_t>(gridDim.x) * kTile) {
        #pragma unroll 1
        for (int k = 0; k < kTile; k += kBlock) {
            const int64_t i = tile + k + threadIdx.x;
            bool go = false;
            if (i < n) {
                PhotonState st;
                photon_init(st, in, i, world_scope);
                Rng rng;
                rng.init(seed, photon_id(id_offset, ids, i), 0u);
                if (st.err == MLGPU_PHOTON_OK && st.n_hist <= max_interactions) {
                    Closest c;
                    first_hit<true>(sc, st.root, st.s, st.d, c, stk);
                    go = photon_bounce(sc, st, c.leaf, c.t, in.wavelength[i], rng, max_interactions);
                }
                if (go) {
                    store_state(ws, i, st, rng.next, in.wavelength[i]);
                } else {
                    photon_finish(st, out, i);
                }
            }
            tile_push(q, go, static_cast<int32_t>(i));
        }
        tile_flush(q, ws.queue[1], ws.counts + 1);
    }
}

// First pass of a level's traces: every query gets `budget` steps, so no
// wave waits long on a few costly rays; the unfinished ones keep their
// traversal state (tr_depth >= 0 marks them) and go on in wf_resume_kernel,
// packed into full waves by wf_defer_kernel.  No barriers here: a block
// barrier would make every wave wait for the slowest of its block.
__global__ __launch_bounds__(kBlock, 5) void wf_trace_kernel(const DeviceScene sc, const Workspace ws,
                                                             const int level, const int32_t budget) {
    MLGPU_LANE_STACK(stk);
    const int32_t count = ws.counts[level];
    const int32_t* queue = ws.queue[level & 1];
    for (int32_t slot = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x); slot < count;
         slot += static_cast<int32_t>(gridDim.x * blockDim.x)) {
        const int32_t i = queue[slot];
        const PhotonRay pr = ws.ray[i];
        const D3 s{pr.sx, pr.sy, pr.sz};
        const D3 d{pr.dx, pr.dy, pr.dz};
        const int32_t scope = pr.root;
        RayF r;
        ray_setup(sc, scope, s, d, r);
        TraceState ts;
        trace_begin(sc, scope, ts);
        const bool done = trace_run<true>(sc, r, s, d, ts, stk, budget);
        ws.hit_leaf[slot] = ts.best.leaf;
        ws.hit_t[slot] = ts.best.t;
        if (done) {
            ws.tr_depth[slot] = -1;
        } else {
            ws.tr_cur[slot] = ts.cur;
            ws.tr_depth[slot] = ts.depth;
            for (int e = 0; e < ts.depth; e++)
                ws.tr_stack[static_cast<int64_t>(e) * ws.cap + slot] = stk.base[e * stk.stride];
        }
    }
}

// The list of a level's deferred slots.
__global__ __launch_bounds__(kBlock) void wf_defer_kernel(const Workspace ws, const int level) {
    __shared__ TileQueue<kTile> q;
    if (threadIdx.x == 0) q.n = 0;
    __syncthreads();
    const int32_t count = ws.counts[level];
    for (int64_t tile = static_cast<int64_t>(blockIdx.x) * kTile; tile < count;
         tile += static_cast<int64_t>(gridDim.x) * kTile) {
        #pragma unroll 1
        for (int k = 0; k < kTile; k += kBlock) {
            const int32_t slot = static_cast<int32_t>(tile + k + threadIdx.x);
            tile_push(q, slot < count && ws.tr_depth[slot] >= 0, slot);
        }
        tile_flush(q, ws.deferred, ws.counts + kDeferredCounts + level);
    }
}

// Second pass: the deferred traversals of a level, run to the end.
__global__ __launch_bounds__(kBlock) void wf_resume_kernel(const DeviceScene sc, const Workspace ws,
                                                           const int level) {
    MLGPU_LANE_STACK(stk);
    const int32_t count = ws.counts[kDeferredCounts + level];
    const int32_t* queue = ws.queue[level & 1];
    for (int32_t j = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x); j < count;
         j += static_cast<int32_t>(gridDim.x * blockDim.x)) {
        const int32_t slot = ws.deferred[j];
        const int32_t i = queue[slot];
        const PhotonRay pr = ws.ray[i];
        const D3 s{pr.sx, pr.sy, pr.sz};
        const D3 d{pr.dx, pr.dy, pr.dz};
        RayF r;
        ray_setup(sc, pr.root, s, d, r);
        TraceState ts;
        ts.cur = ws.tr_cur[slot];
        ts.depth = ws.tr_depth[slot];
        ts.best.leaf = ws.hit_leaf[slot];
        ts.best.t = ws.hit_t[slot];
        for (int e = 0; e < ts.depth; e++)
            stk.base[e * stk.stride] = ws.tr_stack[static_cast<int64_t>(e) * ws.cap + slot];
        trace_run<false>(sc, r, s, d, ts, stk, 0);
        ws.hit_leaf[slot] = ts.best.leaf;
        ws.hit_t[slot] = ts.best.t;
    }
}

__global__ __launch_bounds__(kBlock) void wf_shade_kernel(const DeviceScene sc_arg, const mlgpu_photons in,
                                                          const Workspace ws, const int level,
                                                          const uint32_t max_interactions, const uint64_t seed_arg,
                                                          const uint64_t id_offset_arg, const mlgpu_results out,
                                                          const CallParams* __restrict__ cp,
                                                          const int32_t* __restrict__ ids) {
    const uint64_t seed = cp ? cp->seed : seed_arg;
    const uint64_t id_offset = cp ? cp->id_offset : id_offset_arg;
    MLGPU_STAGE_FUNCTIONS(sc, sc_arg);
    __shared__ TileQueue<kShadeTile> q;
    if (threadIdx.x == 0) q.n = 0;
    __syncthreads();
    const int32_t count = ws.counts[level];
    const int32_t* queue = ws.queue[level & 1];
    int32_t* next = ws.queue[(level + 1) & 1];
    for (int64_t tile = static_cast<int64_t>(blockIdx.x) * kShadeTile; tile < count;
         tile += static_cast<int64_t>(gridDim.x) * kShadeTile) {
        #pragma unroll 1
        for (int k = 0; k < kShadeTile; k += kBlock) {
            const int64_t slot = tile + k + threadIdx.x;
            bool go = false;
            int32_t i = 0;
            if (slot < count) {
                i = queue[slot];
                const PhotonRay pr = ws.ray[i];
                PhotonState st;
                uint32_t draws;
                load_state(pr, ws.sums[i], st, draws);
                Rng rng;
                rng.init(seed, photon_id(id_offset, ids, i), draws);
                go = photon_bounce(sc, st, ws.hit_leaf[slot], ws.hit_t[slot], pr.wavelength, rng, max_interactions);
                if (go) {
                    store_state(ws, i, st, rng.next, pr.wavelength);
                } else {
                    photon_finish(st, out, i);
                }
            }
            tile_push(q, go, i);
        }
        tile_flush(q, next, ws.counts + level + 1);
    }
}

// The last stage of a pass: the photons of a level are shaded with their
// closest hits and then propagated to the end one lane per photon, trace and
// bounce in turn, with no further round trip of their state through the
// workspace.  After the mirror-to-camera level the remaining bounces are
// short traces inside one light guide's scope, so finishing them in the lane
// costs little divergence while every level it saves is one gathered read
// and one write of the photon's state.
__global__ __launch_bounds__(kBlock) void wf_tail_kernel(const DeviceScene sc_arg, const mlgpu_photons in,
                                                         const Workspace ws, const int level,
                                                         const uint32_t max_interactions, const uint64_t seed_arg,
                                                         const uint64_t id_offset_arg, const mlgpu_results out,
                                                         const CallParams* __restrict__ cp,
                                                         const int32_t* __restrict__ ids) {
    MLGPU_STAGE_FUNCTIONS(sc, sc_arg);
    MLGPU_LANE_STACK(stk);
    const uint64_t seed = cp ? cp->seed : seed_arg;
    const uint64_t id_offset = cp ? cp->id_offset : id_offset_arg;
    const int32_t count = ws.counts[level];
    const int32_t* queue = ws.queue[level & 1];
    for (int32_t slot = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x); slot < count;
         slot += static_cast<int32_t>(gridDim.x * blockDim.x)) {
        const int32_t i = queue[slot];
        const PhotonRay pr = ws.ray[i];
        PhotonState st;
        uint32_t draws;
        load_state(pr, ws.sums[i], st, draws);
        const double lambda = pr.wavelength;
        Rng rng;
        rng.init(seed, photon_id(id_offset, ids, i), draws);
        bool go = photon_bounce(sc, st, ws.hit_leaf[slot], ws.hit_t[slot], lambda, rng, max_interactions);
        while (go) {
            Closest c;
            first_hit(sc, st.root, st.s, st.d, c, stk);
            go = photon_bounce(sc, st, c.leaf, c.t, lambda, rng, max_interactions);
        }
        photon_finish(st, out, i);
    }
}

// ----------------------------------------------------- persistent schedule
// One lane per photon at a time, for all of its bounces, with the photon's
// state in registers from its input to its result, so that the only photon
// traffic is the input read and the result write.  A lane whose photon has
// ended takes the next photon of its wave's chunk (kChunk consecutive
// photons that a wave claims with one global atomic), so waves stay full
// until the last chunks.  Traces advance in quanta of `quantum` traversal
// steps; a lane whose trace has ended waits until `shade_lanes` lanes of its
// wave (or every busy one) have, and then they run the bounce physics
// together: the FP64 shading runs on filled waves while long traces (mirror
// to camera, world scope) and short ones (inside one light guide) share a
// wave.

constexpr int kChunk = 256;

__global__ __launch_bounds__(kBlock) void persistent_kernel(const DeviceScene sc_arg, const int32_t world_scope,
                                                            const mlgpu_photons in, const int64_t n_arg,
                                                            const uint32_t max_interactions, const uint64_t seed_arg,
                                                            const uint64_t id_offset_arg, const mlgpu_results out,
                                                            unsigned long long* __restrict__ counter,
                                                            const int32_t quantum, const int32_t shade_lanes,
                                                            const CallParams* __restrict__ cp,
                                                            const int32_t* __restrict__ ids) {
    const int64_t n = cp ? cp->n : n_arg;
    const uint64_t seed = cp ? cp->seed : seed_arg;
    const uint64_t id_offset = cp ? cp->id_offset : id_offset_arg;
    MLGPU_STAGE_FUNCTIONS(sc, sc_arg);
    MLGPU_LANE_STACK(stk);
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const uint64_t below = (1ull << lane) - 1ull;
    // the wave's chunk [chunk_next, chunk_end); wave-uniform
    int64_t chunk_next = 0, chunk_end = 0;
    bool exhausted = false;
    // this lane's photon
    bool busy = false, traced = false;
    int64_t i = 0;
    double lambda = 0.0;
    PhotonState st;
    Rng rng;
    TraceState ts;
    for (;;) {
        // idle lanes take photons until every lane is busy or no photon is left
        for (;;) {
            const uint64_t idle = __ballot(!busy);
            if (idle == 0 || exhausted) break;
            if (chunk_next >= chunk_end) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(counter, static_cast<unsigned long long>(kChunk));
                base = __shfl(base, 0);
                if (static_cast<int64_t>(base) >= n) {
                    exhausted = true;
                    break;
                }
                chunk_next = static_cast<int64_t>(base);
                chunk_end = chunk_next + kChunk < n ? chunk_next + kChunk : n;
            }
            const int64_t want = __popcll(idle);
            const int64_t take = want < chunk_end - chunk_next ? want : chunk_end - chunk_next;
            const int64_t rank = __popcll(idle & below);
            if (!busy && rank < take) {
                i = chunk_next + rank;
                photon_init(st, in, i, world_scope);
                lambda = in.wavelength[i];
                rng.init(seed, photon_id(id_offset, ids, i), 0u);
                if (st.err == MLGPU_PHOTON_OK && st.n_hist <= max_interactions) {
                    trace_begin(sc, st.root, ts);
                    busy = true;
                    traced = false;
                } else {
                    photon_finish(st, out, i);
                }
            }
            chunk_next += take;
        }
        const uint64_t live = __ballot(busy);
        if (live == 0) break;  // only once the photons are exhausted
        const int live_n = __popcll(live);
        const int need = shade_lanes < live_n ? shade_lanes : live_n;
        {
            // the float ray is rebuilt from the photon's state in every round
            // rather than kept: it would otherwise hold registers through the
            // shading of the other lanes
            RayF r;
            if (busy && !traced) ray_setup(sc, st.root, st.s, st.d, r);
            for (;;) {
                if (busy && !traced) traced = trace_run<true>(sc, r, st.s, st.d, ts, stk, quantum);
                if (__popcll(__ballot(busy && traced)) >= need) break;
            }
        }
        if (busy && traced) {
            traced = false;
            if (photon_bounce(sc, st, ts.best.leaf, ts.best.t, lambda, rng, max_interactions)) {
                trace_begin(sc, st.root, ts);
            } else {
                photon_finish(st, out, i);
                busy = false;
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void first_hit_kernel(const DeviceScene sc, const int32_t root_scope,
                                                           const mlgpu_photons rays, const int64_t n,
                                                           const mlgpu_hits out) {
    MLGPU_LANE_STACK(stk);
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const D3 s{rays.support_x[i], rays.support_y[i], rays.support_z[i]};
    const D3 d = normalized(D3{rays.direction_x[i], rays.direction_y[i], rays.direction_z[i]});
    Closest c;
    first_hit(sc, root_scope, s, d, c, stk);
    if (c.leaf < 0) {
        const D3 e = at(s, d, kVoidDistance);
        out.frame[i] = -1;
        out.distance[i] = kVoidDistance;
        out.position_x[i] = e.x;
        out.position_y[i] = e.y;
        out.position_z[i] = e.z;
        return;
    }
    const LeafGeo& g = sc.geo[c.leaf];
    const D3 w = position(g, at(position_inv(g, s), orient_inv(g, d), c.t));
    out.frame[i] = sc.phys[c.leaf].frame;
    out.distance[i] = c.t;
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

// The wavefront workspace of mlgpu_propagate_noalloc (workspace_bytes of
// kNoAllocMaxPhotons, about 58 MB): every GPU the library runs on holds its
// own copy of this array until the process ends, and every call on that GPU
// shares it.  Host code never uses this name as an address: it takes the
// address hipGetSymbolAddress gives for the current device.
static __device__ __align__(256) unsigned char scratch[workspace_bytes(kNoAllocMaxPhotons)];

hipError_t noalloc_workspace(void** ptr) {
    constexpr int kMaxDevices = 64;
    static std::mutex mutex;
    static void* address[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mutex);
    if (!address[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        address[dev] = p;
    }
    *ptr = address[dev];
    return hipSuccess;
}

hipError_t launch_propagate(const DeviceScene& sc, int32_t world_scope, const mlgpu_photons& in, int64_t n,
                            uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                            const mlgpu_results& out, hipStream_t stream, const int32_t* ids) {
    if (n <= 0) return hipSuccess;
    const int block = kBlock;
    const int64_t grid = (n + block - 1) / block;
    hipLaunchKernelGGL(propagate_kernel, dim3(static_cast<uint32_t>(grid)), dim3(block), 0, stream, sc, world_scope,
                       in, n, max_interactions, seed, id_offset, out, ids);
    return hipGetLastError();
}

hipError_t launch_propagate_history(const DeviceScene& sc, int32_t world_scope, const mlgpu_photons& in, int64_t n,
                                    uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                                    const mlgpu_results& out, const mlgpu_history& hist, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t grid = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(propagate_history_kernel, dim3(static_cast<uint32_t>(grid)), dim3(kBlock), 0, stream, sc,
                       world_scope, in, n, max_interactions, seed, id_offset, out, hist);
    return hipGetLastError();
}

hipError_t launch_propagate_wavefront(const DeviceScene& sc, int32_t world_scope, const mlgpu_photons& in,
                                      int64_t n, uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                                      const mlgpu_results& out, const Workspace& ws, int grid_blocks,
                                      int32_t trace_budget, int32_t tail_level, hipStream_t stream,
                                      const CallParams* cp, const int32_t* ids) {
    if (n <= 0) return hipSuccess;
    const int levels = static_cast<int>(max_interactions) + 1;
    if (levels + 1 > kDeferredCounts) return hipErrorInvalidValue;
    hipLaunchKernelGGL(wf_reset_kernel, dim3(1), dim3(256), 0, stream, ws.counts, kNumCounts);
    const int64_t tiles = (n + kTile - 1) / kTile;
    const int tile_grid = static_cast<int>(tiles < grid_blocks ? tiles : grid_blocks);
    const int64_t shade_tiles = (n + kShadeTile - 1) / kShadeTile;
    const int shade_grid = static_cast<int>(shade_tiles < grid_blocks ? shade_tiles : grid_blocks);
    const int64_t blocks = (n + kBlock - 1) / kBlock;
    const int grid = static_cast<int>(blocks < grid_blocks ? blocks : grid_blocks);
    hipLaunchKernelGGL(wf_first_kernel, dim3(tile_grid), dim3(kBlock), 0, stream, sc, world_scope, in, n,
                       max_interactions, seed, id_offset, ws, out, cp, ids);
    for (int level = 1; level < static_cast<int>(max_interactions); level++) {
        const int32_t budget = trace_budget > 0 ? trace_budget : INT32_MAX;
        hipLaunchKernelGGL(wf_trace_kernel, dim3(grid), dim3(kBlock), 0, stream, sc, ws, level, budget);
        if (trace_budget > 0) {
            hipLaunchKernelGGL(wf_defer_kernel, dim3(tile_grid), dim3(kBlock), 0, stream, ws, level);
            hipLaunchKernelGGL(wf_resume_kernel, dim3(grid), dim3(kBlock), 0, stream, sc, ws, level);
        }
        if (tail_level > 0 && level >= tail_level) {
            hipLaunchKernelGGL(wf_tail_kernel, dim3(grid), dim3(kBlock), 0, stream, sc, in, ws, level,
                               max_interactions, seed, id_offset, out, cp, ids);
            break;
        }
        hipLaunchKernelGGL(wf_shade_kernel, dim3(shade_grid), dim3(kBlock), 0, stream, sc, in, ws, level,
                           max_interactions, seed, id_offset, out, cp, ids);
    }
    return hipGetLastError();
}

hipError_t launch_propagate_persistent(const DeviceScene& sc, int32_t world_scope, const mlgpu_photons& in,
                                       int64_t n, uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                                       const mlgpu_results& out, int32_t* counts, int grid_blocks, int32_t quantum,
                                       int32_t shade_lanes, hipStream_t stream, const CallParams* cp,
                                       const int32_t* ids) {
    if (n <= 0) return hipSuccess;
    // the chunk counter: one 64-bit word at the end of the counts, cleared
    // by a kernel so that the sequence can be captured in a graph
    hipLaunchKernelGGL(wf_reset_kernel, dim3(1), dim3(256), 0, stream, counts, kNumCounts);
    auto* counter = reinterpret_cast<unsigned long long*>(counts + kNumCounts - 2);
    const int64_t waves_wanted = (n + kChunk - 1) / kChunk;
    const int64_t blocks = (waves_wanted + kBlock / 64 - 1) / (kBlock / 64);
    const int grid = static_cast<int>(blocks < grid_blocks ? blocks : grid_blocks);
    const int32_t q = quantum > 0 ? quantum : INT32_MAX;
    const int32_t lanes = shade_lanes < 1 ? 1 : (shade_lanes > 64 ? 64 : shade_lanes);
    hipLaunchKernelGGL(persistent_kernel, dim3(grid), dim3(kBlock), 0, stream, sc, world_scope, in, n,
                       max_interactions, seed, id_offset, out, counter, q, lanes, cp, ids);
    return hipGetLastError();
}

hipError_t launch_first_hit(const DeviceScene& sc, int32_t root_scope, const mlgpu_photons& rays, int64_t n,
                            const mlgpu_hits& out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int block = kBlock;
    const int64_t grid = (n + block - 1) / block;
    hipLaunchKernelGGL(first_hit_kernel, dim3(static_cast<uint32_t>(grid)), dim3(block), 0, stream, sc, root_scope,
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
