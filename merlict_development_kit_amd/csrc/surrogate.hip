// HIP kernels of the learned surrogate (surrogate.hpp): a bf16 matrix
// kernel on the CDNA4 matrix cores (v_mfma_f32_32x32x16_bf16, fp32
// accumulation) with the epilogues the network needs, the mixture loss and
// its gradient, Adam, and the packing of weights and features.
#include <hip/hip_runtime.h>

#include <cmath>

#include "arrivals_core.hpp"
#include "sources_core.hpp"
#include "surrogate.hpp"

namespace mlgpu {
namespace sur {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// bf16 <-> fp32, round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<uint16_t>((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// ----------------------------------------------------------- matrix kernel
// Block tile 64 x 64, K step 32 staged through LDS in K-contiguous rows
// (transposed on the way in when an operand is not K-contiguous in memory);
// wave w computes the 32 x 32 sub-tile (w / 2, w % 2) with two MFMAs per
// K step.  Lane l of the 32x32x16 instruction supplies row l % 32 and
// k = 8 (l / 32) .. + 7 of A and of B, and holds C rows
// 8 (v / 4) + 4 (l / 32) + v % 4 of column l % 32 in accumulator v.
constexpr int BM = 64, BN = 64, BK = 32, LDS_STRIDE = BK + 8;

template <bool kAK, bool kBK, int kEpi>
__global__ __launch_bounds__(256) void gemm_kernel(const uint16_t* __restrict__ a, const int lda,
                                                   const uint16_t* __restrict__ b, const int ldb, const int M,
                                                   const int kchunk, const GemmOut out) {
    __shared__ __attribute__((aligned(16))) uint16_t As[BM * LDS_STRIDE];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[BN * LDS_STRIDE];
    const int t = static_cast<int>(threadIdx.x), lane = t & 63, w = t >> 6;
    const int m0 = static_cast<int>(blockIdx.y) * BM, n0 = static_cast<int>(blockIdx.x) * BN;
    const int kbeg = static_cast<int>(blockIdx.z) * kchunk, kend = kbeg + kchunk;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    f32x16 acc;
    for (int v = 0; v < 16; v++) acc[v] = 0.0f;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        if (kAK) {
            const int r = t >> 2, kq = (t & 3) * 8;
            const uint4 v = *reinterpret_cast<const uint4*>(a + static_cast<size_t>(m0 + r) * lda + k0 + kq);
            *reinterpret_cast<uint4*>(As + r * LDS_STRIDE + kq) = v;
        } else {
            const int kk = t >> 3, mq = (t & 7) * 8;
            const uint4 v = *reinterpret_cast<const uint4*>(a + static_cast<size_t>(k0 + kk) * lda + m0 + mq);
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
            for (int j = 0; j < 8; j++) As[(mq + j) * LDS_STRIDE + kk] = e[j];
        }
        if (kBK) {
            const int r = t >> 2, kq = (t & 3) * 8;
            const uint4 v = *reinterpret_cast<const uint4*>(b + static_cast<size_t>(n0 + r) * ldb + k0 + kq);
            *reinterpret_cast<uint4*>(Bs + r * LDS_STRIDE + kq) = v;
        } else {
            const int kk = t >> 3, nq = (t & 7) * 8;
            const uint4 v = *reinterpret_cast<const uint4*>(b + static_cast<size_t>(k0 + kk) * ldb + n0 + nq);
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
            for (int j = 0; j < 8; j++) Bs[(nq + j) * LDS_STRIDE + kk] = e[j];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const bf16x8 av =
                *reinterpret_cast<const bf16x8*>(As + (wm + (lane & 31)) * LDS_STRIDE + kk + (lane >> 5) * 8);
            const bf16x8 bv =
                *reinterpret_cast<const bf16x8*>(Bs + (wn + (lane & 31)) * LDS_STRIDE + kk + (lane >> 5) * 8);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int n = n0 + wn + (lane & 31);
    const float bias = (kEpi != EPI_BF16_MASK && out.bias) ? out.bias[n] : 0.0f;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int m = m0 + wm + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
        const size_t idx = static_cast<size_t>(m) * out.ldc + n;
        const float c = acc[v] + bias;
        if (kEpi == EPI_F32) {
            out.c32[static_cast<size_t>(blockIdx.z) * M * out.ldc + idx] = c;
        } else if (kEpi == EPI_BF16_RELU) {
            out.c16[idx] = f2bf(c > 0.0f ? c : 0.0f);
        } else {
            out.c16[idx] = f2bf(bf2f(out.mask[idx]) > 0.0f ? c : 0.0f);
        }
    }
}

__global__ void sum_slices_kernel(const float* __restrict__ src, const int64_t count, const int splits,
                                  float* __restrict__ dst) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    for (int z = 0; z < splits; z++) s += src[static_cast<int64_t>(z) * count + i];
    dst[i] = s;
}

// Column sums of a bf16 M x ld matrix, first stage: block z sums rows
// [256 z, 256 z + 256) of every column c < ld into part[z * ld + c].
// Threads read 8 consecutive columns (16 bytes) of a row; ld is a multiple
// of 8 dividing 2048.
constexpr int kColRows = 256;
__global__ __launch_bounds__(256) void column_partials_kernel(const uint16_t* __restrict__ g, const int M,
                                                              const int ld, float* __restrict__ part) {
    __shared__ float red[2048];
    const int groups = ld / 8, t = static_cast<int>(threadIdx.x);
    const int cg = t % groups, r0 = t / groups, step = 256 / groups;
    const int m0 = static_cast<int>(blockIdx.x) * kColRows;
    const int m1 = min(M, m0 + kColRows);
    float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int m = m0 + r0; m < m1; m += step) {
        const uint4 v = *reinterpret_cast<const uint4*>(g + static_cast<size_t>(m) * ld + cg * 8);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            acc[2 * j] += __uint_as_float(w[j] << 16);
            acc[2 * j + 1] += __uint_as_float(w[j] & 0xFFFF0000u);
        }
    }
#pragma unroll
    for (int j = 0; j < 8; j++) red[r0 * ld + cg * 8 + j] = acc[j];
    __syncthreads();
    if (t < ld) {
        float s = 0.0f;
        for (int r = 0; r < step; r++) s += red[r * ld + t];
        part[static_cast<size_t>(blockIdx.x) * ld + t] = s;
    }
}

// second stage: db[n] = sum over z < blocks of part[z * ld + n], n < N
__global__ void column_finish_kernel(const float* __restrict__ part, const int blocks, const int ld, const int N,
                                     float* __restrict__ db) {
    const int n = static_cast<int>(blockIdx.x) * blockDim.x + static_cast<int>(threadIdx.x);
    if (n >= N) return;
    float s = 0.0f;
    for (int z = 0; z < blocks; z++) s += part[static_cast<size_t>(z) * ld + n];
    db[n] = s;
}

__global__ void pack_features_kernel(const float* __restrict__ f, const int64_t B, const int64_t B_pad,
                                     uint16_t* __restrict__ x) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= B_pad * kInPad) return;
    const int64_t row = i / kInPad;
    const int col = static_cast<int>(i % kInPad);
    x[i] = (row < B && col < kIn) ? f2bf(f[row * kIn + col]) : static_cast<uint16_t>(0);
}

// log-sigma range of the mixture components (normalised units)
constexpr float kLogSigmaMin = -9.0f, kLogSigmaMax = 6.0f;

__device__ int32_t hex_cell(const mlgpu_hex_cells& c, double x, double y);
__device__ __forceinline__ void image_point(const mlgpu_surrogate_domain& dom, const float* f, double& px,
                                            double& py);

// Gauss-Laguerre nodes and weights of order 4 (Abramowitz & Stegun,
// table 25.9): with u = r^2 / 2, E[g(eps)] of a 2-d standard normal eps is
// sum_j w_j mean over kRuleAngles equally spaced angles of g(sqrt(2 u_j) e_theta).
// The rule is exact for 1, eps and eps^2, so a component entirely on the
// cells gets Z = 1 and no gradient from the truncation.
constexpr int kRuleRadii = 4, kRuleAngles = 8;
__constant__ float kRuleR[kRuleRadii] = {0.803178300f, 1.868561530f, 3.012182031f, 4.334759717f};  // sqrt(2 u_j)
__constant__ float kRuleW[kRuleRadii] = {0.603154104341634f, 0.357418692437800f, 0.0388879085150054f,
                                         0.000539294705561327f};

// Mass Z of the mixture's position marginal on the cells and its gradient
// wrt the outputs o (weight logits, means and log sigmas of x and y) into
// dz (zeroed here); pi are the mixture weights.
__device__ float region_mass(const LossRegion& reg, const float* f, const float* o, const float* pi,
                             const uint32_t rot, float* dz) {
    double px, py;
    image_point(reg.dom, f, px, py);
    const float phi = static_cast<float>(rot >> 8) * (6.2831853071795865f / kRuleAngles / 16777216.0f);
    float c0, s0;
    __sincosf(phi, &s0, &c0);
    constexpr float kHalfSqrt2 = 0.70710678118654752f;
    float zk[kComp], sm[kComp][2], sl[kComp][2];
    float Z = 0.0f;
    for (int k = 0; k < kComp; k++) {
        const float mx = o[2 + 7 * k], my = o[3 + 7 * k];
        const float sx = expf(fminf(fmaxf(o[5 + 7 * k], kLogSigmaMin), kLogSigmaMax));
        const float sy = expf(fminf(fmaxf(o[6 + 7 * k], kLogSigmaMin), kLogSigmaMax));
        float z = 0.0f, a0 = 0.0f, a1 = 0.0f, b0 = 0.0f, b1 = 0.0f;
        for (int j = 0; j < kRuleRadii; j++) {
            const float r = kRuleR[j], w = kRuleW[j] * (1.0f / kRuleAngles);
            float c = c0, sn = s0;
            for (int q = 0; q < kRuleAngles; q++) {
                const float ex = r * c, ey = r * sn;
                const double x = px + static_cast<double>(mx + sx * ex) * reg.dom.position_scale;
                const double y = py + static_cast<double>(my + sy * ey) * reg.dom.position_scale;
                if (hex_cell(reg.cells, x, y) >= 0) {
                    z += w;
                    a0 += w * ex;
                    a1 += w * ey;
                    b0 += w * (ex * ex - 1.0f);
                    b1 += w * (ey * ey - 1.0f);
                }
                const float cn = (c - sn) * kHalfSqrt2;  // rotate by 2 pi / 8
                sn = (c + sn) * kHalfSqrt2;
                c = cn;
            }
        }
        zk[k] = z;
        sm[k][0] = a0 / sx;
        sm[k][1] = a1 / sy;
        sl[k][0] = b0;
        sl[k][1] = b1;
        Z += pi[k] * z;
    }
    for (int j = 0; j < kOutPad; j++) dz[j] = 0.0f;
    for (int k = 0; k < kComp; k++) {
        dz[1 + 7 * k] = pi[k] * (zk[k] - Z);
        for (int d = 0; d < 2; d++) {
            dz[2 + 7 * k + d] = pi[k] * sm[k][d];
            const float raw = o[5 + 7 * k + d];
            if (raw >= kLogSigmaMin && raw <= kLogSigmaMax) dz[5 + 7 * k + d] = pi[k] * sl[k][d];
        }
    }
    return Z;
}

__global__ __launch_bounds__(256) void mixture_loss_kernel(const float* __restrict__ out,
                                                           const float* __restrict__ targets, const int64_t B,
                                                           const int64_t B_pad, const LossRegion reg,
                                                           uint16_t* __restrict__ grad,
                                                           float* __restrict__ loss_sum) {
    __shared__ float part[256];
    const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    float loss = 0.0f;
    if (b < B_pad) {
        float g[kOutPad];
        for (int j = 0; j < kOutPad; j++) g[j] = 0.0f;
        if (b < B) {
            const float* o = out + b * kOutPad;
            const float d = targets[b * 4 + 0];
            const float s = o[0];
            // survival: binary cross entropy of sigmoid(s)
            loss = fmaxf(s, 0.0f) + log1pf(expf(-fabsf(s))) - d * s;
            g[0] = reg.survival_scale * (1.0f / (1.0f + expf(-s)) - d);
            if (d > 0.5f) {
                const float y[3] = {targets[b * 4 + 1], targets[b * 4 + 2], targets[b * 4 + 3]};
                float wmax = -INFINITY;
                for (int k = 0; k < kComp; k++) wmax = fmaxf(wmax, o[1 + 7 * k]);
                float wsum = 0.0f;
                for (int k = 0; k < kComp; k++) wsum += expf(o[1 + 7 * k] - wmax);
                const float lse_w = wmax + logf(wsum);
                float a[kComp], amax = -INFINITY;
                for (int k = 0; k < kComp; k++) {
                    float acc = o[1 + 7 * k] - lse_w - 2.7568155996140185f;  // 1.5 log(2 pi)
                    for (int j = 0; j < 3; j++) {
                        const float ls = fminf(fmaxf(o[5 + 7 * k + j], kLogSigmaMin), kLogSigmaMax);
                        const float z = (y[j] - o[2 + 7 * k + j]) * expf(-ls);
                        acc -= 0.5f * z * z + ls;
                    }
                    a[k] = acc;
                    amax = fmaxf(amax, acc);
                }
                float asum = 0.0f;
                for (int k = 0; k < kComp; k++) asum += expf(a[k] - amax);
                const float lse_a = amax + logf(asum);
                loss -= lse_a;
                for (int k = 0; k < kComp; k++) {
                    const float r = expf(a[k] - lse_a);
                    const float pi = expf(o[1 + 7 * k] - lse_w);
                    g[1 + 7 * k] = pi - r;
                    for (int j = 0; j < 3; j++) {
                        const float raw = o[5 + 7 * k + j];
                        const float ls = fminf(fmaxf(raw, kLogSigmaMin), kLogSigmaMax);
                        const float inv = expf(-ls);
                        const float z = (y[j] - o[2 + 7 * k + j]) * inv;
                        g[2 + 7 * k + j] = -r * z * inv;
                        g[5 + 7 * k + j] = (raw >= kLogSigmaMin && raw <= kLogSigmaMax) ? r * (1.0f - z * z) : 0.0f;
                    }
                }
            } else if (reg.on) {
                // lost photon: -log(1 - sigmoid(s) Z), Z the mixture's mass on the cells
                float wmax = -INFINITY;
                for (int k = 0; k < kComp; k++) wmax = fmaxf(wmax, o[1 + 7 * k]);
                float wsum = 0.0f;
                for (int k = 0; k < kComp; k++) wsum += expf(o[1 + 7 * k] - wmax);
                float pis[kComp];
                for (int k = 0; k < kComp; k++) pis[k] = expf(o[1 + 7 * k] - wmax) / wsum;
                uint32_t h = static_cast<uint32_t>(b) * 0x9E3779B1u ^ reg.rotation_key;
                h ^= h >> 15;
                h *= 0x2C1B3C6Du;
                h ^= h >> 12;
                float dz[kOutPad];
                const float Z = fminf(region_mass(reg, reg.features + b * kIn, o, pis, h, dz), 1.0f);
                const float sg = 1.0f / (1.0f + expf(-s)), om = 1.0f / (1.0f + expf(s));
                const float val = fmaxf(om + sg * (1.0f - Z), 1e-30f);
                loss = -logf(val);
                g[0] = reg.survival_scale * sg * om * Z / val;
                const float gz = -sg / val;
                for (int j = 1; j < kOutPad; j++) g[j] = gz * dz[j];
            }
        }
        const float scale = 1.0f / static_cast<float>(B);
        for (int j = 0; j < kOutPad; j++) grad[b * kOutPad + j] = f2bf(g[j] * scale);
    }
    part[threadIdx.x] = loss;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss_sum, part[0]);
}

__global__ void adam_kernel(const AdamArgs a) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= kNumParams) return;
    const float g = a.grads[i];
    const float m = a.beta1 * a.m[i] + (1.0f - a.beta1) * g;
    const float v = a.beta2 * a.v[i] + (1.0f - a.beta2) * g * g;
    a.m[i] = m;
    a.v[i] = v;
    a.params[i] -= a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
}

// padded element e of all layers: layer by layer, weights then biases.
// The fragment copy holds, for output tile ot (32 rows), K step ks (16
// columns) and lane l, the 8 weights W[32 ot + l % 32][16 ks + 8 (l / 32) + j]:
// element ((ot * ksteps + ks) * 64 + l) * 8 + j; layer 0 uses K = 16.
__device__ __forceinline__ int fragment_ksteps(int l) { return l == 0 ? 1 : kPadIn[l] / 16; }

__global__ void pad_weights_kernel(const float* __restrict__ params, const PaddedWeights pw) {
    int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (int l = 0; l < kLayers; l++) {
        const int64_t nw = static_cast<int64_t>(kPadOut[l]) * kPadIn[l];
        if (e < nw) {
            const int r = static_cast<int>(e / kPadIn[l]), c = static_cast<int>(e % kPadIn[l]);
            const uint16_t v = (r < kLayerOut[l] && c < kLayerIn[l])
                                   ? f2bf(params[weight_offset(l) + static_cast<int64_t>(r) * kLayerIn[l] + c])
                                   : static_cast<uint16_t>(0);
            pw.w[l][e] = v;
            const int ks_n = fragment_ksteps(l);
            if (c < 16 * ks_n) {
                const int ot = r / 32, ks = c / 16, lane = (r % 32) + 32 * ((c % 16) / 8), j = c % 8;
                pw.wf[l][((static_cast<int64_t>(ot) * ks_n + ks) * 64 + lane) * 8 + j] = v;
            }
            return;
        }
        e -= nw;
        if (e < kPadOut[l]) {
            pw.b[l][e] = e < kLayerOut[l] ? params[bias_offset(l) + e] : 0.0f;
            return;
        }
        e -= kPadOut[l];
    }
}

// He-normal initial weights, N(0, 2 / fan_in) (the output layer 0.1 of
// that), zero biases; normals by Box-Muller from Philox block (i, 4, 0, 0).
__global__ void init_params_kernel(float* __restrict__ params, const uint64_t seed) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= kNumParams) return;
    int l = 0;
    while (l + 1 < kLayers && i >= weight_offset(l + 1)) l++;
    if (i >= bias_offset(l)) {
        params[i] = 0.0f;
        return;
    }
    uint32_t w0, w1, w2, w3;
    philox4x32_10(static_cast<uint32_t>(i), 4u, 0u, 0u, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32),
                  w0, w1, w2, w3);
    const double r = sqrt(-2.0 * log_series(1.0 - u53(w0, w1)));
    double sn, cs;
    sincos_turn(u53(w2, w3), sn, cs);
    const double std = sqrt(2.0 / kLayerIn[l]) * (l == kLayers - 1 ? 0.1 : 1.0);
    params[i] = static_cast<float>(std * r * cs);
}

// constants of the feature map of a domain
struct FeatureConsts {
    float inv_r, inv_sa, ilo, lam_scale;
};

__device__ __forceinline__ FeatureConsts feature_consts(const mlgpu_surrogate_domain& dom) {
    FeatureConsts c;
    c.inv_r = static_cast<float>(1.0 / dom.support_radius);
    c.inv_sa = static_cast<float>(1.0 / sin(dom.max_angle));
    c.ilo = static_cast<float>(1.0 / dom.lambda_max);
    c.lam_scale = static_cast<float>(2.0 / (1.0 / dom.lambda_min - 1.0 / dom.lambda_max));
    return c;
}

// the five normalised features of photon i (mlgpu_surrogate_domain), in
// float: the network sees them in bf16
__device__ __forceinline__ void photon_features(const FeatureConsts& c, const mlgpu_photons& ph, const int64_t i,
                                                float* f) {
    const float dx = static_cast<float>(ph.direction_x[i]), dy = static_cast<float>(ph.direction_y[i]);
    const float dz = static_cast<float>(ph.direction_z[i]);
    const float inv_n = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    f[0] = static_cast<float>(ph.support_x[i]) * c.inv_r;
    f[1] = static_cast<float>(ph.support_y[i]) * c.inv_r;
    f[2] = dx * inv_n * c.inv_sa;
    f[3] = dy * inv_n * c.inv_sa;
    // a direction away from -z is outside the domain
    if (!(dz < 0.0f)) f[2] = 2.0f;
    f[4] = (1.0f / static_cast<float>(ph.wavelength[i]) - c.ilo) * c.lam_scale - 1.0f;
}

// nominal image point image_scale (dx / dz, dy / dz) of a photon from its
// features (the direction components are exact to float precision there)
__device__ __forceinline__ void image_point(const mlgpu_surrogate_domain& dom, const float* f, double& px,
                                            double& py) {
    const double sa = sin(dom.max_angle);
    const double dx = static_cast<double>(f[2]) * sa, dy = static_cast<double>(f[3]) * sa;
    const double dz = -sqrt(fmax(1.0 - dx * dx - dy * dy, 1e-12));
    px = dom.image_scale * (dx / dz);
    py = dom.image_scale * (dy / dz);
}

// Targets of photon i: detected (r), the sensor-plane position and the
// arrival time.  The position is where the photon crosses the focal plane
// (the world position of its result in the focal-plane scene, fr) when
// that result is a detection, else the centre of the sensor it ends on.
__global__ void prepare_kernel(const mlgpu_surrogate_domain dom, const mlgpu_photons ph, const mlgpu_results r,
                               const bool with_targets, const double* __restrict__ poses, const mlgpu_results fr,
                               const bool with_focal, const double* __restrict__ focal_poses, const int64_t n,
                               float* __restrict__ features, float* __restrict__ targets) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    photon_features(feature_consts(dom), ph, i, features + i * kIn);
    if (!with_targets) return;
    float* t = targets + i * 4;
    const int32_t frame = r.final_frame[i];
    const bool det = r.sensor_id[i] >= 0 && r.error[i] == 0 && frame >= 0;
    t[0] = det ? 1.0f : 0.0f;
    if (det) {
        double wx, wy;
        const int32_t ff = with_focal ? fr.final_frame[i] : -1;
        if (with_focal && fr.sensor_id[i] >= 0 && fr.error[i] == 0 && ff >= 0) {
            const double* T = focal_poses + 12 * static_cast<int64_t>(ff);
            const double x = fr.x[i], y = fr.y[i];
            wx = T[0] * x + T[1] * y + T[3];
            wy = T[4] * x + T[5] * y + T[7];
        } else {
            const double* T = poses + 12 * static_cast<int64_t>(frame);
            wx = T[3];
            wy = T[7];
        }
        double px, py;
        image_point(dom, features + i * kIn, px, py);
        t[1] = static_cast<float>((wx - px) / dom.position_scale);
        t[2] = static_cast<float>((wy - py) / dom.position_scale);
        t[3] = static_cast<float>((r.time_of_flight[i] - dom.time_offset) / dom.time_scale);
    } else {
        t[1] = t[2] = t[3] = 0.0f;
    }
}

// The cell of a sensor-plane point (mlgpu_hex_cells), or -1: the nearest
// lattice site by cube rounding, then, with cell centres, among that site
// and its six neighbours the cell whose centre is nearest of those whose
// hexagon holds the point (every one for inner radius 0).
template <typename T>
__device__ __forceinline__ int32_t cell_at(const mlgpu_hex_cells& c, const int ii, const int jj) {
    if (ii < -c.radius || ii > c.radius || jj < -c.radius || jj > c.radius) return -1;
    return c.cell[(ii + c.radius) * (2 * c.radius + 1) + (jj + c.radius)];
}

// does the hexagon of inner radius r about (cx, cy), flat sides facing
// n1, n2, n3, hold (x, y)?
template <typename T>
__device__ __forceinline__ bool in_cell_hexagon(const T x, const T y, const T cx, const T cy, const T* n, const T r) {
    const T dx = x - cx, dy = y - cy;
    const T u = dx * n[0] + dy * n[1], v = dx * n[2] + dy * n[3], w = dx * n[4] + dy * n[5];
    return u < r && u > -r && v < r && v > -r && w < r && w > -r;
}

template <typename T>
__device__ __forceinline__ int32_t cell_of_site(const mlgpu_hex_cells& c, const int ii, const int jj, const T x,
                                                const T y, const T* n, const T r) {
    if (!c.centre) return cell_at<T>(c, ii, jj);
    constexpr int kNi[7] = {0, 1, -1, 0, 0, 1, -1};
    constexpr int kNj[7] = {0, 0, 0, 1, -1, -1, 1};
    int32_t best = -1;
    T best_d2 = static_cast<T>(INFINITY);
#pragma unroll
    for (int q = 0; q < 7; q++) {
        const int32_t id = cell_at<T>(c, ii + kNi[q], jj + kNj[q]);
        if (id < 0) continue;
        const T cx = static_cast<T>(c.centre[2 * id]), cy = static_cast<T>(c.centre[2 * id + 1]);
        if (r > static_cast<T>(0) && !in_cell_hexagon<T>(x, y, cx, cy, n, r)) continue;
        const T d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy);
        if (d2 < best_d2) { best_d2 = d2; best = id; }
    }
    return best;
}

// flat-side normals of the cells: a1 turned by 0, 60 and 120 degrees
template <typename T>
__device__ __forceinline__ void cell_normals(const mlgpu_hex_cells& c, T* n) {
    const double l = sqrt(c.a1[0] * c.a1[0] + c.a1[1] * c.a1[1]);
    const double e0 = c.a1[0] / l, e1 = c.a1[1] / l;
    const double h = 0.86602540378443865;  // sin 60
    n[0] = static_cast<T>(e0);
    n[1] = static_cast<T>(e1);
    n[2] = static_cast<T>(0.5 * e0 - h * e1);
    n[3] = static_cast<T>(h * e0 + 0.5 * e1);
    n[4] = static_cast<T>(-0.5 * e0 - h * e1);
    n[5] = static_cast<T>(h * e0 - 0.5 * e1);
}

__device__ int32_t hex_cell(const mlgpu_hex_cells& c, const double x, const double y) {
    const double px = x - c.origin[0], py = y - c.origin[1];
    const double det = c.a1[0] * c.a2[1] - c.a1[1] * c.a2[0];
    const double fi = (px * c.a2[1] - py * c.a2[0]) / det;
    const double fj = (c.a1[0] * py - c.a1[1] * px) / det;
    const double fk = -fi - fj;
    double ri = rint(fi), rj = rint(fj), rk = rint(fk);
    const double di = fabs(ri - fi), dj = fabs(rj - fj), dk = fabs(rk - fk);
    if (di > dj && di > dk) ri = -rj - rk;
    else if (dj > dk) rj = -ri - rk;
    double n[6];
    if (c.centre) cell_normals<double>(c, n);
    return cell_of_site<double>(c, static_cast<int>(ri), static_cast<int>(rj), x, y, n, c.inner_radius);
}

// float form of the draws' constants: the lattice's inverse basis, the
// cells' flat-side normals and inner radius, the image-point and position
// scales
struct DrawConsts {
    float m00, m01, m10, m11;  // inverse of [a1 a2]
    float ox, oy;              // lattice origin
    float n[6];                // flat-side normals of the cells
    float r_in;                // their inner radius
    float sin_angle, image_scale, position_scale;
};

__device__ __forceinline__ DrawConsts draw_consts(const SampleArgs& a) {
    DrawConsts k;
    const double det = a.cells.a1[0] * a.cells.a2[1] - a.cells.a1[1] * a.cells.a2[0];
    k.m00 = static_cast<float>(a.cells.a2[1] / det);
    k.m01 = static_cast<float>(-a.cells.a2[0] / det);
    k.m10 = static_cast<float>(-a.cells.a1[1] / det);
    k.m11 = static_cast<float>(a.cells.a1[0] / det);
    k.ox = static_cast<float>(a.cells.origin[0]);
    k.oy = static_cast<float>(a.cells.origin[1]);
    cell_normals<float>(a.cells, k.n);
    k.r_in = static_cast<float>(a.cells.inner_radius);
    k.sin_angle = static_cast<float>(sin(a.dom.max_angle));
    k.image_scale = static_cast<float>(a.dom.image_scale);
    k.position_scale = static_cast<float>(a.dom.position_scale);
    return k;
}

// hex_cell in float (sensor planes of metres, lattice pitches of millimetres)
__device__ __forceinline__ int32_t hex_cell_f(const mlgpu_hex_cells& c, const DrawConsts& k, const float x,
                                             const float y) {
    const float px = x - k.ox, py = y - k.oy;
    const float fi = k.m00 * px + k.m01 * py;
    const float fj = k.m10 * px + k.m11 * py;
    const float fk = -fi - fj;
    float ri = rintf(fi), rj = rintf(fj), rk = rintf(fk);
    const float di = fabsf(ri - fi), dj = fabsf(rj - fj), dk = fabsf(rk - fk);
    if (di > dj && di > dk) ri = -rj - rk;
    else if (dj > dk) rj = -ri - rk;
    return cell_of_site<float>(c, static_cast<int>(ri), static_cast<int>(rj), x, y, k.n, k.r_in);
}

// uniform in (0, 1] from the top 24 bits of a word
__device__ __forceinline__ float u24_open(const uint32_t w) {
    return static_cast<float>((w >> 8) + 1u) * (1.0f / 16777216.0f);
}
__device__ __forceinline__ float u24(const uint32_t w) { return static_cast<float>(w >> 8) * (1.0f / 16777216.0f); }

// The surrogate's draw for photon i of the call (features f, outputs o):
// false when it is drawn (and binned if it lands on a cell), true when it
// must fall back to propagation.  Float arithmetic with 24-bit uniforms:
// Philox stream 3 block 0 gives the survival and component uniforms and
// the x, y normals (Box-Muller), block 1 the time normal.
__device__ bool sample_one(const SampleArgs& a, const DrawConsts& k, const float* f, const float* o,
                           const int64_t i) {
    const bool in_domain = f[0] * f[0] + f[1] * f[1] <= 1.0f && f[2] * f[2] + f[3] * f[3] <= 1.0f &&
                           f[4] >= -1.0f && f[4] <= 1.0f;
    float wmax = -INFINITY;
    for (int q = 0; q < kComp; q++) wmax = fmaxf(wmax, o[1 + 7 * q]);
    float ew[kComp], wsum = 0.0f;
#pragma unroll
    for (int q = 0; q < kComp; q++) {
        ew[q] = __expf(o[1 + 7 * q] - wmax);
        wsum += ew[q];
    }
    if (a.max_entropy < static_cast<double>(INFINITY)) {
        // entropy bound of the mixture: H(weights) + sum_k pi_k H(component k)
        const float lse = wmax + __logf(wsum);
        float h = 0.0f;
#pragma unroll
        for (int q = 0; q < kComp; q++) {
            const float lp = o[1 + 7 * q] - lse, p = ew[q] / wsum;
            float hk = 3.0f * 1.4189385332046727f;  // 0.5 log(2 pi e) per dimension
            for (int j = 0; j < 3; j++) hk += fminf(fmaxf(o[5 + 7 * q + j], kLogSigmaMin), kLogSigmaMax);
            h += p * (hk - lp);
        }
        if (!(h <= static_cast<float>(a.max_entropy))) return true;
    }
    if (!in_domain) return true;
    const uint64_t id = a.id_offset + static_cast<uint64_t>(i);
    const uint32_t id_lo = static_cast<uint32_t>(id), id_hi = static_cast<uint32_t>(id >> 32);
    const uint32_t k_lo = static_cast<uint32_t>(a.seed), k_hi = static_cast<uint32_t>(a.seed >> 32);
    uint32_t w[4];
    philox4x32_10(0u, 3u, id_lo, id_hi, k_lo, k_hi, w[0], w[1], w[2], w[3]);
    // u < 1 / (1 + exp(-s))
    if (!(u24(w[0]) * (1.0f + __expf(-o[0])) < 1.0f)) return false;
    // component by the cumulative weights
    const float target = u24(w[1]) * wsum;
    int comp = kComp - 1;
    float acc = 0.0f;
#pragma unroll
    for (int q = 0; q < kComp; q++) {
        acc += ew[q];
        if (comp == kComp - 1 && target < acc) comp = q;
    }
    const float r1 = sqrtf(-2.0f * __logf(u24_open(w[2])));
    float s1, c1;
    __sincosf(6.2831853071795865f * u24(w[3]), &s1, &c1);
    const float* oc = o + 7 * comp;
    const float ls0 = fminf(fmaxf(oc[5], kLogSigmaMin), kLogSigmaMax);
    const float ls1 = fminf(fmaxf(oc[6], kLogSigmaMin), kLogSigmaMax);
    const float y0 = oc[2] + __expf(ls0) * (r1 * c1);
    const float y1 = oc[3] + __expf(ls1) * (r1 * s1);
    // nominal image point from the direction features
    const float dx = f[2] * k.sin_angle, dy = f[3] * k.sin_angle;
    const float inv_dz = -1.0f / sqrtf(fmaxf(1.0f - dx * dx - dy * dy, 1e-12f));
    const float x = k.image_scale * dx * inv_dz + y0 * k.position_scale;
    const float y = k.image_scale * dy * inv_dz + y1 * k.position_scale;
    // a position off the cells is a photon that misses the sensors
    const int32_t cell = hex_cell_f(a.cells, k, x, y);
    if (cell < 0) return false;
    philox4x32_10(1u, 3u, id_lo, id_hi, k_lo, k_hi, w[0], w[1], w[2], w[3]);
    const float r2 = sqrtf(-2.0f * __logf(u24_open(w[0])));
    const float z2 = r2 * __cosf(6.2831853071795865f * u24(w[1]));
    const float ls2 = fminf(fmaxf(oc[7], kLogSigmaMin), kLogSigmaMax);
    const float yt = oc[4] + __expf(ls2) * z2;
    const double t = static_cast<double>(yt) * a.dom.time_scale + a.dom.time_offset;
    const int64_t slot = arrival_cell(a.hist, cell, 0, t, a.hist.photon_id_offset + static_cast<uint64_t>(i));
    if (slot >= 0) atomicAdd(&a.hist.counts[slot], 1u);
    return false;
}

// wave-aggregated append of the fallback photons (every lane of the wave calls it)
__device__ __forceinline__ void append_fallback(const SampleArgs& a, const bool fallback, const int64_t i) {
    const uint64_t mask = __ballot(fallback);
    if (mask == 0) return;
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(a.fallback_count, __popcll(mask));
    base = __shfl(base, leader);
    if (fallback) a.fallback[base + __popcll(mask & ((1ull << lane) - 1ull))] = static_cast<int32_t>(i);
}

__global__ __launch_bounds__(256) void sample_kernel(const SampleArgs a, const int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const DrawConsts k = draw_consts(a);
    const bool fallback = i < n && sample_one(a, k, a.features + i * kIn, a.out + i * kOutPad, i);
    append_fallback(a, fallback, i);
}

// ------------------------------------------------------- fused inference
// Features, the whole network and the draws for tiles of kTileP photons in
// one kernel.  Activations stay in LDS as bf16, photon-major (kTileP rows
// of kHStride).  The matrix instructions take the weights as the A operand
// (rows = output features) and the activations as B (columns = photons),
// so a lane's accumulators hold four consecutive outputs of one photon,
// stored to LDS as one 8-byte write.  The weights come from a copy laid out
// in fragment order (pad_weights_kernel): one 1 KB coalesced load feeds a
// wave's instruction.  Wave w computes outputs [64 w, 64 w + 64) of all
// photons in the hidden layers and output tile w % 2 of photon tiles
// 2 (w / 2), + 1 in the last layer.  Optionally writes the outputs
// (FusedArgs::out) and skips the draws (FusedArgs::sample false).
constexpr int kTileP = 128;
constexpr int kHStride = kHidden + 8;
constexpr int kFStride = 16 + 8;
constexpr int kOStride = kOut + 8;  // 65, odd

typedef float float2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
// two floats to packed bf16 (round to nearest even) in one conversion
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    const float2v f = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f, bf16x2v));
}

// ReLU of two packed bf16 values: as 16-bit integers every negative bf16
// (sign bit set) is below 0 and the non-negative ones keep their order, so
// a packed integer max with 0 is the ReLU (-0 becomes +0).
typedef short short2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t v) {
    const short2v x = __builtin_bit_cast(short2v, v);
    const short2v zero = {0, 0};
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(x, zero));
}

__global__ __launch_bounds__(256, 2) void fused_kernel(const FusedArgs a, const int64_t n) {
    __shared__ __attribute__((aligned(16))) uint16_t H[kTileP * kHStride];
    __shared__ __attribute__((aligned(16))) uint16_t F16[kTileP * kFStride];
    __shared__ float F[2][kTileP * kIn];  // features of the current and the next tile
    float* O = reinterpret_cast<float*>(H);
    const int t = static_cast<int>(threadIdx.x), lane = t & 63, w = t >> 6;
    const int lr = lane & 31, lk = (lane >> 5) * 8, lq = (lane >> 5) * 4;
    // the per-call constants of the features and draws live in LDS, read
    // where they are used, so that no register holds them across the
    // matrix loops
    __shared__ DrawConsts dk_lds;
    __shared__ FeatureConsts fc_lds;
    if (t == 0) {
        dk_lds = draw_consts(a.s);
        fc_lds = feature_consts(a.s.dom);
    }
    __syncthreads();
    // features of the tile at `base` into Fb and F16, by one thread per photon
    auto features = [&](const int64_t base, const int p, float* Fb) {
        const int64_t i = base + p;
        float* f = Fb + p * kIn;
        if (i < n) {
            if (a.features) {
#pragma unroll
                for (int j = 0; j < kIn; j++) f[j] = a.features[i * kIn + j];
            } else {
                const FeatureConsts fc = fc_lds;
                photon_features(fc, a.ph, i, f);
            }
        } else {
#pragma unroll
            for (int j = 0; j < kIn; j++) f[j] = 0.0f;
        }
        uint32_t* row = reinterpret_cast<uint32_t*>(F16 + p * kFStride);
        row[0] = pack_bf16(f[0], f[1]);
        row[1] = pack_bf16(f[2], f[3]);
        row[2] = pack_bf16(f[4], 0.0f);
#pragma unroll
        for (int j = 3; j < 8; j++) row[j] = 0u;
    };
    // Waves 2 and 3 compute a tile's features while waves 0 and 1 draw the
    // previous tile's photons (F double-buffered; F16 is free once layer 0
    // has read it).
    const int64_t step = static_cast<int64_t>(gridDim.x) * kTileP;
    int buf = 0;
    if (t >= kTileP && static_cast<int64_t>(blockIdx.x) * kTileP < n)
        features(static_cast<int64_t>(blockIdx.x) * kTileP, t - kTileP, F[0]);
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kTileP; base < n; base += step, buf ^= 1) {
        // weight fragments of the first k-steps of the next layer, loaded
        // ahead of the barrier that precedes it (here layer 0: one k-step)
        bf16x8 pf0, pf1, pf2, pf3;
        {
            const bf16x8* P = reinterpret_cast<const bf16x8*>(a.wf[0]) + (2 * w) * 64 + lane;
            pf0 = P[0];
            pf1 = P[64];
            pf2 = pf0;
            pf3 = pf1;
        }
        __syncthreads();
        // layer 0 (K = 16, five of them non-zero) and the hidden layers.
        // The accumulators start at the bias; the weights of k-step ks + 1
        // are in flight while the instructions of k-step ks run, and the
        // four activation fragments of a k-step are read before its eight
        // matrix instructions.
#pragma unroll 1
        for (int l = 0; l < kLayers - 1; l++) {
            const int ksteps = l == 0 ? 1 : kHidden / 16;
            const bf16x8* W0 = reinterpret_cast<const bf16x8*>(a.wf[l]) + (2 * w) * ksteps * 64 + lane;
            const bf16x8* W1 = W0 + ksteps * 64;
            const uint16_t* src = l == 0 ? F16 : H;
            const int stride = l == 0 ? kFStride : kHStride;
            bf16x8 wc0 = pf0, wc1 = pf1;
            static_assert(kHidden / 16 % 2 == 0, "the k-loop takes k-steps in pairs");
            // the accumulators start at zero; the bias is added when they are
            // packed into the next layer's activations
            const f32x16 zero = {};
            f32x16 acc[2][4];
            // activation fragments of half `h` (photon tiles 2 h, 2 h + 1) of
            // k-step ks
            auto load_h = [&](const int ks, const int h, bf16x8 (&hv)[2]) {
#pragma unroll
                for (int q = 0; q < 2; q++)
                    hv[q] = *reinterpret_cast<const bf16x8*>(src + ((2 * h + q) * 32 + lr) * stride + ks * 16 + lk);
            };
            // the four instructions of half h of a k-step on the weights (w0, w1)
            auto mma = [&](const bf16x8& w0, const bf16x8& w1, const bf16x8 (&hv)[2], const int h, const bool first) {
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const int pt = 2 * h + q;
                    acc[0][pt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w0, hv[q], first ? zero : acc[0][pt], 0, 0, 0);
                    acc[1][pt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1, hv[q], first ? zero : acc[1][pt], 0, 0, 0);
                }
            };
            // The activation fragments of the next half k-step are read from
            // LDS while the current half's instructions run (two sets of two
            // fragments), the weights a k-step ahead (two sets); the
            // scheduling barriers keep the compiler from sinking the loads
            // to their use.
            bf16x8 ha[2], hb[2];
            load_h(0, 0, ha);
            if (ksteps == 1) {
                load_h(0, 1, hb);
                __builtin_amdgcn_sched_barrier(0);
                mma(wc0, wc1, ha, 0, true);
                mma(wc0, wc1, hb, 1, true);
            } else {
                bf16x8 wb0 = pf2, wb1 = pf3;
                load_h(0, 1, hb);
                __builtin_amdgcn_sched_barrier(0);
                mma(wc0, wc1, ha, 0, true);
                load_h(1, 0, ha);
                __builtin_amdgcn_sched_barrier(0);
                mma(wc0, wc1, hb, 1, true);
                const int k2 = ksteps > 2 ? 2 : 1;
                wc0 = W0[k2 * 64];
                wc1 = W1[k2 * 64];
                load_h(1, 1, hb);
                __builtin_amdgcn_sched_barrier(0);
                mma(wb0, wb1, ha, 0, false);
                load_h(k2, 0, ha);
                __builtin_amdgcn_sched_barrier(0);
                mma(wb0, wb1, hb, 1, false);
#pragma unroll 1
                for (int ks = 2; ks < ksteps; ks += 2) {
                    wb0 = W0[(ks + 1) * 64];
                    wb1 = W1[(ks + 1) * 64];
                    load_h(ks, 1, hb);
                    __builtin_amdgcn_sched_barrier(0);
                    mma(wc0, wc1, ha, 0, false);
                    load_h(ks + 1, 0, ha);
                    __builtin_amdgcn_sched_barrier(0);
                    mma(wc0, wc1, hb, 1, false);
                    // past the last k-step the loads repeat it (no branch)
                    const int kn = ks + 2 < ksteps ? ks + 2 : ks + 1;
                    wc0 = W0[kn * 64];
                    wc1 = W1[kn * 64];
                    load_h(ks + 1, 1, hb);
                    __builtin_amdgcn_sched_barrier(0);
                    mma(wb0, wb1, ha, 0, false);
                    load_h(kn, 0, ha);
                    __builtin_amdgcn_sched_barrier(0);
                    mma(wb0, wb1, hb, 1, false);
                }
            }
            if (l + 1 < kLayers - 1) {
                const bf16x8* P = reinterpret_cast<const bf16x8*>(a.wf[l + 1]) + (2 * w) * (kHidden / 16) * 64 + lane;
                pf0 = P[0];
                pf1 = P[(kHidden / 16) * 64];
                pf2 = P[64];
                pf3 = P[(kHidden / 16) * 64 + 64];
            } else {
                const bf16x8* P = reinterpret_cast<const bf16x8*>(a.wf[kLayers - 1]) + (w & 1) * (kHidden / 16) * 64 + lane;
                pf0 = P[0];
                pf2 = P[64];
            }
            __syncthreads();
#pragma unroll
            for (int ot = 0; ot < 2; ot++) {
#pragma unroll
                for (int g = 0; g < 4; g++) {
                    const int col = w * 64 + ot * 32 + g * 8 + lq;
                    const float4 bias = *reinterpret_cast<const float4*>(a.b[l] + col);
#pragma unroll
                    for (int pt = 0; pt < 4; pt++) {
                        const int row = pt * 32 + lr;
                        uint2 packed;
                        packed.x = relu_bf16x2(pack_bf16(acc[ot][pt][4 * g + 0] + bias.x, acc[ot][pt][4 * g + 1] + bias.y));
                        packed.y = relu_bf16x2(pack_bf16(acc[ot][pt][4 * g + 2] + bias.z, acc[ot][pt][4 * g + 3] + bias.w));
                        *reinterpret_cast<uint2*>(H + row * kHStride + col) = packed;
                    }
                }
            }
            __syncthreads();
        }
        {
            const int ot = w & 1, pt0 = (w >> 1) * 2;
            const bf16x8* W = reinterpret_cast<const bf16x8*>(a.wf[kLayers - 1]) + ot * (kHidden / 16) * 64 + lane;
            f32x16 bt;
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const float4 bias = *reinterpret_cast<const float4*>(a.b[kLayers - 1] + ot * 32 + g * 8 + lq);
                bt[4 * g + 0] = bias.x;
                bt[4 * g + 1] = bias.y;
                bt[4 * g + 2] = bias.z;
                bt[4 * g + 3] = bias.w;
            }
            f32x16 acc[2];
            auto load_o = [&](const int ks, bf16x8 (&hv)[2]) {
#pragma unroll
                for (int p = 0; p < 2; p++)
                    hv[p] = *reinterpret_cast<const bf16x8*>(H + ((pt0 + p) * 32 + lr) * kHStride + ks * 16 + lk);
            };
            auto mma = [&](const bf16x8& wv, const bf16x8 (&hv)[2], const bool first) {
#pragma unroll
                for (int p = 0; p < 2; p++)
                    acc[p] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wv, hv[p], first ? bt : acc[p], 0, 0, 0);
            };
            // weights and activations of the next k-step in flight while the
            // current k-step's instructions run (two register sets each)
            bf16x8 wc = pf0, wb = pf2;
            bf16x8 ha[2], hb[2];
            load_o(0, ha);
            load_o(1, hb);
            __builtin_amdgcn_sched_barrier(0);
            mma(wc, ha, true);
            wc = W[2 * 64];
            load_o(2, ha);
            __builtin_amdgcn_sched_barrier(0);
            mma(wb, hb, false);
#pragma unroll 1
            for (int ks = 2; ks < kHidden / 16; ks += 2) {
                wb = W[(ks + 1) * 64];
                load_o(ks + 1, hb);
                __builtin_amdgcn_sched_barrier(0);
                mma(wc, ha, false);
                const int kn = ks + 2 < kHidden / 16 ? ks + 2 : ks + 1;
                wc = W[kn * 64];
                load_o(kn, ha);
                __builtin_amdgcn_sched_barrier(0);
                mma(wb, hb, false);
            }
            __syncthreads();
            // an odd row stride: the draws read one photon's outputs per
            // lane without LDS bank collisions
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int col = ot * 32 + g * 8 + lq;
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const int row = (pt0 + p) * 32 + lr;
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        if (col + e < kOut) O[row * kOStride + col + e] = acc[p][4 * g + e];
                }
            }
            __syncthreads();
        }
        if (a.out) {
            for (int e = t; e < kTileP * kOut; e += 256) {
                const int row = e / kOut, col = e % kOut;
                if (base + row < n) a.out[(base + row) * kOut + col] = O[row * kOStride + col];
            }
        }
        if (t < kTileP) {
            if (a.sample) {
                const int64_t i = base + t;
                const DrawConsts dk = dk_lds;
                const bool fallback = i < n && sample_one(a.s, dk, F[buf] + t * kIn, O + t * kOStride, i);
                append_fallback(a.s, fallback, i);
            }
        } else if (base + step < n) {
            features(base + step, t - kTileP, F[buf ^ 1]);
        }
        __syncthreads();
    }
}

// ------------------------------------------------- fp8 fused inference
// The same tile, waves and epilogues as fused_kernel with layers 1.. on
// v_mfma_scale_f32_32x32x64_f8f6f4: weights and activations e4m3 (OCP),
// fp32 accumulation.  Layer 0 (K = 16) stays bf16.  The activations are
// kept in LDS as e4m3, photon-major (kTileP rows of kQStride bytes): the
// ReLU outputs clamped to 448, the largest e4m3 value.  Row r of a
// weight matrix is stored as W[r] * 2^s (quantize_weights_kernel) and the
// instruction's A scale 2^-s (an e8m0 byte, per lane: every lane holds one
// row) undoes it in the product, so the accumulators hold the plain sums.
// Lane l of the instruction supplies row (or photon) l % 32 and 32 bytes
// of K; the K order inside a k-step of 64 is the same for A and B, which
// is all the sum needs: byte j of lane l is K position 64 ks + 32 (l / 32) + j.
// The activations are stored in a permuted order (act_position) that puts
// the 16 outputs a lane holds for one photon in 16 consecutive bytes, one
// 16-byte LDS write; the weights' K positions follow the same permutation.
// The biases live in LDS; the outputs are stored with an odd row stride
// (kO8Stride floats) so that the draws' per-photon reads do not collide in
// the LDS banks.
typedef int i32x8 __attribute__((__vector_size__(8 * sizeof(int))));
constexpr int kQStride = kHidden + 16;
constexpr int kQSteps = kHidden / 64;  // k-steps of the fp8 layers
constexpr int kO8Stride = kOut + 8;    // 65
constexpr int kBiasFloats = (kLayers - 1) * kHidden + kOutPad;
constexpr float kFp8Max = 448.0f;
constexpr int kFp8Tile = 128;  // photons per block of the fp8 kernel (128 or 256)
static_assert(kTileP * kO8Stride * 4 <= kTileP * kQStride, "the outputs reuse the activations' LDS");
static_assert(kO8Stride % 2 == 1, "odd output stride");
static_assert(kQSteps == 4, "the fp8 k-loop is written out for four k-steps");

// Stored K position of input feature c of layers 1..: inside each block of
// 32, feature 8 a + 4 h + b (the accumulator 4 a + b of the lanes of half
// h) goes to 16 h + 4 a + b.
__device__ __forceinline__ int act_position(int c) {
    return (c & ~31) | (((c >> 2) & 1) << 4) | (((c >> 3) & 3) << 2) | (c & 3);
}

// the largest s with m 2^s <= 448 = 1.75 2^8, for m > 0 (0 otherwise),
// within [-100, 100]
__device__ __forceinline__ int fp8_row_shift(float m) {
    if (!(m > 0.0f)) return 0;
    int e;
    const float f = frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
    const int s = (2.0f * f <= 1.75f ? 8 : 7) - (e - 1);
    return s < -100 ? -100 : (s > 100 ? 100 : s);
}

// four floats to four e4m3 bytes (round to nearest even), the first in the
// lowest byte
__device__ __forceinline__ uint32_t pack_fp8x4(float x0, float x1, float x2, float x3) {
    int p = __builtin_amdgcn_cvt_pk_fp8_f32(x0, x1, 0, false);
    p = __builtin_amdgcn_cvt_pk_fp8_f32(x2, x3, p, true);
    return static_cast<uint32_t>(p);
}

// ReLU clamped to the e4m3 range
__device__ __forceinline__ float act_fp8(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, kFp8Max); }

// C + sum over a k-step of (A 2^(scale_a - 127)) B; both operands e4m3
__device__ __forceinline__ f32x16 mfma_fp8(const i32x8& a, const i32x8& b, const f32x16& c, int scale_a) {
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, scale_a, 0, 127);
}

// One wave per row of layers 1..: the row's largest magnitude, its shift s
// and the row's e4m3 values W[r][c] 2^s at their fragment positions
// (element ((ot * kQSteps + ks) * 64 + l) * 32 + j holds W[32 ot + l % 32][c]
// with act_position(c) = 64 ks + 32 (l / 32) + j); padded rows are zero.
__global__ __launch_bounds__(256) void quantize_weights_kernel(const float* __restrict__ params,
                                                               const PaddedWeights pw) {
    int row = static_cast<int>(blockIdx.x) * 4 + static_cast<int>(threadIdx.x >> 6);
    const int lane = static_cast<int>(threadIdx.x & 63u);
    int l = 1;
    while (l < kLayers && row >= kPadOut[l]) {
        row -= kPadOut[l];
        l++;
    }
    if (l >= kLayers) return;
    static_assert(kHidden == 4 * 64, "one wave quantizes a row of 256 in fours");
    const int c = 4 * lane;
    float v[4];
    float m = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        v[j] = row < kLayerOut[l] ? params[weight_offset(l) + static_cast<int64_t>(row) * kLayerIn[l] + c + j] : 0.0f;
        m = fmaxf(m, fabsf(v[j]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    const int s = fp8_row_shift(m);
    const uint32_t q = pack_fp8x4(ldexpf(v[0], s), ldexpf(v[1], s), ldexpf(v[2], s), ldexpf(v[3], s));
    const int pc = act_position(c);  // c % 4 == 0: the four features stay consecutive
    const int ot = row / 32, ks = pc / 64, fl = row % 32 + 32 * ((pc % 64) / 32), j = pc % 32;
    *reinterpret_cast<uint32_t*>(pw.wq[l] + ((ot * kQSteps + ks) * 64 + fl) * 32 + j) = q;
    if (lane == 0) pw.wscale[l][row] = static_cast<uint8_t>(127 - s);
}

// TP photons per block of 2 TP threads: waves w and w + 4 share neuron
// group w % 4 (the same weight fragments, close in time) on photon halves
// w / 4.
template <int TP>
__global__ __launch_bounds__(2 * TP, 256 / TP) void fused_fp8_kernel(const FusedArgs a, const int64_t n) {
    constexpr int kThreads = 2 * TP;
    __shared__ __attribute__((aligned(16))) uint8_t Q[TP * kQStride];
    __shared__ __attribute__((aligned(16))) uint16_t F16[TP * kFStride];
    __shared__ float F[2][TP * kIn];  // features of the current and the next tile
    __shared__ __attribute__((aligned(16))) float Bs[kBiasFloats];  // the biases, layer by layer
    float* O = reinterpret_cast<float*>(Q);
    const int t = static_cast<int>(threadIdx.x), lane = t & 63, w = t >> 6;
    const int lr = lane & 31, lk = (lane >> 5) * 8, lq = (lane >> 5) * 4, lh = (lane >> 5) * 32;
    const int ng = w & 3, ph = w >> 2;  // neuron group, photon half
    __shared__ DrawConsts dk_lds;
    __shared__ FeatureConsts fc_lds;
    if (t == 0) {
        dk_lds = draw_consts(a.s);
        fc_lds = feature_consts(a.s.dom);
    }
    for (int e = t; e < kBiasFloats; e += kThreads) {
        const int l = e / kHidden < kLayers - 1 ? e / kHidden : kLayers - 1;
        Bs[e] = a.b[l][e - l * kHidden];
    }
    __syncthreads();
    const f32x16 zero = {};
    // features of the tile at `base` into F[buf] and F16, by one thread per photon
    auto features = [&](const int64_t base, const int p, float* Fb) {
        const int64_t i = base + p;
        float* f = Fb + p * kIn;
        if (i < n) {
            if (a.features) {
#pragma unroll
                for (int j = 0; j < kIn; j++) f[j] = a.features[i * kIn + j];
            } else {
                const FeatureConsts fc = fc_lds;
                photon_features(fc, a.ph, i, f);
            }
        } else {
#pragma unroll
            for (int j = 0; j < kIn; j++) f[j] = 0.0f;
        }
        uint32_t* row = reinterpret_cast<uint32_t*>(F16 + p * kFStride);
        row[0] = pack_bf16(f[0], f[1]);
        row[1] = pack_bf16(f[2], f[3]);
        row[2] = pack_bf16(f[4], 0.0f);
#pragma unroll
        for (int j = 3; j < 8; j++) row[j] = 0u;
    };
    // Waves 2 and 3 compute a tile's features while waves 0 and 1 draw the
    // previous tile's photons (F double-buffered; F16 is free once layer 0
    // has read it).
    const int64_t step = static_cast<int64_t>(gridDim.x) * TP;
    int buf = 0;
    if (t >= TP && static_cast<int64_t>(blockIdx.x) * TP < n)
        features(static_cast<int64_t>(blockIdx.x) * TP, t - TP, F[0]);
    __syncthreads();
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * TP; base < n; base += step, buf ^= 1) {
        bf16x8 w00, w01;
        {
            const bf16x8* P = reinterpret_cast<const bf16x8*>(a.wf[0]) + (2 * ng) * 64 + lane;
            w00 = P[0];
            w01 = P[64];
        }
        // layer 0 in bf16 (one k-step of 16)
        f32x16 acc[2][4];
        {
            bf16x8 hv[4];
#pragma unroll
            for (int q = 0; q < 4; q++) hv[q] = *reinterpret_cast<const bf16x8*>(F16 + ((4 * ph + q) * 32 + lr) * kFStride + lk);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                acc[0][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w00, hv[q], zero, 0, 0, 0);
                acc[1][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w01, hv[q], zero, 0, 0, 0);
            }
        }
        // Layer l's outputs (acc, without the bias) become layer l + 1's
        // inputs in Q, then layer l + 1 runs.  The next layer's first weight
        // fragments and its scales load before the barriers.
        i32x8 wc0, wc1;
        int sc0, sc1;
#pragma unroll 1
        for (int l = 0; l < kLayers - 1; l++) {
            const int nl = l + 1;
            if (nl < kLayers - 1) {
                const i32x8* P = reinterpret_cast<const i32x8*>(a.wq[nl]) + (2 * ng) * kQSteps * 64 + lane;
                wc0 = P[0];
                wc1 = P[kQSteps * 64];
                sc0 = a.wscale[nl][ng * 64 + lr];
                sc1 = a.wscale[nl][ng * 64 + 32 + lr];
            } else {
                const int ot = w & 1;
                wc0 = reinterpret_cast<const i32x8*>(a.wq[nl])[ot * kQSteps * 64 + lane];
                wc1 = wc0;
                sc0 = a.wscale[nl][ot * 32 + lr];
                sc1 = sc0;
            }
            __syncthreads();
            // the 16 outputs of a lane for photon tile pt and output tile ot
            // as 16 consecutive e4m3 bytes (act_position)
            const float* bias_l = Bs + l * kHidden + ng * 64;
#pragma unroll
            for (int ot = 0; ot < 2; ot++) {
#pragma unroll
                for (int pt = 0; pt < 4; pt++) {
                    uint4 v;
                    uint32_t* vv = reinterpret_cast<uint32_t*>(&v);
#pragma unroll
                    for (int g = 0; g < 4; g++) {
                        const float4 bias = *reinterpret_cast<const float4*>(bias_l + ot * 32 + g * 8 + lq);
                        vv[g] = pack_fp8x4(act_fp8(acc[ot][pt][4 * g + 0] + bias.x), act_fp8(acc[ot][pt][4 * g + 1] + bias.y),
                                           act_fp8(acc[ot][pt][4 * g + 2] + bias.z), act_fp8(acc[ot][pt][4 * g + 3] + bias.w));
                    }
                    *reinterpret_cast<uint4*>(Q + ((4 * ph + pt) * 32 + lr) * kQStride + ng * 64 + ot * 32 + (lane >> 5) * 16) = v;
                }
            }
            __syncthreads();
            if (nl == kLayers - 1) break;
            const i32x8* W0 = reinterpret_cast<const i32x8*>(a.wq[nl]) + (2 * ng) * kQSteps * 64 + lane;
            const i32x8* W1 = W0 + kQSteps * 64;
            // activation fragments of half h (photon tiles 2 h, 2 h + 1) of k-step ks
            auto load_h = [&](const int ks, const int h, i32x8 (&hv)[2]) {
#pragma unroll
                for (int q = 0; q < 2; q++)
                    hv[q] = *reinterpret_cast<const i32x8*>(Q + ((4 * ph + 2 * h + q) * 32 + lr) * kQStride + ks * 64 + lh);
            };
            auto mma = [&](const i32x8& x0, const i32x8& x1, const i32x8 (&hv)[2], const int h, const bool first) {
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const int pt = 2 * h + q;
                    acc[0][pt] = mfma_fp8(x0, hv[q], first ? zero : acc[0][pt], sc0);
                    acc[1][pt] = mfma_fp8(x1, hv[q], first ? zero : acc[1][pt], sc1);
                }
            };
            // the next half k-step's activations and the next k-step's
            // weights are in flight while the current instructions run
            i32x8 ha[2], hb[2], wn0, wn1;
            load_h(0, 0, ha);
            wn0 = W0[64];
            wn1 = W1[64];
            load_h(0, 1, hb);
            __builtin_amdgcn_sched_barrier(0);
            mma(wc0, wc1, ha, 0, true);
            load_h(1, 0, ha);
            __builtin_amdgcn_sched_barrier(0);
            mma(wc0, wc1, hb, 1, true);
            wc0 = W0[2 * 64];
            wc1 = W1[2 * 64];
            load_h(1, 1, hb);
            __builtin_amdgcn_sched_barrier(0);
            mma(wn0, wn1, ha, 0, false);
            load_h(2, 0, ha);
            __builtin_amdgcn_sched_barrier(0);
            mma(wn0, wn1, hb, 1, false);
            wn0 = W0[3 * 64];
            wn1 = W1[3 * 64];
            load_h(2, 1, hb);
            __builtin_amdgcn_sched_barrier(0);
            mma(wc0, wc1, ha, 0, false);
            load_h(3, 0, ha);
            __builtin_amdgcn_sched_barrier(0);
            mma(wc0, wc1, hb, 1, false);
            load_h(3, 1, hb);
            __builtin_amdgcn_sched_barrier(0);
            mma(wn0, wn1, ha, 0, false);
            __builtin_amdgcn_sched_barrier(0);
            mma(wn0, wn1, hb, 1, false);
        }
        {
            // output layer: wave w, output tile w % 2 of photon tiles 2 (w / 2), + 1
            const int ot = w & 1, pt0 = (w >> 1) * 2;
            const i32x8* W = reinterpret_cast<const i32x8*>(a.wq[kLayers - 1]) + ot * kQSteps * 64 + lane;
            f32x16 bt;
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const float4 bias = *reinterpret_cast<const float4*>(Bs + (kLayers - 1) * kHidden + ot * 32 + g * 8 + lq);
                bt[4 * g + 0] = bias.x;
                bt[4 * g + 1] = bias.y;
                bt[4 * g + 2] = bias.z;
                bt[4 * g + 3] = bias.w;
            }
            f32x16 o2[2];
            auto load_o = [&](const int ks, i32x8 (&hv)[2]) {
#pragma unroll
                for (int p = 0; p < 2; p++)
                    hv[p] = *reinterpret_cast<const i32x8*>(Q + ((pt0 + p) * 32 + lr) * kQStride + ks * 64 + lh);
            };
            auto mma = [&](const i32x8& wv, const i32x8 (&hv)[2], const bool first) {
#pragma unroll
                for (int p = 0; p < 2; p++) o2[p] = mfma_fp8(wv, hv[p], first ? bt : o2[p], sc0);
            };
            i32x8 ha[2], hb[2];
            i32x8 wn = W[64];
            load_o(0, ha);
            load_o(1, hb);
            __builtin_amdgcn_sched_barrier(0);
            mma(wc0, ha, true);
            wc0 = W[2 * 64];
            load_o(2, ha);
            __builtin_amdgcn_sched_barrier(0);
            mma(wn, hb, false);
            wn = W[3 * 64];
            load_o(3, hb);
            __builtin_amdgcn_sched_barrier(0);
            mma(wc0, ha, false);
            __builtin_amdgcn_sched_barrier(0);
            mma(wn, hb, false);
            __syncthreads();
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int col = ot * 32 + g * 8 + lq;
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const int row = (pt0 + p) * 32 + lr;
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        if (col + e < kOut) O[row * kO8Stride + col + e] = o2[p][4 * g + e];
                }
            }
            __syncthreads();
        }
        if (a.out) {
            for (int e = t; e < TP * kOut; e += kThreads) {
                const int row = e / kOut, col = e % kOut;
                if (base + row < n) a.out[(base + row) * kOut + col] = O[row * kO8Stride + col];
            }
        }
        if (t < TP) {
            if (a.sample) {
                const int64_t i = base + t;
                const DrawConsts dk = dk_lds;
                const bool fallback = i < n && sample_one(a.s, dk, F[buf] + t * kIn, O + t * kO8Stride, i);
                append_fallback(a.s, fallback, i);
            }
        } else if (base + step < n) {
            features(base + step, t - TP, F[buf ^ 1]);
        }
        __syncthreads();
    }
}

__global__ void gather_photons_kernel(const mlgpu_photons src, const int32_t* __restrict__ idx,
                                      const int32_t* __restrict__ count, const int64_t n,
                                      const mlgpu_photon_buffers dst) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n || j >= *count) return;
    const int64_t i = idx[j];
    dst.support_x[j] = src.support_x[i];
    dst.support_y[j] = src.support_y[i];
    dst.support_z[j] = src.support_z[i];
    dst.direction_x[j] = src.direction_x[i];
    dst.direction_y[j] = src.direction_y[i];
    dst.direction_z[j] = src.direction_z[i];
    dst.wavelength[j] = src.wavelength[i];
}

int64_t padded_elements() {
    int64_t s = 0;
    for (int l = 0; l < kLayers; l++) s += static_cast<int64_t>(kPadOut[l]) * kPadIn[l] + kPadOut[l];
    return s;
}

}  // namespace

hipError_t launch_gemm(bool a_kcontig, bool b_kcontig, int epilogue, const uint16_t* a, int lda, const uint16_t* b,
                       int ldb, int M, int N, int K, int splits, const GemmOut& out, hipStream_t stream) {
    if (M % BM || N % BN || splits < 1 || K % (BK * splits)) return hipErrorInvalidValue;
    if (splits > 1 && epilogue != EPI_F32) return hipErrorInvalidValue;
    const dim3 grid(static_cast<uint32_t>(N / BN), static_cast<uint32_t>(M / BM), static_cast<uint32_t>(splits));
    const int kchunk = K / splits;
#define MLGPU_GEMM(AK, BKC, E) \
    hipLaunchKernelGGL((gemm_kernel<AK, BKC, E>), grid, dim3(256), 0, stream, a, lda, b, ldb, M, kchunk, out)
    if (a_kcontig && b_kcontig) {
        if (epilogue == EPI_F32) MLGPU_GEMM(true, true, EPI_F32);
        else if (epilogue == EPI_BF16_RELU) MLGPU_GEMM(true, true, EPI_BF16_RELU);
        else MLGPU_GEMM(true, true, EPI_BF16_MASK);
    } else if (a_kcontig && !b_kcontig) {
        if (epilogue == EPI_F32) MLGPU_GEMM(true, false, EPI_F32);
        else if (epilogue == EPI_BF16_RELU) MLGPU_GEMM(true, false, EPI_BF16_RELU);
        else MLGPU_GEMM(true, false, EPI_BF16_MASK);
    } else if (!a_kcontig && !b_kcontig) {
        if (epilogue != EPI_F32) return hipErrorInvalidValue;
        MLGPU_GEMM(false, false, EPI_F32);
    } else {
        if (epilogue != EPI_F32) return hipErrorInvalidValue;
        MLGPU_GEMM(false, true, EPI_F32);
    }
#undef MLGPU_GEMM
    return hipGetLastError();
}

hipError_t launch_sum_slices(const float* src, int64_t count, int splits, float* dst, hipStream_t stream) {
    hipLaunchKernelGGL(sum_slices_kernel, dim3(static_cast<uint32_t>((count + 255) / 256)), dim3(256), 0, stream, src,
                       count, splits, dst);
    return hipGetLastError();
}

hipError_t launch_column_sums(const uint16_t* g, int M, int N, int ld, float* part, float* db, hipStream_t stream) {
    if (ld % 8 != 0 || 2048 % ld != 0 || N > ld) return hipErrorInvalidValue;
    const int blocks = (M + kColRows - 1) / kColRows;
    hipLaunchKernelGGL(column_partials_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, stream, g, M, ld,
                       part);
    hipLaunchKernelGGL(column_finish_kernel, dim3(static_cast<uint32_t>((N + 255) / 256)), dim3(256), 0, stream, part,
                       blocks, ld, N, db);
    return hipGetLastError();
}

hipError_t launch_pack_features(const float* f, int64_t B, int64_t B_pad, uint16_t* x, hipStream_t stream) {
    const int64_t total = B_pad * kInPad;
    hipLaunchKernelGGL(pack_features_kernel, dim3(static_cast<uint32_t>((total + 255) / 256)), dim3(256), 0, stream,
                       f, B, B_pad, x);
    return hipGetLastError();
}

hipError_t launch_mixture_loss(const float* out, const float* targets, int64_t B, int64_t B_pad,
                               const LossRegion& region, uint16_t* grad, float* loss_sum, hipStream_t stream) {
    hipLaunchKernelGGL(mixture_loss_kernel, dim3(static_cast<uint32_t>((B_pad + 255) / 256)), dim3(256), 0, stream,
                       out, targets, B, B_pad, region, grad, loss_sum);
    return hipGetLastError();
}

hipError_t launch_adam(const AdamArgs& a, hipStream_t stream) {
    hipLaunchKernelGGL(adam_kernel, dim3(static_cast<uint32_t>((kNumParams + 255) / 256)), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_init_params(float* params, uint64_t seed, hipStream_t stream) {
    hipLaunchKernelGGL(init_params_kernel, dim3(static_cast<uint32_t>((kNumParams + 255) / 256)), dim3(256), 0, stream,
                       params, seed);
    return hipGetLastError();
}

hipError_t launch_prepare(const mlgpu_surrogate_domain& dom, const mlgpu_photons& ph, const mlgpu_results* r,
                          const double* poses, const mlgpu_results* focal, const double* focal_poses, int64_t n,
                          float* features, float* targets, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const mlgpu_results none{};
    hipLaunchKernelGGL(prepare_kernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, stream, dom, ph,
                       r ? *r : none, r != nullptr, poses, focal ? *focal : none, focal != nullptr, focal_poses, n,
                       features, targets);
    return hipGetLastError();
}

hipError_t launch_sample(const SampleArgs& a, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(sample_kernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, stream, a, n);
    return hipGetLastError();
}

hipError_t launch_fused(const FusedArgs& a, int64_t n, int num_cus, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t tiles = (n + kTileP - 1) / kTileP;
    const int64_t cap = static_cast<int64_t>(num_cus) * 2;
    const dim3 grid(static_cast<uint32_t>(tiles < cap ? tiles : cap));
    if (a.fp8) {
        for (int l = 1; l < kLayers; l++)
            if (!a.wq[l] || !a.wscale[l]) return hipErrorInvalidValue;
        const int64_t tiles8 = (n + kFp8Tile - 1) / kFp8Tile;
        const int64_t cap8 = static_cast<int64_t>(num_cus) * (256 / kFp8Tile);
        hipLaunchKernelGGL(fused_fp8_kernel<kFp8Tile>, dim3(static_cast<uint32_t>(tiles8 < cap8 ? tiles8 : cap8)),
                           dim3(2 * kFp8Tile), 0, stream, a, n);
    } else {
        hipLaunchKernelGGL(fused_kernel, grid, dim3(256), 0, stream, a, n);
    }
    return hipGetLastError();
}

hipError_t launch_gather_photons(const mlgpu_photons& src, const int32_t* idx, const int32_t* count, int64_t n,
                                 const mlgpu_photon_buffers& dst, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_photons_kernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, stream, src,
                       idx, count, n, dst);
    return hipGetLastError();
}

hipError_t launch_pad_weights(const float* params, const PaddedWeights& pw, hipStream_t stream) {
    const int64_t total = padded_elements();
    hipLaunchKernelGGL(pad_weights_kernel, dim3(static_cast<uint32_t>((total + 255) / 256)), dim3(256), 0, stream,
                       params, pw);
    if (pw.wq[1]) {
        int rows = 0;
        for (int l = 1; l < kLayers; l++) rows += kPadOut[l];
        hipLaunchKernelGGL(quantize_weights_kernel, dim3(static_cast<uint32_t>((rows + 3) / 4)), dim3(256), 0, stream,
                           params, pw);
    }
    return hipGetLastError();
}

}  // namespace sur
}  // namespace mlgpu
