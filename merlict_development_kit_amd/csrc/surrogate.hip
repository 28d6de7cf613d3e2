// HIP kernels of the learned surrogate (surrogate.hpp): a bf16 matrix
// kernel on the CDNA4 matrix cores (v_mfma_f32_32x32x16_bf16, fp32
// accumulation) with the epilogues the network needs, the mixture loss and
// its gradient, Adam, and the packing of weights and features.
#include <hip/hip_runtime.h>

#include <cmath>

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

// 64 columns per block; threads (c, r) sum rows r, r + 4, ... of column c.
__global__ __launch_bounds__(256) void column_sums_kernel(const uint16_t* __restrict__ g, const int M, const int N,
                                                          const int ld, float* __restrict__ db) {
    __shared__ float part[4][64];
    const int c = static_cast<int>(threadIdx.x) & 63, r = static_cast<int>(threadIdx.x) >> 6;
    const int n = static_cast<int>(blockIdx.x) * 64 + c;
    float s = 0.0f;
    if (n < N)
        for (int m = r; m < M; m += 4) s += bf2f(g[static_cast<size_t>(m) * ld + n]);
    part[r][c] = s;
    __syncthreads();
    if (r == 0 && n < N) db[n] = (part[0][c] + part[1][c]) + (part[2][c] + part[3][c]);
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

__global__ __launch_bounds__(256) void mixture_loss_kernel(const float* __restrict__ out,
                                                           const float* __restrict__ targets, const int64_t B,
                                                           const int64_t B_pad, uint16_t* __restrict__ grad,
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
            g[0] = 1.0f / (1.0f + expf(-s)) - d;
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

// padded element e of all layers: layer by layer, weights then biases
__global__ void pad_weights_kernel(const float* __restrict__ params, const PaddedWeights pw) {
    int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (int l = 0; l < kLayers; l++) {
        const int64_t nw = static_cast<int64_t>(kPadOut[l]) * kPadIn[l];
        if (e < nw) {
            const int r = static_cast<int>(e / kPadIn[l]), c = static_cast<int>(e % kPadIn[l]);
            pw.w[l][e] = (r < kLayerOut[l] && c < kLayerIn[l])
                             ? f2bf(params[weight_offset(l) + static_cast<int64_t>(r) * kLayerIn[l] + c])
                             : static_cast<uint16_t>(0);
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

hipError_t launch_column_sums(const uint16_t* g, int M, int N, int ld, float* db, hipStream_t stream) {
    hipLaunchKernelGGL(column_sums_kernel, dim3(static_cast<uint32_t>((N + 63) / 64)), dim3(256), 0, stream, g, M, N,
                       ld, db);
    return hipGetLastError();
}

hipError_t launch_pack_features(const float* f, int64_t B, int64_t B_pad, uint16_t* x, hipStream_t stream) {
    const int64_t total = B_pad * kInPad;
    hipLaunchKernelGGL(pack_features_kernel, dim3(static_cast<uint32_t>((total + 255) / 256)), dim3(256), 0, stream,
                       f, B, B_pad, x);
    return hipGetLastError();
}

hipError_t launch_mixture_loss(const float* out, const float* targets, int64_t B, int64_t B_pad, uint16_t* grad,
                               float* loss_sum, hipStream_t stream) {
    hipLaunchKernelGGL(mixture_loss_kernel, dim3(static_cast<uint32_t>((B_pad + 255) / 256)), dim3(256), 0, stream,
                       out, targets, B, B_pad, grad, loss_sum);
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

hipError_t launch_pad_weights(const float* params, const PaddedWeights& pw, hipStream_t stream) {
    const int64_t total = padded_elements();
    hipLaunchKernelGGL(pad_weights_kernel, dim3(static_cast<uint32_t>((total + 255) / 256)), dim3(256), 0, stream,
                       params, pw);
    return hipGetLastError();
}

}  // namespace sur
}  // namespace mlgpu
