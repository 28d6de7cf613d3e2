// Learned surrogate of the propagator: shapes, parameter layout and the
// launchers of its HIP kernels (surrogate.hip).
//
// Network: 5 -> 256 -> 256 -> 256 -> 256 -> 57, ReLU between layers.  The
// 57 outputs are the survival logit and 8 mixture components of
// (weight logit, mean x y t, log sigma x y t) over the sensor-plane
// position and arrival time of a photon that reaches a sensor.
// Weights and activations are bf16 in the matrix kernels (MFMA, fp32
// accumulation); the optimiser keeps fp32 master weights.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mlgpu.h"

namespace mlgpu {
namespace sur {

constexpr int kIn = 5;         // x, y, cx, cy, wavelength (normalised)
constexpr int kInPad = 64;     // input width as the matrix kernels see it
constexpr int kHidden = 256;
constexpr int kComp = 8;       // mixture components
constexpr int kOut = 1 + 7 * kComp;  // 57
constexpr int kOutPad = 64;
constexpr int kLayers = 5;
constexpr int kBatchAlign = 64;  // batches are padded to this many rows

// Logical (unpadded) layer shapes: out x in.
constexpr int kLayerOut[kLayers] = {kHidden, kHidden, kHidden, kHidden, kOut};
constexpr int kLayerIn[kLayers] = {kIn, kHidden, kHidden, kHidden, kHidden};
// Padded shapes of the bf16 weights the kernels read.
constexpr int kPadOut[kLayers] = {kHidden, kHidden, kHidden, kHidden, kOutPad};
constexpr int kPadIn[kLayers] = {kInPad, kHidden, kHidden, kHidden, kHidden};

// Offsets of W_l (row-major out x in) and b_l in the flat fp32 parameter
// vector: W0 b0 W1 b1 ... W4 b4.
constexpr int64_t weight_offset(int l) {
    int64_t o = 0;
    for (int i = 0; i < l; i++) o += static_cast<int64_t>(kLayerOut[i]) * kLayerIn[i] + kLayerOut[i];
    return o;
}
constexpr int64_t bias_offset(int l) { return weight_offset(l) + static_cast<int64_t>(kLayerOut[l]) * kLayerIn[l]; }
constexpr int64_t kNumParams = weight_offset(kLayers);

// Epilogues of the matrix kernel.
enum Epilogue : int {
    EPI_F32 = 0,         // C (+ bias) as fp32
    EPI_BF16_RELU = 1,   // relu(C + bias) as bf16
    EPI_BF16_MASK = 2,   // C where mask > 0 else 0, as bf16 (back through a ReLU)
};

struct GemmOut {
    float* c32;           // EPI_F32 output (split-K: slice z at c32 + z * M * ldc)
    uint16_t* c16;        // bf16 outputs
    const float* bias;    // per column n, or null
    const uint16_t* mask; // EPI_BF16_MASK: bf16 matrix, same shape and ldc as C
    int ldc;
};

// C[M x N] = sum_k A(m, k) B(n, k): A(m, k) at a[m * lda + k] when a_kcontig
// else a[k * lda + m]; B(n, k) at b[n * ldb + k] when b_kcontig else
// b[k * ldb + n].  M, N multiples of 64; K a multiple of 32 * splits.
hipError_t launch_gemm(bool a_kcontig, bool b_kcontig, int epilogue, const uint16_t* a, int lda, const uint16_t* b,
                       int ldb, int M, int N, int K, int splits, const GemmOut& out, hipStream_t stream);

// dst[i] = sum over z < splits of src[z * count + i]
hipError_t launch_sum_slices(const float* src, int64_t count, int splits, float* dst, hipStream_t stream);

// bias gradient: db[n] = sum_m g(m, n) of a bf16 M x ld matrix, n < N;
// part holds ceil(M / 256) * ld floats of per-block partial sums
hipError_t launch_column_sums(const uint16_t* g, int M, int N, int ld, float* part, float* db, hipStream_t stream);

// Features (B x kInPad bf16, zero padded) from fp32 B x kIn.
hipError_t launch_pack_features(const float* f, int64_t B, int64_t B_pad, uint16_t* x, hipStream_t stream);

// Truncation of the mixture to sensor cells in the loss (mlgpu_surrogate_set_region).
struct LossRegion {
    bool on;
    mlgpu_surrogate_domain dom;
    mlgpu_hex_cells cells;
    const float* features;  // B x kIn fp32, for the photons' image points
    uint32_t rotation_key;  // varies the quadrature's rotation between steps
    float survival_scale;   // factor on the gradient wrt the survival logit
};

// Loss of the outputs (B x kOutPad fp32) for targets (B x 4: detected,
// x, y, t), its gradient as bf16 (B x kOutPad, scaled by 1 / B) and the
// summed loss (added to *loss_sum).
hipError_t launch_mixture_loss(const float* out, const float* targets, int64_t B, int64_t B_pad,
                               const LossRegion& region, uint16_t* grad, float* loss_sum, hipStream_t stream);

// Adam on the flat fp32 parameters; writes the padded bf16 weight copies.
struct AdamArgs {
    float* params;
    const float* grads;
    float* m;
    float* v;
    float lr, beta1, beta2, eps;
    float c1, c2;  // 1 / (1 - beta^t)
};
hipError_t launch_adam(const AdamArgs& a, hipStream_t stream);

// fp32 flat parameters -> padded bf16 weights (per layer kPadOut x kPadIn)
// and padded fp32 biases (per layer kPadOut).
struct PaddedWeights {
    uint16_t* w[kLayers];
    float* b[kLayers];
    uint16_t* wf[kLayers];   // fragment-ordered copy for the fused inference kernel
    // fp8 (e4m3) fragment-ordered copies of layers 1.. for the fp8 inference
    // kernel, row r of layer l stored as W[r] * 2^s (the largest s with
    // max |W[r]| * 2^s <= 448), and the rows' scale bytes 127 - s (layer 0:
    // null)
    uint8_t* wq[kLayers];
    uint8_t* wscale[kLayers];
};
hipError_t launch_pad_weights(const float* params, const PaddedWeights& pw, hipStream_t stream);

// He-normal initial parameters from Philox (seed).
hipError_t launch_init_params(float* params, uint64_t seed, hipStream_t stream);

// Features and (when r is not null) targets of photons (mlgpu_surrogate_prepare).
hipError_t launch_prepare(const mlgpu_surrogate_domain& dom, const mlgpu_photons& ph, const mlgpu_results* r,
                          const double* poses, const mlgpu_results* focal, const double* focal_poses, int64_t n,
                          float* features, float* targets, hipStream_t stream);

// Surrogate draws of n photons from their outputs (rows of kOutPad fp32):
// photons in the domain with a low entropy bound are drawn and binned into
// hist when they land on a cell; the others are appended to fallback
// (count in *fallback_count).
struct SampleArgs {
    const float* out;
    const float* features;   // n x kIn, for the domain test
    mlgpu_surrogate_domain dom;
    mlgpu_hex_cells cells;
    mlgpu_histogram hist;
    double max_entropy;
    uint64_t seed;
    uint64_t id_offset;      // Philox id of photon 0
    int32_t* fallback;
    int32_t* fallback_count;
};
hipError_t launch_sample(const SampleArgs& a, int64_t n, hipStream_t stream);

// Features, network and draws in one kernel (the inference path): photons
// (or ready features) in, outputs (n x kOut, optional) and draws out.
struct FusedArgs {
    mlgpu_photons ph;
    const float* features;     // n x kIn instead of photons, or null
    SampleArgs s;              // its out and features are unused
    bool sample;
    float* out;                // n x kOut, or null
    const uint16_t* wf[kLayers];  // fragment-ordered bf16 weights
    const float* b[kLayers];
    bool fp8;                     // layers 1.. on the fp8 copies (PaddedWeights::wq, wscale)
    const uint8_t* wq[kLayers];
    const uint8_t* wscale[kLayers];
};
hipError_t launch_fused(const FusedArgs& a, int64_t n, int num_cus, hipStream_t stream);

// dst[j] = src[idx[j]] for the 7 photon arrays, j < *count (count <= n)
hipError_t launch_gather_photons(const mlgpu_photons& src, const int32_t* idx, const int32_t* count, int64_t n,
                                 const mlgpu_photon_buffers& dst, hipStream_t stream);

}  // namespace sur
}  // namespace mlgpu
