// C ABI of the learned surrogate (include/mlgpu.h): parameters, forward,
// gradients and Adam training steps on the device.  The network's matrix
// products run in surrogate.hip on the matrix cores.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "scene.hpp"
#include "surrogate.hpp"

using namespace mlgpu::sur;

struct mlgpu_surrogate {
    int device = 0;
    float* params = nullptr;  // fp32 master parameters, flat
    float* grads = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    PaddedWeights pw{};
    void* weights_mem = nullptr;
    int64_t step = 0;
    bool fp8 = false;  // inference precision of layers 1.. (mlgpu_surrogate_set_inference)
    std::mutex mutex;  // guards the batch workspace, the step counter and the region
    // factor on the loss gradient wrt the survival logit
    // (mlgpu_surrogate_set_survival_scale)
    float survival_scale = 1.0f;
    // truncation region of the training loss (mlgpu_surrogate_set_region)
    bool region_on = false;
    mlgpu_surrogate_domain region_dom{};
    mlgpu_hex_cells region_cells{};
    // batch workspace, rows = padded batch
    void* ws = nullptr;
    int64_t ws_rows = 0;
    uint16_t* x0 = nullptr;            // rows x kInPad
    uint16_t* h[kLayers - 1] = {};     // rows x kHidden, post-ReLU activations
    float* out = nullptr;              // rows x kOutPad
    uint16_t* gout = nullptr;          // rows x kOutPad
    uint16_t* gh[2] = {};              // rows x kHidden, gradients wrt pre-activations
    float* partial = nullptr;          // split-K slices of a weight gradient
    float* loss = nullptr;             // scalar
    // hybrid buffers, one chunk of photons
    void* hy = nullptr;
    float* hy_features = nullptr;      // chunk x kIn
    int32_t* hy_fallback = nullptr;    // chunk
    int32_t* hy_count = nullptr;       // 1
    mlgpu_photon_buffers hy_photons{}; // chunk, the gathered fallback photons
    mlgpu_results hy_results{};        // chunk, their plain results
    int32_t* hy_count_host = nullptr;  // pinned
};

namespace {

using mlgpu::set_last_error;

constexpr int kKChunk = 512;  // rows of the batch per split of a weight-gradient product

int hip_fail(hipError_t e, const char* what) {
    set_last_error(std::string(what) + ": " + hipGetErrorString(e));
    return MLGPU_ERR_HIP;
}

int64_t padded_rows(int64_t n) { return (n + kBatchAlign - 1) / kBatchAlign * kBatchAlign; }

size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

int ensure_workspace(mlgpu_surrogate* s, int64_t rows) {
    if (s->ws_rows >= rows) return MLGPU_OK;
    if (s->ws) {
        (void)hipDeviceSynchronize();
        (void)hipFree(s->ws);
        s->ws = nullptr;
        s->ws_rows = 0;
    }
    const size_t r = static_cast<size_t>(rows);
    const int splits = static_cast<int>((rows + kKChunk - 1) / kKChunk);
    size_t bytes = align256(r * kInPad * 2) + (kLayers - 1) * align256(r * kHidden * 2) +
                   align256(r * kOutPad * 4) + align256(r * kOutPad * 2) + 2 * align256(r * kHidden * 2) +
                   align256(static_cast<size_t>(splits) * kHidden * kHidden * 4) + 256;
    hipError_t e = hipMalloc(&s->ws, bytes);
    if (e != hipSuccess) { s->ws = nullptr; return hip_fail(e, "hipMalloc surrogate workspace"); }
    char* p = static_cast<char*>(s->ws);
    auto take = [&](size_t b) { char* q = p; p += align256(b); return q; };
    s->x0 = reinterpret_cast<uint16_t*>(take(r * kInPad * 2));
    for (int l = 0; l < kLayers - 1; l++) s->h[l] = reinterpret_cast<uint16_t*>(take(r * kHidden * 2));
    s->out = reinterpret_cast<float*>(take(r * kOutPad * 4));
    s->gout = reinterpret_cast<uint16_t*>(take(r * kOutPad * 2));
    for (int k = 0; k < 2; k++) s->gh[k] = reinterpret_cast<uint16_t*>(take(r * kHidden * 2));
    s->partial = reinterpret_cast<float*>(take(static_cast<size_t>(splits) * kHidden * kHidden * 4));
    s->loss = reinterpret_cast<float*>(take(256));
    s->ws_rows = rows;
    return MLGPU_OK;
}

// activations h[0..3] and outputs `out` of the padded batch in x0
hipError_t forward(mlgpu_surrogate* s, int64_t rows, hipStream_t st) {
    const int M = static_cast<int>(rows);
    hipError_t e = hipSuccess;
    const uint16_t* in = s->x0;
    for (int l = 0; l < kLayers; l++) {
        GemmOut o{};
        o.bias = s->pw.b[l];
        if (l < kLayers - 1) {
            o.c16 = s->h[l];
            o.ldc = kHidden;
            e = launch_gemm(true, true, EPI_BF16_RELU, in, kPadIn[l], s->pw.w[l], kPadIn[l], M, kPadOut[l], kPadIn[l],
                            1, o, st);
            in = s->h[l];
        } else {
            o.c32 = s->out;
            o.ldc = kOutPad;
            e = launch_gemm(true, true, EPI_F32, in, kPadIn[l], s->pw.w[l], kPadIn[l], M, kPadOut[l], kPadIn[l], 1, o,
                            st);
        }
        if (e != hipSuccess) return e;
    }
    return e;
}

// gradient of layer l's weights: dW_l = G^T X over the batch, G (rows x
// kPadOut[l], ld ldg) the gradient wrt the layer's outputs, X its input
hipError_t weight_gradient(mlgpu_surrogate* s, int l, const uint16_t* g, int ldg, const uint16_t* x, int rows,
                           hipStream_t st) {
    const int splits = rows / kKChunk > 0 && rows % kKChunk == 0 ? rows / kKChunk : 1;
    GemmOut o{};
    o.c32 = s->partial;
    o.ldc = kPadIn[l];
    // C[kPadOut x kPadIn] = sum_b G(b, n) X(b, k): A = G^T (n contiguous), B = X (k contiguous per row b)
    hipError_t e = launch_gemm(false, false, EPI_F32, g, ldg, x, kPadIn[l], kPadOut[l], kPadIn[l], rows, splits, o, st);
    if (e != hipSuccess) return e;
    // fold the splits into slice 0, then copy the logical block
    e = launch_sum_slices(s->partial, static_cast<int64_t>(kPadOut[l]) * kPadIn[l], splits, s->partial, st);
    if (e != hipSuccess) return e;
    // logical rows / columns into the flat gradient vector
    const size_t row_bytes = static_cast<size_t>(kLayerIn[l]) * 4;
    e = hipMemcpy2DAsync(s->grads + weight_offset(l), row_bytes, s->partial, static_cast<size_t>(kPadIn[l]) * 4,
                         row_bytes, static_cast<size_t>(kLayerOut[l]), hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
    // the slices are free again once the copy above has run (same stream)
    return launch_column_sums(g, rows, kLayerOut[l], ldg, s->partial, s->grads + bias_offset(l), st);
}

// loss and gradients of the batch (features already packed in x0)
hipError_t backward(mlgpu_surrogate* s, const float* features, const float* targets, int64_t n, int64_t rows,
                    hipStream_t st) {
    const int M = static_cast<int>(rows);
    LossRegion reg{};
    reg.on = s->region_on;
    reg.dom = s->region_dom;
    reg.cells = s->region_cells;
    reg.features = features;
    reg.rotation_key = static_cast<uint32_t>(static_cast<uint64_t>(s->step) * 0x9E3779B97F4A7C15ull >> 32);
    reg.survival_scale = s->survival_scale;
    hipError_t e = hipMemsetAsync(s->loss, 0, sizeof(float), st);
    if (e == hipSuccess) e = forward(s, rows, st);
    if (e == hipSuccess) e = launch_mixture_loss(s->out, targets, n, rows, reg, s->gout, s->loss, st);
    if (e != hipSuccess) return e;
    const uint16_t* g = s->gout;
    int ldg = kOutPad;
    for (int l = kLayers - 1; l >= 0; l--) {
        const uint16_t* x = l > 0 ? s->h[l - 1] : s->x0;
        e = weight_gradient(s, l, g, ldg, x, M, st);
        if (e != hipSuccess) return e;
        if (l == 0) break;
        // gradient wrt layer l's input, through the ReLU of layer l - 1:
        // C[rows x kHidden] = sum_n G(b, n) W_l(n, k), W_l row-major kPadOut x kPadIn
        uint16_t* gh = s->gh[l & 1];
        GemmOut o{};
        o.c16 = gh;
        o.mask = s->h[l - 1];
        o.ldc = kHidden;
        e = launch_gemm(true, false, EPI_BF16_MASK, g, ldg, s->pw.w[l], kPadIn[l], M, kPadIn[l], kPadOut[l], 1, o, st);
        if (e != hipSuccess) return e;
        g = gh;
        ldg = kHidden;
    }
    return hipSuccess;
}

constexpr int64_t kHybridChunk = int64_t{1} << 24;

int ensure_hybrid(mlgpu_surrogate* s) {
    if (s->hy) return MLGPU_OK;
    const size_t c = static_cast<size_t>(kHybridChunk);
    const size_t bytes = align256(c * kIn * 4) + align256(c * 4) + 256 + 7 * align256(c * 8) + 2 * align256(c * 4) +
                         3 * align256(c) + align256(c * 8) + 4 * align256(c * 4);
    hipError_t e = hipMalloc(&s->hy, bytes);
    if (e != hipSuccess) { s->hy = nullptr; return hip_fail(e, "hipMalloc hybrid buffers"); }
    e = hipHostMalloc(reinterpret_cast<void**>(&s->hy_count_host), sizeof(int32_t));
    if (e != hipSuccess) { (void)hipFree(s->hy); s->hy = nullptr; return hip_fail(e, "hipHostMalloc"); }
    char* p = static_cast<char*>(s->hy);
    auto take = [&](size_t b) { char* q = p; p += align256(b); return q; };
    s->hy_features = reinterpret_cast<float*>(take(c * kIn * 4));
    s->hy_fallback = reinterpret_cast<int32_t*>(take(c * 4));
    s->hy_count = reinterpret_cast<int32_t*>(take(256));
    double** ph[7] = {&s->hy_photons.support_x, &s->hy_photons.support_y, &s->hy_photons.support_z,
                      &s->hy_photons.direction_x, &s->hy_photons.direction_y, &s->hy_photons.direction_z,
                      &s->hy_photons.wavelength};
    for (double** q : ph) *q = reinterpret_cast<double*>(take(c * 8));
    s->hy_results.final_frame = reinterpret_cast<int32_t*>(take(c * 4));
    s->hy_results.sensor_id = reinterpret_cast<int32_t*>(take(c * 4));
    s->hy_results.final_interaction = reinterpret_cast<uint8_t*>(take(c));
    s->hy_results.num_interactions = reinterpret_cast<uint8_t*>(take(c));
    s->hy_results.error = reinterpret_cast<uint8_t*>(take(c));
    s->hy_results.time_of_flight = reinterpret_cast<double*>(take(c * 8));
    s->hy_results.x = reinterpret_cast<float*>(take(c * 4));
    s->hy_results.y = reinterpret_cast<float*>(take(c * 4));
    s->hy_results.cx = reinterpret_cast<float*>(take(c * 4));
    s->hy_results.cy = reinterpret_cast<float*>(take(c * 4));
    s->hy_results.accumulated_distance = nullptr;
    return MLGPU_OK;
}

bool domain_ok(const mlgpu_surrogate_domain* d) {
    return d && d->support_radius > 0.0 && d->max_angle > 0.0 && d->max_angle < 1.5707963267948966 &&
           d->lambda_min > 0.0 && d->lambda_max > d->lambda_min && d->position_scale > 0.0 && d->time_scale > 0.0 &&
           std::isfinite(d->image_scale) && std::isfinite(d->support_radius) && std::isfinite(d->lambda_max) &&
           std::isfinite(d->position_scale) &&
           std::isfinite(d->time_offset) && std::isfinite(d->time_scale);
}

int num_cus_of(int device) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 256;
    return cus;
}

FusedArgs fused_args(const mlgpu_surrogate* s) {
    FusedArgs a{};
    for (int l = 0; l < kLayers; l++) {
        a.wf[l] = s->pw.wf[l];
        a.b[l] = s->pw.b[l];
        a.wq[l] = s->pw.wq[l];
        a.wscale[l] = s->pw.wscale[l];
    }
    a.fp8 = s->fp8;
    return a;
}

// bf16 elements of layer l's fragment-ordered weight copy
size_t fragment_elements(int l) {
    const int ksteps = l == 0 ? 1 : kPadIn[l] / 16;
    return static_cast<size_t>(kPadOut[l] / 32) * ksteps * 64 * 8;
}

int check_batch(const mlgpu_surrogate* s, const float* features, int64_t n) {
    if (!s || !features || n < 1) { set_last_error("null argument or empty batch"); return MLGPU_ERR_ARGUMENT; }
    if (n > (int64_t{1} << 24)) { set_last_error("at most 2^24 photons per surrogate batch"); return MLGPU_ERR_ARGUMENT; }
    return MLGPU_OK;
}

}  // namespace

namespace {

// a lattice table, and a hexagon radius >= 0 (0: nearest centre)
bool cells_ok(const mlgpu_hex_cells& c) {
    return c.cell && c.radius >= 0 && c.inner_radius >= 0.0 && std::isfinite(c.inner_radius);
}

}  // namespace

extern "C" {

int64_t mlgpu_surrogate_num_params(void) { return kNumParams; }

int mlgpu_surrogate_create(uint64_t seed, mlgpu_surrogate** out) {
    if (!out) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    *out = nullptr;
    auto* s = new mlgpu_surrogate();
    (void)hipGetDevice(&s->device);
    const size_t pbytes = static_cast<size_t>(kNumParams) * 4;
    hipError_t e = hipMalloc(&s->params, 4 * align256(pbytes));
    if (e != hipSuccess) { delete s; return hip_fail(e, "hipMalloc surrogate parameters"); }
    char* p = reinterpret_cast<char*>(s->params);
    s->grads = reinterpret_cast<float*>(p + align256(pbytes));
    s->adam_m = reinterpret_cast<float*>(p + 2 * align256(pbytes));
    s->adam_v = reinterpret_cast<float*>(p + 3 * align256(pbytes));
    size_t wbytes = 0;
    for (int l = 0; l < kLayers; l++)
        wbytes += align256(static_cast<size_t>(kPadOut[l]) * kPadIn[l] * 2) + align256(static_cast<size_t>(kPadOut[l]) * 4) +
                  align256(fragment_elements(l) * 2) +
                  (l > 0 ? align256(static_cast<size_t>(kPadOut[l]) * kPadIn[l]) + align256(kPadOut[l]) : 0);
    e = hipMalloc(&s->weights_mem, wbytes);
    if (e != hipSuccess) { (void)hipFree(s->params); delete s; return hip_fail(e, "hipMalloc surrogate weights"); }
    char* q = static_cast<char*>(s->weights_mem);
    for (int l = 0; l < kLayers; l++) {
        s->pw.w[l] = reinterpret_cast<uint16_t*>(q);
        q += align256(static_cast<size_t>(kPadOut[l]) * kPadIn[l] * 2);
        s->pw.b[l] = reinterpret_cast<float*>(q);
        q += align256(static_cast<size_t>(kPadOut[l]) * 4);
        s->pw.wf[l] = reinterpret_cast<uint16_t*>(q);
        q += align256(fragment_elements(l) * 2);
        if (l > 0) {
            s->pw.wq[l] = reinterpret_cast<uint8_t*>(q);
            q += align256(static_cast<size_t>(kPadOut[l]) * kPadIn[l]);
            s->pw.wscale[l] = reinterpret_cast<uint8_t*>(q);
            q += align256(kPadOut[l]);
        }
    }
    e = launch_init_params(s->params, seed, nullptr);
    if (e == hipSuccess) e = hipMemsetAsync(s->adam_m, 0, pbytes, nullptr);
    if (e == hipSuccess) e = hipMemsetAsync(s->adam_v, 0, pbytes, nullptr);
    if (e == hipSuccess) e = launch_pad_weights(s->params, s->pw, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) { mlgpu_surrogate_destroy(s); return hip_fail(e, "surrogate init"); }
    *out = s;
    return MLGPU_OK;
}

void mlgpu_surrogate_destroy(mlgpu_surrogate* s) {
    if (!s) return;
    (void)hipDeviceSynchronize();
    if (s->ws) (void)hipFree(s->ws);
    if (s->hy) (void)hipFree(s->hy);
    if (s->hy_count_host) (void)hipHostFree(s->hy_count_host);
    if (s->weights_mem) (void)hipFree(s->weights_mem);
    if (s->params) (void)hipFree(s->params);
    delete s;
}

int mlgpu_surrogate_get_params(const mlgpu_surrogate* s, float* dst, void* stream) {
    if (!s || !dst) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    const hipError_t e = hipMemcpyAsync(dst, s->params, static_cast<size_t>(kNumParams) * 4, hipMemcpyDefault,
                                        static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "copy parameters");
    return MLGPU_OK;
}

int mlgpu_surrogate_set_params(mlgpu_surrogate* s, const float* src, void* stream) {
    if (!s || !src) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lock(s->mutex);
    const size_t pbytes = static_cast<size_t>(kNumParams) * 4;
    hipError_t e = hipMemcpyAsync(s->params, src, pbytes, hipMemcpyDefault, st);
    if (e == hipSuccess) e = hipMemsetAsync(s->adam_m, 0, pbytes, st);
    if (e == hipSuccess) e = hipMemsetAsync(s->adam_v, 0, pbytes, st);
    if (e == hipSuccess) e = launch_pad_weights(s->params, s->pw, st);
    if (e != hipSuccess) return hip_fail(e, "set parameters");
    s->step = 0;
    return MLGPU_OK;
}

int mlgpu_surrogate_set_inference(mlgpu_surrogate* s, int32_t precision) {
    if (!s || (precision != MLGPU_SURROGATE_BF16 && precision != MLGPU_SURROGATE_FP8)) {
        set_last_error("null surrogate or unknown precision");
        return MLGPU_ERR_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(s->mutex);
    s->fp8 = precision == MLGPU_SURROGATE_FP8;
    return MLGPU_OK;
}

int32_t mlgpu_surrogate_get_inference(const mlgpu_surrogate* s) {
    return s && s->fp8 ? MLGPU_SURROGATE_FP8 : MLGPU_SURROGATE_BF16;
}

int mlgpu_surrogate_forward(mlgpu_surrogate* s, const float* features, int64_t n, float* out, void* stream) {
    int rc = check_batch(s, features, n);
    if (rc) return rc;
    if (!out) { set_last_error("null output"); return MLGPU_ERR_ARGUMENT; }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    FusedArgs a = fused_args(s);
    a.features = features;
    a.out = out;
    a.sample = false;
    const hipError_t e = launch_fused(a, n, num_cus_of(s->device), st);
    if (e != hipSuccess) return hip_fail(e, "surrogate forward");
    return MLGPU_OK;
}

int mlgpu_surrogate_gradients(mlgpu_surrogate* s, const float* features, const float* targets, int64_t n,
                              float* grads, float* loss, void* stream) {
    int rc = check_batch(s, features, n);
    if (rc) return rc;
    if (!targets) { set_last_error("null targets"); return MLGPU_ERR_ARGUMENT; }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lock(s->mutex);
    const int64_t rows = padded_rows(n);
    if ((rc = ensure_workspace(s, rows))) return rc;
    hipError_t e = launch_pack_features(features, n, rows, s->x0, st);
    if (e == hipSuccess) e = backward(s, features, targets, n, rows, st);
    if (e == hipSuccess && grads)
        e = hipMemcpyAsync(grads, s->grads, static_cast<size_t>(kNumParams) * 4, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, s->loss, 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "surrogate gradients");
    return MLGPU_OK;
}

int mlgpu_surrogate_train_step(mlgpu_surrogate* s, const float* features, const float* targets, int64_t n,
                               float learning_rate, float* loss, void* stream) {
    int rc = check_batch(s, features, n);
    if (rc) return rc;
    if (!targets || !(learning_rate > 0.0f) || !std::isfinite(learning_rate)) {
        set_last_error("null targets or bad learning rate");
        return MLGPU_ERR_ARGUMENT;
    }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lock(s->mutex);
    const int64_t rows = padded_rows(n);
    if ((rc = ensure_workspace(s, rows))) return rc;
    hipError_t e = launch_pack_features(features, n, rows, s->x0, st);
    if (e == hipSuccess) e = backward(s, features, targets, n, rows, st);
    if (e == hipSuccess) {
        s->step++;
        AdamArgs a{};
        a.params = s->params;
        a.grads = s->grads;
        a.m = s->adam_m;
        a.v = s->adam_v;
        a.lr = learning_rate;
        a.beta1 = 0.9f;
        a.beta2 = 0.999f;
        a.eps = 1e-8f;
        a.c1 = static_cast<float>(1.0 / (1.0 - std::pow(0.9, static_cast<double>(s->step))));
        a.c2 = static_cast<float>(1.0 / (1.0 - std::pow(0.999, static_cast<double>(s->step))));
        e = launch_adam(a, st);
    }
    if (e == hipSuccess) e = launch_pad_weights(s->params, s->pw, st);
    if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, s->loss, 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "surrogate train step");
    return MLGPU_OK;
}

int mlgpu_surrogate_set_survival_scale(mlgpu_surrogate* s, float scale) {
    if (!s) { set_last_error("null surrogate"); return MLGPU_ERR_ARGUMENT; }
    if (!(scale > 0.0f) || !std::isfinite(scale)) {
        set_last_error("survival scale must be finite and > 0");
        return MLGPU_ERR_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(s->mutex);
    s->survival_scale = scale;
    return MLGPU_OK;
}

int mlgpu_surrogate_set_region(mlgpu_surrogate* s, const mlgpu_surrogate_domain* dom, const mlgpu_hex_cells* cells) {
    if (!s) { set_last_error("null surrogate"); return MLGPU_ERR_ARGUMENT; }
    std::lock_guard<std::mutex> lock(s->mutex);
    if (!cells) {
        s->region_on = false;
        return MLGPU_OK;
    }
    if (!domain_ok(dom) || !cells_ok(*cells)) {
        set_last_error("bad domain or cells");
        return MLGPU_ERR_ARGUMENT;
    }
    s->region_dom = *dom;
    s->region_cells = *cells;
    s->region_on = true;
    return MLGPU_OK;
}

int mlgpu_surrogate_prepare(const mlgpu_surrogate_domain* dom, const mlgpu_photons* ph, const mlgpu_results* r,
                            const double* poses, const mlgpu_results* focal, const double* focal_poses, int64_t n,
                            float* features, float* targets, void* stream) {
    if (!domain_ok(dom) || !ph || !features || n < 0 || (r && (!targets || !poses)) || (focal && (!r || !focal_poses))) {
        set_last_error("bad domain or null argument");
        return MLGPU_ERR_ARGUMENT;
    }
    const hipError_t e = launch_prepare(*dom, *ph, r, poses, focal, focal_poses, n, features, targets,
                                        static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "surrogate prepare");
    return MLGPU_OK;
}

int mlgpu_propagate_hybrid(const mlgpu_scene* scene, mlgpu_surrogate* s, const mlgpu_surrogate_domain* dom,
                           const mlgpu_hex_cells* cells, int32_t world_frame, const mlgpu_photons* ph, int64_t n,
                           const mlgpu_config* cfg, const mlgpu_hybrid_config* hc, const mlgpu_histogram* hist,
                           int64_t* num_fallback, void* stream) {
    if (!scene || !s || !domain_ok(dom) || !cells || !cells_ok(*cells) || !ph || !cfg || !hc ||
        !hist || !hist->counts || n < 0) {
        set_last_error("bad domain or null argument");
        return MLGPU_ERR_ARGUMENT;
    }
    if (num_fallback) *num_fallback = 0;
    if (n == 0) return MLGPU_OK;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lock(s->mutex);
    int rc;
    if ((rc = ensure_hybrid(s))) return rc;
    const int cus = num_cus_of(s->device);
    int64_t fallbacks = 0;
    for (int64_t c0 = 0; c0 < n; c0 += kHybridChunk) {
        const int64_t m = std::min(kHybridChunk, n - c0);
        const mlgpu_photons chunk{ph->support_x + c0, ph->support_y + c0, ph->support_z + c0, ph->direction_x + c0,
                                  ph->direction_y + c0, ph->direction_z + c0, ph->wavelength + c0};
        hipError_t e = hipMemsetAsync(s->hy_count, 0, sizeof(int32_t), st);
        if (e == hipSuccess) {
            FusedArgs a = fused_args(s);
            a.ph = chunk;
            a.sample = true;
            a.s.dom = *dom;
            a.s.cells = *cells;
            a.s.hist = *hist;
            a.s.hist.photon_id_offset = hist->photon_id_offset + static_cast<uint64_t>(c0);
            a.s.max_entropy = hc->max_entropy;
            a.s.seed = hc->seed;
            a.s.id_offset = cfg->photon_id_offset + static_cast<uint64_t>(c0);
            a.s.fallback = s->hy_fallback;
            a.s.fallback_count = s->hy_count;
            e = launch_fused(a, m, cus, st);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(s->hy_count_host, s->hy_count, sizeof(int32_t), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "hybrid surrogate pass");
        const int64_t k = *s->hy_count_host;
        if (k == 0) continue;
        fallbacks += k;
        // k (<= m) is known here: one thread per fallback photon
        e = launch_gather_photons(chunk, s->hy_fallback, s->hy_count, k, s->hy_photons, st);
        if (e != hipSuccess) return hip_fail(e, "hybrid gather");
        const mlgpu_photons fb{s->hy_photons.support_x, s->hy_photons.support_y, s->hy_photons.support_z,
                               s->hy_photons.direction_x, s->hy_photons.direction_y, s->hy_photons.direction_z,
                               s->hy_photons.wavelength};
        // every fallback photon is propagated and binned under its own id
        // (photon_id_offset + its index in the call), whatever slot the
        // fallback list gave it: its draws and jitter are those of the same
        // photon under mlgpu_propagate, and the histogram does not depend on
        // the order in which the list filled
        mlgpu_config fcfg = *cfg;
        fcfg.photon_id_offset = cfg->photon_id_offset + static_cast<uint64_t>(c0);
        if ((rc = mlgpu::propagate_gathered(scene, world_frame, &fb, k, &fcfg, &s->hy_results, s->hy_fallback,
                                            stream)))
            return rc;
        mlgpu_histogram fh = *hist;
        fh.photon_id_offset = hist->photon_id_offset + static_cast<uint64_t>(c0);
        if ((rc = mlgpu::histogram_gathered(&s->hy_results, k, &fh, s->hy_fallback, stream))) return rc;
    }
    if (num_fallback) *num_fallback = fallbacks;
    return MLGPU_OK;
}

}  // extern "C"
