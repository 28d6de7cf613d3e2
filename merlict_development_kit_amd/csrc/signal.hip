// The sensor signal chain on the GPU: per-sensor photon pipelines sorted by
// arrival time (merlict_signal_processing/PipelinePhoton.cpp:25-66), the
// photo-electric converter (PhotoElectricConverter.cpp:10-95: quantum
// efficiency, crosstalk chain, dark-rate accidental pulses) and the
// single-photon pulse extraction (pulse_extraction.cpp:9-30).
//
// The night-sky background is the step between the pipelines and the
// converter (Injector.cpp:11-107); sensor s draws its background photons
// from Philox stream 9 of (seed, s).
//
// The reference draws every random number of an event from one generator
// in sensor order, which serialises the chain.  Here each photon of the
// sorted pipelines (its conversion and crosstalk chain), each sensor (its
// accidental pulses) and each pulse (its extraction jitter) owns a Philox
// stream, consumed one 24-bit uniform at a time in the reference's order, so
// every photon, sensor and pulse is a lane of its own.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "kernels.hpp"
#include "propagate_core.hpp"
#include "pool.hpp"
#include "signal.hpp"
#include "stream_rng.hpp"

namespace mlgpu {
namespace {

constexpr uint32_t kConverterStream = 6;
constexpr uint32_t kExtractionStream = 7;
constexpr uint32_t kAccidentalStream = 8;
constexpr uint32_t kNsbStream = 9;
constexpr int32_t kNightSky = -100;    // simulation_truth.h:15
constexpr double kMaxMeanPerSensor = 1e7;
constexpr int32_t kCrosstalk = -202;   // simulation_truth.h:17
constexpr int32_t kAccidental = -201;  // simulation_truth.h:16
constexpr int32_t kNumTimeSlices = 100;  // pulse_extraction.h:13
constexpr int32_t kMaxChain = 1 << 20;

// Func1::evaluate (merlict/function/Func1.cpp:66-71) on a table sorted by
// x with the slopes computed before sorting; false outside [x0, xN)
__device__ bool table_eval(const FuncPoint* p, int32_t n, double x, double& y) {
    if (!(x >= p[0].x) || !(x < p[n - 1].x)) return false;
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (p[mid].x > x) hi = mid; else lo = mid + 1;
    }
    const FuncPoint p0 = p[lo - 1];
    y = p0.slope * x + p0.y - p0.slope * p0.x;
    return true;
}

// order-preserving map of a double to an unsigned key
__device__ __forceinline__ uint64_t time_key(double t) {
    const uint64_t b = __double_as_longlong(t);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// photon i is detected: a sensor in range and no error flag
struct Detected {
    const int32_t* sensor;
    const uint8_t* error;
    int32_t num_sensors;
    __device__ bool operator()(const int32_t i) const {
        const int32_t s = sensor[i];
        return s >= 0 && s < num_sensors && error[i] == 0;
    }
};

// time keys of the detected photons (index order)
__global__ void keys_kernel(const int32_t* __restrict__ index, const int64_t m, const mlgpu_results r,
                            uint64_t* __restrict__ tkey) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= m) return;
    tkey[k] = time_key(r.time_of_flight[index[k]]);
}

// sensor of each photon in time order
__global__ void sensor_of_kernel(const int32_t* __restrict__ order, const mlgpu_results r, const int64_t m,
                                 uint32_t* __restrict__ sk) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < m) sk[k] = static_cast<uint32_t>(r.sensor_id[order[k]]);
}

// offsets[s] = first position of sensor s in the sorted keys (s <= num_sensors)
__global__ void offsets_kernel(const uint32_t* __restrict__ sorted_sensor, const int64_t m, const int32_t num_sensors,
                               int64_t* __restrict__ offsets) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s > num_sensors) return;
    int64_t lo = 0, hi = m;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (static_cast<int64_t>(sorted_sensor[mid]) < s) lo = mid + 1; else hi = mid;
    }
    offsets[s] = lo;
}

__global__ void gather_kernel(const int32_t* __restrict__ index, const int64_t m, const mlgpu_results r,
                              const double* __restrict__ wavelength, const int32_t* __restrict__ truth,
                              const mlgpu_pipelines out) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int32_t i = index[k];
    out.arrival_time[k] = r.time_of_flight[i];
    out.wavelength[k] = wavelength[i];
    out.truth_id[k] = truth ? truth[i] : i;
}

// The converter's random numbers: photon j of the sorted pipelines draws
// its quantum-efficiency uniform and its crosstalk chain from Philox stream
// 6 of (seed, j); sensor s draws its accidental pulses from stream 8 of
// (seed, s).  Every photon and every sensor is then independent.

// pulses of photon j (0 or 1 + its crosstalk chain); *bad for a wavelength
// outside the table (Func1 throws std::out_of_range, Limits.cpp:37-46)
__device__ int32_t photon_pulses(const mlgpu_pipelines& p, const int64_t j, const FuncPoint* qe, const int32_t qe_n,
                                 const double crosstalk, const uint64_t seed, bool& bad) {
    StreamRng rng;
    rng.init(seed, static_cast<uint64_t>(j), kConverterStream);
    double q;
    bad = !table_eval(qe, qe_n, p.wavelength[j], q);
    if (bad || !(q >= rng.uniform())) return 0;
    // Converter::add_pulse: a crosstalk pulse follows while p >= u (the
    // chain is cut at kMaxChain, which p < 1 - 1e-6 makes unreachable in
    // practice)
    int32_t k = 1;
    for (int32_t c = 0; c < kMaxChain && crosstalk >= rng.uniform(); c++) k++;
    return k;
}

__global__ void photon_count_kernel(const mlgpu_pipelines p, const int64_t m, const FuncPoint* __restrict__ qe,
                                    const int32_t qe_n, const double crosstalk, const uint64_t seed,
                                    int64_t* __restrict__ count, int32_t* __restrict__ bad) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= m) return;
    bool b = false;
    count[j] = photon_pulses(p, j, qe, qe_n, crosstalk, seed, b);
    if (b) atomicOr(bad, 1);
}

// Converter::add_accidental_pulse: the sensor's accidental pulses, each
// with its crosstalk chain; with Emit they are written from `at` on
template <bool Emit>
__device__ int64_t accidentals(const uint64_t seed, const int32_t s, const double dark_rate, const double crosstalk,
                               const double exposure_time, const mlgpu_pulses& out, const int64_t at) {
    StreamRng rng;
    rng.init(seed, static_cast<uint64_t>(s), kAccidentalStream);
    int64_t k = 0;
    double sum = rng.expovariate(dark_rate);
    while (sum < exposure_time) {
        const double next = rng.expovariate(dark_rate);
        if (Emit) {
            out.arrival_time[at + k] = sum;
            out.truth_id[at + k] = kAccidental;
        }
        k++;
        for (int32_t c = 0; c < kMaxChain && crosstalk >= rng.uniform(); c++) {
            if (Emit) {
                out.arrival_time[at + k] = sum;
                out.truth_id[at + k] = kCrosstalk;
            }
            k++;
        }
        sum += next;
    }
    return k;
}

// per sensor: its photons' pulses (from the scan of the photon counts) plus
// its accidental pulses
__global__ void sensor_count_kernel(const mlgpu_pipelines p, const int32_t num_sensors,
                                    const int64_t* __restrict__ photon_scan, const uint64_t seed,
                                    const double dark_rate, const double crosstalk, const double exposure_time,
                                    int64_t* __restrict__ sensor_count) {
    const int32_t s = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= num_sensors) return;
    const int64_t own = photon_scan[p.offsets[s + 1]] - photon_scan[p.offsets[s]];
    sensor_count[s] = own + accidentals<false>(seed, s, dark_rate, crosstalk, exposure_time, mlgpu_pulses{}, 0);
}

// sensor of pipeline entry j (binary search of the offsets)
__device__ __forceinline__ int32_t sensor_of(const int64_t* offsets, const int32_t num_sensors, const int64_t j) {
    int32_t lo = 0, hi = num_sensors - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (offsets[mid] <= j) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void photon_emit_kernel(const mlgpu_pipelines p, const int64_t m, const int32_t num_sensors,
                                   const FuncPoint* __restrict__ qe, const int32_t qe_n, const double crosstalk,
                                   const uint64_t seed, const int64_t* __restrict__ photon_scan,
                                   const mlgpu_pulses out) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= m) return;
    bool b = false;
    const int32_t k = photon_pulses(p, j, qe, qe_n, crosstalk, seed, b);
    if (k == 0) return;
    const int32_t s = sensor_of(p.offsets, num_sensors, j);
    const int64_t at = out.offsets[s] + photon_scan[j] - photon_scan[p.offsets[s]];
    out.arrival_time[at] = p.arrival_time[j];
    out.truth_id[at] = p.truth_id[j];
    for (int32_t c = 1; c < k; c++) {
        out.arrival_time[at + c] = p.arrival_time[j];
        out.truth_id[at + c] = kCrosstalk;
    }
}

__global__ void accidental_emit_kernel(const mlgpu_pipelines p, const int32_t num_sensors,
                                       const int64_t* __restrict__ photon_scan, const uint64_t seed,
                                       const double dark_rate, const double crosstalk, const double exposure_time,
                                       const mlgpu_pulses out) {
    const int32_t s = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= num_sensors) return;
    const int64_t at = out.offsets[s] + photon_scan[p.offsets[s + 1]] - photon_scan[p.offsets[s]];
    (void)accidentals<true>(seed, s, dark_rate, crosstalk, exposure_time, out, at);
}

// extract_pulses (pulse_extraction.cpp:9-30) for pulse k, with the normal
// from Philox stream 7 of (seed, k): slice = round((t + N(0, std)) /
// duration), -1 outside [0, 100)
__global__ void extract_kernel(const mlgpu_pulses p, const int64_t count, const double duration,
                               const double std_dev, const uint64_t seed, int16_t* __restrict__ slice) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= count) return;
    StreamRng rng;
    rng.init(seed, static_cast<uint64_t>(k), kExtractionStream);
    const double t = p.arrival_time[k] + rng.normal(0.0, std_dev);
    const int32_t sl = static_cast<int32_t>(round(t / duration));
    slice[k] = (sl >= 0 && sl < kNumTimeSlices) ? static_cast<int16_t>(sl) : static_cast<int16_t>(-1);
}

// ------------------------------------------------ night-sky background

__device__ __forceinline__ double key_time(uint64_t k) {
    const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double(static_cast<long long>(b));
}

// earliest and latest arrival time (as order-preserving keys)
__global__ __launch_bounds__(256) void time_range_kernel(const double* __restrict__ t, const int64_t m,
                                                         unsigned long long* __restrict__ range) {
    __shared__ unsigned long long lo_s[256], hi_s[256];
    unsigned long long lo = ~0ull, hi = 0ull;
    for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < m;
         j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const unsigned long long k = time_key(t[j]);
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    }
    lo_s[threadIdx.x] = lo;
    hi_s[threadIdx.x] = hi;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (static_cast<int>(threadIdx.x) < off) {
            const unsigned long long a = lo_s[threadIdx.x + off], b = hi_s[threadIdx.x + off];
            if (a < lo_s[threadIdx.x]) lo_s[threadIdx.x] = a;
            if (b > hi_s[threadIdx.x]) hi_s[threadIdx.x] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMin(&range[0], lo_s[0]);
        atomicMax(&range[1], hi_s[0]);
    }
}

// Injector.cpp:26-45: min starts at the largest double, max at the
// smallest positive one
struct ModeEdges {
    double lo, step;
    uint32_t count;  // number of bin edges
    __device__ double edge(uint32_t i) const { return lo + step * static_cast<double>(i); }
};

__device__ __forceinline__ ModeEdges mode_edges(const unsigned long long* range, const uint32_t count) {
    double lo = key_time(range[0]), hi = key_time(range[1]);
    if (!(lo < 1.7976931348623157e308)) lo = 1.7976931348623157e308;
    if (!(hi > 2.2250738585072014e-308)) hi = 2.2250738585072014e-308;
    ModeEdges e;
    e.lo = lo;
    e.count = count;
    e.step = (hi - lo) / static_cast<double>(count - 1u);  // numeric::linspace (numeric.cpp:35-44)
    return e;
}

// Histogram1::fill_in (Histogram1.cpp:29-47): the bin below the first edge
// greater than t; -1 for under- and overflow
__device__ __forceinline__ int64_t mode_bin(const ModeEdges& e, const double t) {
    const int64_t last = static_cast<int64_t>(e.count) - 1;
    int64_t k = 0;
    if (e.step > 0.0) {
        const double g = floor((t - e.lo) / e.step);
        k = g < 0.0 ? 0 : (g > static_cast<double>(last) ? last : static_cast<int64_t>(g));
    }
    while (k < last && e.edge(static_cast<uint32_t>(k + 1)) <= t) k++;
    while (k >= 0 && e.edge(static_cast<uint32_t>(k)) > t) k--;
    return (k < 0 || k >= last) ? -1 : k;
}

// Up to kModeLdsBins bins each block counts in LDS and adds its counts
// once (the photons crowd into a few bins around the peak, where global
// atomics would serialise); beyond, global atomics.
constexpr uint32_t kModeLdsBins = 12288;

__global__ __launch_bounds__(256) void mode_histogram_kernel(const double* __restrict__ t, const int64_t m,
                                                             const unsigned long long* __restrict__ range,
                                                             const uint32_t count, uint32_t* __restrict__ bins) {
    __shared__ uint32_t local[kModeLdsBins];
    const ModeEdges e = mode_edges(range, count);
    const uint32_t nb = count - 1u;
    const bool lds = nb <= kModeLdsBins;
    if (lds)
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) local[i] = 0u;
    __syncthreads();
    for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < m;
         j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t b = mode_bin(e, t[j]);
        if (b < 0) continue;
        if (lds) atomicAdd(&local[b], 1u); else atomicAdd(&bins[b], 1u);
    }
    __syncthreads();
    if (lds)
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
            if (local[i]) atomicAdd(&bins[i], local[i]);
}

// Histogram1::mode / arg_max (Histogram1.cpp:49-65): the first bin of the
// largest count; start = mode - exposure / 2 (Injector.cpp:68-71)
__global__ __launch_bounds__(1024) void mode_start_kernel(const uint32_t* __restrict__ bins,
                                                          const unsigned long long* __restrict__ range,
                                                          const uint32_t count, const int64_t m,
                                                          const double exposure_time, double* __restrict__ start) {
    __shared__ uint32_t best_c[1024];
    __shared__ uint32_t best_i[1024];
    uint32_t c = 0u, idx = 0xffffffffu;
    const uint32_t nb = m > 0 ? count - 1u : 0u;
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
        if (bins[i] > c) { c = bins[i]; idx = i; }
    best_c[threadIdx.x] = c;
    best_i[threadIdx.x] = idx;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (static_cast<int>(threadIdx.x) < off) {
            const uint32_t c2 = best_c[threadIdx.x + off], i2 = best_i[threadIdx.x + off];
            if (c2 > best_c[threadIdx.x] || (c2 == best_c[threadIdx.x] && i2 < best_i[threadIdx.x])) {
                best_c[threadIdx.x] = c2;
                best_i[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double mode = 0.0;
        if (m > 0) {
            const ModeEdges e = mode_edges(range, count);
            const uint32_t i = best_c[0] > 0u ? best_i[0] : 0u;
            mode = 0.5 * (e.edge(i + 1u) + e.edge(i));
        }
        *start = mode - 0.5 * exposure_time;
    }
}

// the Poisson process of sensor s (Injector.cpp:73-96); Emit writes the
// arrival times (start + running sum) and the wavelengths from `at` on
template <bool Emit>
__device__ int64_t background(const uint64_t seed, const int32_t s, const double rate, const double exposure_time,
                              const double start, const FuncPoint* cdf, const int32_t cdf_n, const double cdf_total,
                              double* time, double* lambda, const int64_t at, bool& bad) {
    StreamRng rng;
    rng.init(seed, static_cast<uint64_t>(s), kNsbStream);
    int64_t k = 0;
    double sum = rng.expovariate(rate);
    while (sum < exposure_time) {
        const double next = rng.expovariate(rate);
        if (Emit) time[at + k] = start + sum;
        k++;
        sum += next;
    }
    if (Emit) {
        // SamplesFromDistribution::draw: the inverse cdf at u * total
        for (int64_t q = 0; q < k; q++) {
            double y;
            if (!table_eval(cdf, cdf_n, rng.uniform() * cdf_total, y)) { bad = true; y = 0.0; }
            lambda[at + q] = y;
        }
    }
    return k;
}

__device__ __forceinline__ bool sensor_rate(const float* eff, const int32_t s, const int32_t num_sensors,
                                            const double rate, const double exposure_time, double& r) {
    r = rate * static_cast<double>(eff[s]) / static_cast<double>(num_sensors);
    return r >= 0.0 && r * exposure_time <= kMaxMeanPerSensor;
}

__global__ void nsb_count_kernel(const int32_t num_sensors, const float* __restrict__ eff, const double rate,
                                 const double exposure_time, const uint64_t seed, int64_t* __restrict__ count,
                                 int32_t* __restrict__ bad) {
    const int32_t s = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= num_sensors) return;
    double r;
    if (!sensor_rate(eff, s, num_sensors, rate, exposure_time, r)) {
        count[s] = 0;
        atomicOr(bad, 1);
        return;
    }
    bool b = false;
    count[s] = background<false>(seed, s, r, exposure_time, 0.0, nullptr, 0, 0.0, nullptr, nullptr, 0, b);
}

__global__ void nsb_emit_kernel(const int32_t num_sensors, const float* __restrict__ eff, const double rate,
                                const double exposure_time, const uint64_t seed, const double* __restrict__ start,
                                const FuncPoint* __restrict__ cdf, const int32_t cdf_n, const double cdf_total,
                                const int64_t* __restrict__ offsets, double* __restrict__ time,
                                double* __restrict__ lambda, int32_t* __restrict__ bad) {
    const int32_t s = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= num_sensors) return;
    double r;
    if (!sensor_rate(eff, s, num_sensors, rate, exposure_time, r)) return;
    bool b = false;
    (void)background<true>(seed, s, r, exposure_time, *start, cdf, cdf_n, cdf_total, time, lambda, offsets[s], b);
    if (b) atomicOr(bad, 1);
}

// number of entries of t[lo, hi) below x (Strict) or not above x
template <bool Strict>
__device__ __forceinline__ int64_t rank_in(const double* t, int64_t lo, int64_t hi, const double x) {
    const int64_t base = lo;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (Strict ? (t[mid] < x) : (t[mid] <= x)) lo = mid + 1; else hi = mid;
    }
    return lo - base;
}

// merged position of pipeline photon j: its rank in its pipeline + the
// background photons of its sensor that arrive strictly earlier
__global__ void merge_photons_kernel(const mlgpu_pipelines in, const int64_t m, const int32_t num_sensors,
                                     const int64_t* __restrict__ nsb_offsets, const double* __restrict__ nsb_time,
                                     const double* __restrict__ start, const mlgpu_pipelines out) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int32_t s = sensor_of(in.offsets, num_sensors, j);
    const double t = in.arrival_time[j];
    const int64_t pos = in.offsets[s] + nsb_offsets[s] + (j - in.offsets[s]) +
                        rank_in<true>(nsb_time, nsb_offsets[s], nsb_offsets[s + 1], t);
    out.arrival_time[pos] = t - *start;
    out.wavelength[pos] = in.wavelength[j];
    out.truth_id[pos] = in.truth_id[j];
}

// merged position of background photon q: its rank among its sensor's
// background photons + the sensor's pipeline photons not later than it
__global__ void merge_background_kernel(const mlgpu_pipelines in, const int32_t num_sensors, const int64_t total,
                                        const int64_t* __restrict__ nsb_offsets, const double* __restrict__ nsb_time,
                                        const double* __restrict__ nsb_lambda, const double* __restrict__ start,
                                        const mlgpu_pipelines out) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= total) return;
    const int32_t s = sensor_of(nsb_offsets, num_sensors, q);
    const double t = nsb_time[q];
    const int64_t pos = in.offsets[s] + nsb_offsets[s] + (q - nsb_offsets[s]) +
                        rank_in<false>(in.arrival_time, in.offsets[s], in.offsets[s + 1], t);
    out.arrival_time[pos] = t - *start;
    out.wavelength[pos] = nsb_lambda[q];
    out.truth_id[pos] = kNightSky;
}

__global__ void add_offsets_kernel(const int64_t* __restrict__ a, const int64_t* __restrict__ b, const int32_t n,
                                   int64_t* __restrict__ out) {
    const int32_t s = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (s <= n) out[s] = a[s] + b[s];
}

// ------------------------------------------------------ photon stream

constexpr uint8_t kNextChannelMarker = 255;  // PhotonStream.h:15; FACT's NEXT_PIXEL_MARKER
constexpr int32_t kFactSliceOffset = 30;     // merlict_fact/apps/photon_stream.h:10

struct Kept {
    const int16_t* slice;
    __device__ int64_t operator()(const int64_t k) const { return slice[k] >= 0 ? 1 : 0; }
};

// kept pulse k of channel s goes to symbol kscan[k] + s (the kept pulses
// of the channels before it and their markers come first)
__global__ void stream_pulses_kernel(const mlgpu_pulses p, const int16_t* __restrict__ slice, const int64_t count,
                                     const int32_t num_sensors, const int64_t* __restrict__ kscan,
                                     uint8_t* __restrict__ symbols, int32_t* __restrict__ truth) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= count || slice[k] < 0) return;
    const int32_t s = sensor_of(p.offsets, num_sensors, k);
    symbols[kscan[k] + s] = static_cast<uint8_t>(slice[k]);
    if (truth) truth[kscan[k]] = p.truth_id[k];
}

// the marker after channel s (s < num_markers: every channel but the last,
// or every channel)
__global__ void stream_markers_kernel(const mlgpu_pulses p, const int32_t num_markers,
                                      const int64_t* __restrict__ kscan, uint8_t* __restrict__ symbols) {
    const int32_t s = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= num_markers) return;
    symbols[kscan[p.offsets[s + 1]] + s] = kNextChannelMarker;
}

// FACT's photon stream (merlict_fact/apps/fact_propagation.cpp:34-94): the
// slice of pulse k relative to the median arrival time of the event, kept
// within [30, 130), else -1
__global__ void fact_slices_kernel(const double* __restrict__ t, const int64_t count,
                                   const double* __restrict__ median, const double trigger_time,
                                   const double slice_duration, int16_t* __restrict__ slice) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const double x = round((t[k] - *median + trigger_time) / slice_duration);
    int16_t out = -1;
    if (x >= static_cast<double>(kFactSliceOffset) && x < static_cast<double>(kFactSliceOffset + kNumTimeSlices))
        out = static_cast<int16_t>(x);
    slice[k] = out;
}

__global__ void time_keys_kernel(const double* __restrict__ t, const int64_t count, uint64_t* __restrict__ keys) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < count) keys[k] = time_key(t[k]);
}

// the median arrival time: the sorted time key at floor(count / 2)
__global__ void key_at_kernel(const uint64_t* __restrict__ sorted, const int64_t at, double* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = key_time(sorted[at]);
}

// Radix selection of the k-th smallest of m keys, 8 bits a pass from the
// top: the histogram of the next digit over the keys that share the bits
// chosen so far, then the bucket that holds rank k.  After eight passes
// `prefix` is the key itself.
struct SelectState {
    unsigned long long prefix, mask;
    long long k;
    unsigned int hist[256];
};

__global__ void select_init_kernel(SelectState* st, const long long k) {
    st->hist[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
        st->prefix = 0ull;
        st->mask = 0ull;
        st->k = k;
    }
}

__global__ __launch_bounds__(256) void select_hist_kernel(const uint64_t* __restrict__ keys, const int64_t m,
                                                          SelectState* __restrict__ st, const int shift) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const unsigned long long prefix = st->prefix, mask = st->mask;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < m;
         i += static_cast<int64_t>(gridDim.x) * 256) {
        const unsigned long long key = keys[i];
        if ((key & mask) == prefix) atomicAdd(&h[(key >> shift) & 255ull], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&st->hist[threadIdx.x], h[threadIdx.x]);
}

__global__ void select_pick_kernel(SelectState* st, const int shift) {
    // inclusive prefix sums of the 256 counts in LDS; the thread whose
    // bucket holds rank k (the last bucket if none does) records it
    __shared__ long long c[256];
    const int t = static_cast<int>(threadIdx.x);
    const long long k = st->k;
    const long long v = st->hist[t];
    st->hist[t] = 0u;
    c[t] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const long long x = t >= off ? c[t - off] : 0;
        __syncthreads();
        c[t] += x;
        __syncthreads();
    }
    const long long incl = c[t], excl = incl - v;
    if (excl <= k && (k < incl || t == 255)) {
        st->k = k - excl;
        st->prefix |= static_cast<unsigned long long>(t) << shift;
        st->mask |= 255ull << shift;
    }
}

size_t align_up(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

}  // namespace

hipError_t build_pipelines(const mlgpu_results& r, const double* wavelength, const int32_t* truth, int64_t n,
                           int32_t num_sensors, const mlgpu_pipelines& out, int64_t* num_photons,
                           hipStream_t stream, int* too_small) {
    *num_photons = 0;
    *too_small = 0;
    const size_t cap = static_cast<size_t>(n);
    // keys and values, doubled for the sort's alternate buffers
    size_t temp_bytes = 0, temp2 = 0, temp3 = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, temp_bytes, static_cast<uint64_t*>(nullptr),
                                             static_cast<uint64_t*>(nullptr), static_cast<int32_t*>(nullptr),
                                             static_cast<int32_t*>(nullptr), cap, 0, 64, stream);
    if (e != hipSuccess) return e;
    e = rocprim::radix_sort_pairs(nullptr, temp2, static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                  static_cast<int32_t*>(nullptr), static_cast<int32_t*>(nullptr), cap, 0, 32, stream);
    if (e != hipSuccess) return e;
    const Detected detected{r.sensor_id, r.error, num_sensors};
    e = rocprim::select(nullptr, temp3, rocprim::counting_iterator<int32_t>(0), static_cast<int32_t*>(nullptr),
                        static_cast<int32_t*>(nullptr), cap, detected, stream);
    if (e != hipSuccess) return e;
    temp_bytes = temp_bytes > temp2 ? temp_bytes : temp2;
    temp_bytes = temp_bytes > temp3 ? temp_bytes : temp3;
    const size_t bytes = 2 * align_up(cap * 4) + 2 * align_up(cap * 8) + 2 * align_up(cap * 4) + align_up(temp_bytes) +
                         256;
    PoolBlock block;
    e = block.get(bytes);
    if (e != hipSuccess) return e;
    char* mem = block.as();
    char* q = mem;
    auto take = [&](size_t b) { char* x = q; q += align_up(b); return x; };
    auto* skey = reinterpret_cast<uint32_t*>(take(cap * 4));
    auto* skey2 = reinterpret_cast<uint32_t*>(take(cap * 4));
    auto* tkey = reinterpret_cast<uint64_t*>(take(cap * 8));
    auto* tkey2 = reinterpret_cast<uint64_t*>(take(cap * 8));
    auto* idx = reinterpret_cast<int32_t*>(take(cap * 4));
    auto* idx2 = reinterpret_cast<int32_t*>(take(cap * 4));
    void* temp = take(temp_bytes);
    auto* cnt = reinterpret_cast<int32_t*>(take(4));
    int32_t m32 = 0;
    {
        // the detected photons' indices in index order (an order-preserving
        // select), so equal arrival times keep photon order through the
        // stable sorts
        size_t tb = temp_bytes;
        e = rocprim::select(temp, tb, rocprim::counting_iterator<int32_t>(0), idx, cnt, cap, detected, stream);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&m32, cnt, 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    const size_t m = static_cast<size_t>(m32);
    *num_photons = m32;
    if (e == hipSuccess && static_cast<int64_t>(m) > out.capacity) *too_small = 1;
    if (e == hipSuccess && !*too_small && m > 0) {
        hipLaunchKernelGGL(keys_kernel, dim3(static_cast<uint32_t>((m + 255) / 256)), dim3(256), 0, stream, idx,
                           static_cast<int64_t>(m), r, tkey);
        e = hipGetLastError();
    }
    if (e == hipSuccess && !*too_small && m > 0) {
        // by time, then stably by sensor: (sensor, time, photon index)
        size_t tb = temp_bytes;
        e = rocprim::radix_sort_pairs(temp, tb, tkey, tkey2, idx, idx2, m, 0, 64, stream);
        // the sensor keys follow the time order through a gather by index:
        // sort (sensor of idx2[k], idx2[k]) stably by sensor
        if (e == hipSuccess) {
            // sensor key of each time-sorted photon
            hipLaunchKernelGGL(sensor_of_kernel, dim3(static_cast<uint32_t>((m + 255) / 256)), dim3(256), 0, stream,
                               idx2, r, static_cast<int64_t>(m), skey);
            e = hipGetLastError();
        }
        if (e == hipSuccess) {
            tb = temp_bytes;
            e = rocprim::radix_sort_pairs(temp, tb, skey, skey2, idx2, idx, m, 0, 32, stream);
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(offsets_kernel, dim3(static_cast<uint32_t>((num_sensors + 1 + 255) / 256)), dim3(256),
                               0, stream, skey2, static_cast<int64_t>(m), num_sensors, out.offsets);
            hipLaunchKernelGGL(gather_kernel, dim3(static_cast<uint32_t>((m + 255) / 256)), dim3(256), 0, stream,
                               idx, static_cast<int64_t>(m), r, wavelength, truth, out);
            e = hipGetLastError();
        }
    } else if (e == hipSuccess && !*too_small) {
        hipLaunchKernelGGL(offsets_kernel, dim3(static_cast<uint32_t>((num_sensors + 1 + 255) / 256)), dim3(256), 0,
                           stream, skey2, static_cast<int64_t>(0), num_sensors, out.offsets);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipStreamSynchronize(stream);  // the borrowed blocks are idle when returned
    return e;
}

hipError_t convert_pulses(const mlgpu_pipelines& p, int32_t num_sensors, int64_t num_photons, const FuncPoint* qe,
                          int32_t qe_n, double dark_rate, double crosstalk, double exposure_time, uint64_t seed,
                          const mlgpu_pulses& out, int64_t* num_pulses, int* bad_wavelength, int* too_small,
                          hipStream_t stream) {
    *num_pulses = 0;
    *bad_wavelength = 0;
    *too_small = 0;
    const size_t ns = static_cast<size_t>(num_sensors), m = static_cast<size_t>(num_photons);
    size_t scan_a = 0, scan_b = 0;
    hipError_t e = rocprim::exclusive_scan(nullptr, scan_a, static_cast<int64_t*>(nullptr),
                                           static_cast<int64_t*>(nullptr), int64_t{0}, m + 1,
                                           rocprim::plus<int64_t>(), stream);
    if (e == hipSuccess)
        e = rocprim::exclusive_scan(nullptr, scan_b, static_cast<int64_t*>(nullptr), static_cast<int64_t*>(nullptr),
                                    int64_t{0}, ns + 1, rocprim::plus<int64_t>(), stream);
    if (e != hipSuccess) return e;
    const size_t scan_bytes = scan_a > scan_b ? scan_a : scan_b;
    const size_t bytes = 2 * align_up((m + 1) * 8) + align_up((ns + 1) * 8) + align_up(scan_bytes) + 256;
    PoolBlock block;
    e = block.get(bytes);
    if (e != hipSuccess) return e;
    char* mem = block.as();
    char* q = mem;
    auto take = [&](size_t b) { char* x = q; q += align_up(b); return x; };
    auto* pcount = reinterpret_cast<int64_t*>(take((m + 1) * 8));
    auto* pscan = reinterpret_cast<int64_t*>(take((m + 1) * 8));
    auto* scount = reinterpret_cast<int64_t*>(take((ns + 1) * 8));
    void* temp = take(scan_bytes);
    auto* bad = reinterpret_cast<int32_t*>(take(4));
    int32_t host_bad = 0;
    int64_t total = 0;
    e = hipMemsetAsync(pcount, 0, (m + 1) * 8, stream);
    if (e == hipSuccess) e = hipMemsetAsync(scount, 0, (ns + 1) * 8, stream);
    if (e == hipSuccess) e = hipMemsetAsync(bad, 0, 4, stream);
    if (e == hipSuccess && m > 0) {
        hipLaunchKernelGGL(photon_count_kernel, dim3(static_cast<uint32_t>((m + 255) / 256)), dim3(256), 0, stream, p,
                           num_photons, qe, qe_n, crosstalk, seed, pcount, bad);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = rocprim::exclusive_scan(temp, scan_a, pcount, pscan, int64_t{0}, m + 1, rocprim::plus<int64_t>(), stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sensor_count_kernel, dim3(static_cast<uint32_t>((ns + 63) / 64)), dim3(64), 0, stream, p,
                           num_sensors, pscan, seed, dark_rate, crosstalk, exposure_time, scount);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = rocprim::exclusive_scan(temp, scan_b, scount, out.offsets, int64_t{0}, ns + 1, rocprim::plus<int64_t>(),
                                    stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&total, out.offsets + ns, 8, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&host_bad, bad, 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess) {
        *num_pulses = total;
        *bad_wavelength = host_bad;
        if (total > out.capacity) *too_small = 1;
    }
    if (e == hipSuccess && !*too_small && !host_bad && total > 0) {
        if (m > 0)
            hipLaunchKernelGGL(photon_emit_kernel, dim3(static_cast<uint32_t>((m + 255) / 256)), dim3(256), 0, stream,
                               p, num_photons, num_sensors, qe, qe_n, crosstalk, seed, pscan, out);
        hipLaunchKernelGGL(accidental_emit_kernel, dim3(static_cast<uint32_t>((ns + 63) / 64)), dim3(64), 0, stream,
                           p, num_sensors, pscan, seed, dark_rate, crosstalk, exposure_time, out);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
    }
    (void)hipStreamSynchronize(stream);  // the borrowed blocks are idle when returned
    return e;
}

hipError_t extract_pulses(const mlgpu_pulses& p, int64_t count, double duration, double std_dev, uint64_t seed,
                          int16_t* slice, hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(extract_kernel, dim3(static_cast<uint32_t>((count + 255) / 256)), dim3(256), 0, stream, p,
                       count, duration, std_dev, seed, slice);
    return hipGetLastError();
}

hipError_t inject_nsb(const mlgpu_pipelines& in, int32_t num_sensors, int64_t m, const float* efficiencies,
                      const FuncPoint* cdf, int32_t cdf_n, double cdf_total, double rate, double exposure_time,
                      uint64_t seed, const mlgpu_pipelines& out, int64_t* num_photons, double* start, int* bad,
                      int* too_small, hipStream_t stream) {
    *num_photons = 0;
    *start = 0.0;
    *bad = 0;
    *too_small = 0;
    const size_t ns = static_cast<size_t>(num_sensors);
    // Injector.cpp:52-55: 2 + sqrt(m) bin edges
    const uint32_t count = m > 0 ? static_cast<uint32_t>(2u + std::sqrt(static_cast<double>(
                                                                   static_cast<uint32_t>(m))))
                                 : 2u;
    size_t scan_bytes = 0;
    hipError_t e = rocprim::exclusive_scan(nullptr, scan_bytes, static_cast<int64_t*>(nullptr),
                                           static_cast<int64_t*>(nullptr), int64_t{0}, ns + 1,
                                           rocprim::plus<int64_t>(), stream);
    if (e != hipSuccess) return e;
    const size_t bytes = align_up(16) + align_up(static_cast<size_t>(count) * 4) + 2 * align_up((ns + 1) * 8) +
                         align_up(scan_bytes) + align_up(8) + align_up(4) + 256;
    PoolBlock block;
    e = block.get(bytes);
    if (e != hipSuccess) return e;
    char* mem = block.as();
    char* q = mem;
    auto take = [&](size_t b) { char* x = q; q += align_up(b); return x; };
    auto* range = reinterpret_cast<unsigned long long*>(take(16));
    auto* bins = reinterpret_cast<uint32_t*>(take(static_cast<size_t>(count) * 4));
    auto* ncount = reinterpret_cast<int64_t*>(take((ns + 1) * 8));
    auto* noff = reinterpret_cast<int64_t*>(take((ns + 1) * 8));
    void* temp = take(scan_bytes);
    auto* dstart = reinterpret_cast<double*>(take(8));
    auto* dbad = reinterpret_cast<int32_t*>(take(4));
    PoolBlock nsb_block;
    char* nsb_mem = nullptr;
    int64_t total = 0;
    int32_t host_bad = 0;
    const unsigned long long init[2] = {~0ull, 0ull};
    e = hipMemcpyAsync(range, init, 16, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemsetAsync(bins, 0, static_cast<size_t>(count) * 4, stream);
    if (e == hipSuccess) e = hipMemsetAsync(ncount, 0, (ns + 1) * 8, stream);
    if (e == hipSuccess) e = hipMemsetAsync(dbad, 0, 4, stream);
    if (e == hipSuccess && m > 0) {
        const uint32_t grid = static_cast<uint32_t>(std::min<int64_t>((m + 255) / 256, 2048));
        hipLaunchKernelGGL(time_range_kernel, dim3(grid), dim3(256), 0, stream, in.arrival_time, m, range);
        // LDS counts: a few photons per thread before each block's flush
        const uint32_t hgrid = static_cast<uint32_t>(std::min<int64_t>((m + 4095) / 4096, 1024));
        hipLaunchKernelGGL(mode_histogram_kernel, dim3(hgrid), dim3(256), 0, stream, in.arrival_time, m, range, count,
                           bins);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mode_start_kernel, dim3(1), dim3(1024), 0, stream, bins, range, count, m, exposure_time,
                           dstart);
        hipLaunchKernelGGL(nsb_count_kernel, dim3(static_cast<uint32_t>((ns + 255) / 256)), dim3(256), 0, stream,
                           num_sensors, efficiencies, rate, exposure_time, seed, ncount, dbad);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = rocprim::exclusive_scan(temp, scan_bytes, ncount, noff, int64_t{0}, ns + 1, rocprim::plus<int64_t>(),
                                    stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&total, noff + ns, 8, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(start, dstart, 8, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&host_bad, dbad, 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess) {
        *num_photons = m + total;
        if (host_bad) *bad = 1;
        if (m + total > out.capacity) *too_small = 1;
    }
    if (e == hipSuccess && !*bad && !*too_small) {
        const size_t nt = static_cast<size_t>(total > 0 ? total : 1);
        e = nsb_block.get(2 * align_up(nt * 8));
        nsb_mem = nsb_block.as();
        double* nsb_time = reinterpret_cast<double*>(nsb_mem);
        double* nsb_lambda = reinterpret_cast<double*>(nsb_mem + align_up(nt * 8));
        if (e == hipSuccess && total > 0) {
            hipLaunchKernelGGL(nsb_emit_kernel, dim3(static_cast<uint32_t>((ns + 255) / 256)), dim3(256), 0, stream,
                               num_sensors, efficiencies, rate, exposure_time, seed, dstart, cdf, cdf_n, cdf_total,
                               noff, nsb_time, nsb_lambda, dbad);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipMemcpyAsync(&host_bad, dbad, 4, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e == hipSuccess && host_bad) *bad = 1;
        }
        if (e == hipSuccess && !*bad) {
            hipLaunchKernelGGL(add_offsets_kernel, dim3(static_cast<uint32_t>((ns + 1 + 255) / 256)), dim3(256), 0,
                               stream, in.offsets, noff, num_sensors, out.offsets);
            if (m > 0)
                hipLaunchKernelGGL(merge_photons_kernel, dim3(static_cast<uint32_t>((m + 255) / 256)), dim3(256), 0,
                                   stream, in, m, num_sensors, noff, nsb_time, dstart, out);
            if (total > 0)
                hipLaunchKernelGGL(merge_background_kernel, dim3(static_cast<uint32_t>((total + 255) / 256)),
                                   dim3(256), 0, stream, in, num_sensors, total, noff, nsb_time, nsb_lambda, dstart,
                                   out);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
        }
    }
    (void)hipStreamSynchronize(stream);  // the borrowed blocks are idle when returned
    return e;
}

hipError_t photon_stream(const mlgpu_pulses& p, const int16_t* slice, int32_t num_sensors, int64_t count,
                         uint8_t* symbols, int32_t* truth, int64_t capacity, int64_t* num_symbols, int* too_small,
                         hipStream_t stream, bool trailing_marker) {
    *num_symbols = 0;
    *too_small = 0;
    const size_t m = static_cast<size_t>(count);
    size_t scan_bytes = 0;
    const auto kept = rocprim::make_transform_iterator(rocprim::counting_iterator<int64_t>(0), Kept{slice});
    hipError_t e = rocprim::exclusive_scan(nullptr, scan_bytes, kept, static_cast<int64_t*>(nullptr), int64_t{0},
                                           m + 1, rocprim::plus<int64_t>(), stream);
    if (e != hipSuccess) return e;
    PoolBlock block;
    e = block.get(align_up((m + 2) * 8) + align_up(scan_bytes));
    if (e != hipSuccess) return e;
    char* mem = block.as();
    auto* kscan = reinterpret_cast<int64_t*>(mem);
    void* temp = mem + align_up((m + 2) * 8);
    int64_t total = 0;
    // kscan[k]: kept pulses before pulse k; kscan[count] (the total) is
    // filled in from the last pulse below
    e = rocprim::exclusive_scan(temp, scan_bytes, kept, kscan, int64_t{0}, m, rocprim::plus<int64_t>(), stream);
    int64_t last = 0;
    int16_t last_slice = -1;
    if (e == hipSuccess && m > 0) e = hipMemcpyAsync(&last, kscan + m - 1, 8, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess && m > 0) e = hipMemcpyAsync(&last_slice, slice + m - 1, 2, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess) {
        total = m > 0 ? last + (last_slice >= 0 ? 1 : 0) : 0;  // kept pulses
        e = hipMemcpyAsync(kscan + m, &total, 8, hipMemcpyHostToDevice, stream);
    }
    if (e == hipSuccess) {
        // PhotonStream::num_symbols_to_represent (PhotonStream.cpp:181-193);
        // FACT's stream ends every pixel with a marker
        const int64_t sym = trailing_marker ? total + num_sensors
                                            : (total + num_sensors > 0 ? total + num_sensors - 1 : 0);
        *num_symbols = sym;
        if (sym > capacity) *too_small = 1;
    }
    if (e == hipSuccess && !*too_small) {
        if (m > 0)
            hipLaunchKernelGGL(stream_pulses_kernel, dim3(static_cast<uint32_t>((m + 255) / 256)), dim3(256), 0,
                               stream, p, slice, count, num_sensors, kscan, symbols, truth);
        const int32_t markers = trailing_marker ? num_sensors : num_sensors - 1;
        if (markers > 0)
            hipLaunchKernelGGL(stream_markers_kernel, dim3(static_cast<uint32_t>((markers + 255) / 256)),
                               dim3(256), 0, stream, p, markers, kscan, symbols);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
    }
    (void)hipStreamSynchronize(stream);  // the borrowed blocks are idle when returned
    return e;
}

hipError_t fact_slices(const mlgpu_pulses& p, int64_t count, double trigger_time, double slice_duration,
                       int16_t* slice, hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    const size_t m = static_cast<size_t>(count);
    PoolBlock block;
    hipError_t e = block.get(align_up(m * 8) + align_up(sizeof(SelectState)) + 256);
    if (e != hipSuccess) return e;
    char* q = block.as();
    auto* keys = reinterpret_cast<uint64_t*>(q);
    auto* st = reinterpret_cast<SelectState*>(q + align_up(m * 8));
    auto* median = reinterpret_cast<double*>(q + align_up(m * 8) + align_up(sizeof(SelectState)));
    // estimate_median_arrival_time (fact_propagation.cpp:34-60): the sorted
    // time at unsigned(count * 0.5), found by radix selection (no sort)
    const int64_t half = static_cast<int64_t>(static_cast<uint32_t>(static_cast<double>(count) * 0.5));
    const uint32_t grid = static_cast<uint32_t>(std::min<size_t>((m + 1023) / 1024, 2048));
    hipLaunchKernelGGL(time_keys_kernel, dim3(static_cast<uint32_t>((m + 255) / 256)), dim3(256), 0, stream,
                       p.arrival_time, count, keys);
    hipLaunchKernelGGL(select_init_kernel, dim3(1), dim3(256), 0, stream, st, static_cast<long long>(half));
    for (int shift = 56; shift >= 0; shift -= 8) {
        hipLaunchKernelGGL(select_hist_kernel, dim3(grid), dim3(256), 0, stream, keys, count, st, shift);
        hipLaunchKernelGGL(select_pick_kernel, dim3(1), dim3(256), 0, stream, st, shift);
    }
    hipLaunchKernelGGL(key_at_kernel, dim3(1), dim3(64), 0, stream, reinterpret_cast<const uint64_t*>(&st->prefix),
                       int64_t{0}, median);
    hipLaunchKernelGGL(fact_slices_kernel, dim3(static_cast<uint32_t>((m + 255) / 256)), dim3(256), 0, stream,
                       p.arrival_time, count, median, trigger_time, slice_duration, slice);
    e = hipGetLastError();
    (void)hipStreamSynchronize(stream);  // the borrowed block is idle when returned
    return e;
}

}  // namespace mlgpu
