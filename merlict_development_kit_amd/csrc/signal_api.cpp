// C ABI of the sensor signal chain (include/mlgpu.h): pipelines, photo-
// electric conversion, pulse extraction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/mlgpu.h"
#include "device_scene.hpp"
#include "signal.hpp"

namespace mlgpu {
void set_last_error(const std::string& msg);
}

namespace {

using mlgpu::set_last_error;

int hip_fail(hipError_t e, const char* what) {
    set_last_error(std::string(what) + ": " + hipGetErrorString(e));
    return MLGPU_ERR_HIP;
}

}  // namespace

extern "C" {

int mlgpu_sensor_pipelines(const mlgpu_results* r, const double* wavelength, const int32_t* truth_id, int64_t n,
                           int32_t num_sensors, const mlgpu_pipelines* out, int64_t* num_photons, void* stream) {
    if (!r || !out || !num_photons || n < 0 || num_sensors < 0 || !out->offsets) {
        set_last_error("null argument or negative size");
        return MLGPU_ERR_ARGUMENT;
    }
    *num_photons = 0;
    if (n > INT32_MAX) { set_last_error("at most 2^31 - 1 photons"); return MLGPU_ERR_ARGUMENT; }
    if (n > 0 && (!r->sensor_id || !r->error || !r->time_of_flight || !wavelength)) {
        set_last_error("null result or wavelength array");
        return MLGPU_ERR_ARGUMENT;
    }
    if (n == 0) {
        const hipError_t e = hipMemsetAsync(out->offsets, 0, sizeof(int64_t) * (static_cast<size_t>(num_sensors) + 1),
                                            static_cast<hipStream_t>(stream));
        if (e == hipSuccess) (void)hipStreamSynchronize(static_cast<hipStream_t>(stream));
        return e == hipSuccess ? MLGPU_OK : hip_fail(e, "sensor pipelines");
    }
    int too_small = 0;
    const hipError_t e = mlgpu::build_pipelines(*r, wavelength, truth_id, n, num_sensors, *out, num_photons,
                                                static_cast<hipStream_t>(stream), &too_small);
    if (e != hipSuccess) return hip_fail(e, "sensor pipelines");
    if (too_small) { set_last_error("capacity below the number of detected photons"); return MLGPU_ERR_ARGUMENT; }
    return MLGPU_OK;
}

int mlgpu_photoelectric_convert(const mlgpu_pipelines* p, int32_t num_sensors, const mlgpu_converter* c,
                                const mlgpu_pulses* out, int64_t* num_pulses, void* stream) {
    if (!p || !c || !out || !num_pulses || num_sensors < 0 || !out->offsets || !c->qe_x || !c->qe_y || c->qe_n < 1) {
        set_last_error("null argument, negative size or empty quantum efficiency table");
        return MLGPU_ERR_ARGUMENT;
    }
    *num_pulses = 0;
    // Converter::Converter (PhotoElectricConverter.cpp:18-29) rejects p >= 1
    if (!(c->crosstalk >= 0.0) || !(c->crosstalk < 1.0 - 1e-6)) {
        set_last_error("need 0 <= crosstalk < 1 - 1e-6");
        return MLGPU_ERR_ARGUMENT;
    }
    if (!(c->dark_rate >= 0.0) || !(c->exposure_time >= 0.0) || !(c->dark_rate * c->exposure_time <= 1e7)) {
        set_last_error("need dark_rate >= 0, exposure_time >= 0 and dark_rate * exposure_time <= 1e7");
        return MLGPU_ERR_ARGUMENT;
    }
    if (num_sensors == 0) return MLGPU_OK;
    // the Func1 table (merlict/function/Func1.cpp:22-50): slopes before sorting
    const int n = c->qe_n;
    std::vector<mlgpu::FuncPoint> pts(static_cast<size_t>(n));
    for (int i = 0; i < n; i++) {
        pts[i].x = c->qe_x[i];
        pts[i].y = c->qe_y[i];
        pts[i].slope = (i != n - 1) ? (c->qe_y[i] - c->qe_y[i + 1]) / (c->qe_x[i] - c->qe_x[i + 1]) : 0.0;
        pts[i].pad = 0.0;
    }
    std::sort(pts.begin(), pts.end(), [](const mlgpu::FuncPoint& a, const mlgpu::FuncPoint& b) { return a.x < b.x; });
    for (int i = 0; i + 1 < n; i++)
        if (pts[i].x == pts[i + 1].x) { set_last_error("quantum efficiency table has a duplicate wavelength"); return MLGPU_ERR_ARGUMENT; }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    mlgpu::FuncPoint* qe = nullptr;
    hipError_t e = hipMalloc(&qe, sizeof(mlgpu::FuncPoint) * pts.size());
    if (e != hipSuccess) return hip_fail(e, "hipMalloc quantum efficiency");
    e = hipMemcpyAsync(qe, pts.data(), sizeof(mlgpu::FuncPoint) * pts.size(), hipMemcpyHostToDevice, st);
    int bad = 0, too_small = 0;
    int64_t m = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&m, p->offsets + num_sensors, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && m > 0 && (!p->arrival_time || !p->wavelength || !p->truth_id)) {
        (void)hipFree(qe);
        set_last_error("null pipeline array");
        return MLGPU_ERR_ARGUMENT;
    }
    if (e == hipSuccess)
        e = mlgpu::convert_pulses(*p, num_sensors, m, qe, n, c->dark_rate, c->crosstalk, c->exposure_time, c->seed,
                                  *out, num_pulses, &bad, &too_small, st);
    (void)hipStreamSynchronize(st);
    (void)hipFree(qe);
    if (e != hipSuccess) return hip_fail(e, "photo-electric conversion");
    if (bad) { set_last_error("a wavelength outside the quantum efficiency table"); return MLGPU_ERR_ARGUMENT; }
    if (too_small) { set_last_error("capacity below the number of pulses"); return MLGPU_ERR_ARGUMENT; }
    return MLGPU_OK;
}

int mlgpu_extract_pulses(const mlgpu_pulses* p, int32_t num_sensors, double duration, double std_dev, uint64_t seed,
                         int16_t* slice, void* stream) {
    if (!p || !slice || num_sensors < 0 || !p->offsets || !(duration > 0.0) || !(std_dev >= 0.0)) {
        set_last_error("null argument or bad slice duration / spread");
        return MLGPU_ERR_ARGUMENT;
    }
    if (num_sensors == 0) return MLGPU_OK;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    int64_t count = 0;
    hipError_t e = hipMemcpyAsync(&count, p->offsets + num_sensors, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && count > 0 && !p->arrival_time) {
        set_last_error("null pulse array");
        return MLGPU_ERR_ARGUMENT;
    }
    if (e == hipSuccess) e = mlgpu::extract_pulses(*p, count, duration, std_dev, seed, slice, st);
    if (e != hipSuccess) return hip_fail(e, "pulse extraction");
    return MLGPU_OK;
}

}  // extern "C"
