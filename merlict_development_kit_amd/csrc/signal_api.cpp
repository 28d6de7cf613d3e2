// C ABI of the sensor signal chain (include/mlgpu.h): pipelines, photo-
// electric conversion, pulse extraction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/mlgpu.h"
#include "device_scene.hpp"
#include "pool.hpp"
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

// A merlict Func1 on the host (merlict/function/Func1.cpp:22-71): slopes
// from the table as given, then sorted by x; no duplicate x.
struct HostFunc {
    std::vector<mlgpu::FuncPoint> p;
    double lower = 0.0, upper = 0.0;
    bool make(const std::vector<double>& x, const std::vector<double>& y) {
        const size_t n = x.size();
        if (n == 0) return false;
        p.resize(n);
        for (size_t i = 0; i < n; i++) {
            p[i].x = x[i];
            p[i].y = y[i];
            p[i].slope = (i != n - 1) ? (y[i] - y[i + 1]) / (x[i] - x[i + 1]) : 0.0;
            p[i].pad = 0.0;
        }
        std::sort(p.begin(), p.end(), [](const mlgpu::FuncPoint& a, const mlgpu::FuncPoint& b) { return a.x < b.x; });
        for (size_t i = 0; i + 1 < n; i++)
            if (p[i].x == p[i + 1].x) return false;
        lower = p.front().x;
        upper = p.back().x;
        return true;
    }
    // Func1::evaluate: false outside [lower, upper)
    bool eval(double x, double& y) const {
        if (x < lower || x >= upper) return false;
        const auto it = std::upper_bound(p.begin(), p.end(), x,
                                         [](double v, const mlgpu::FuncPoint& q) { return q.x > v; });
        if (it == p.begin()) return false;
        const mlgpu::FuncPoint& q = *(it - 1);
        y = q.slope * x + q.y - q.slope * q.x;
        return true;
    }
    double max() const {
        double m = p.front().y;
        for (const auto& q : p) if (q.y > m) m = q.y;
        return m;
    }
};

constexpr uint64_t kSteps = 4096;  // function::DEFAULT_STEPS (function/tools.h:12)

// function::integral (function/tools.cpp:37-56): the running sum of
// f(x) * step on the grid lower + i * step, the last point at the upper
// limit with f just below it
bool integral(const HostFunc& f, HostFunc& out) {
    const double step = (f.upper - f.lower) / static_cast<double>(kSteps);
    std::vector<double> x{f.lower}, y{0.0};
    for (uint64_t i = 1; i < kSteps; i++) {
        const double xi = f.lower + static_cast<double>(i) * step;
        double v;
        if (!f.eval(xi, v)) return false;
        x.push_back(xi);
        y.push_back(v * step + y.back());
    }
    double v;
    if (!f.eval(std::nextafter(f.upper, f.lower), v)) return false;
    x.push_back(f.upper);
    y.push_back(v * step + y.back());
    return out.make(x, y);
}

// function::inverse (function/tools.cpp:58-76): (f(x), x) on the same grid
bool inverse(const HostFunc& f, HostFunc& out) {
    const double step = (f.upper - f.lower) / static_cast<double>(kSteps);
    std::vector<double> x, y;
    double v;
    if (!f.eval(f.lower, v)) return false;
    x.push_back(v);
    y.push_back(f.lower);
    for (uint64_t i = 1; i < kSteps; i++) {
        const double xi = f.lower + static_cast<double>(i) * step;
        if (!f.eval(xi, v)) return false;
        x.push_back(v);
        y.push_back(xi);
    }
    if (!f.eval(std::nextafter(f.upper, f.lower), v)) return false;
    x.push_back(v);
    y.push_back(f.upper);
    return out.make(x, y);
}

// SamplesFromDistribution (random/SamplesFromDistribution.cpp:8-17): the
// inverse cumulative table and the total integral of the flux
int flux_tables(const double* fx, const double* fy, int32_t n, HostFunc& cdf, double& total) {
    if (!fx || !fy || n < 2) {
        set_last_error("flux table needs at least two points");
        return MLGPU_ERR_ARGUMENT;
    }
    HostFunc f, F;
    if (!f.make(std::vector<double>(fx, fx + n), std::vector<double>(fy, fy + n))) {
        set_last_error("flux table has a duplicate wavelength");
        return MLGPU_ERR_ARGUMENT;
    }
    if (!integral(f, F) || !inverse(F, cdf)) {
        set_last_error("flux table: its cumulative distribution is not invertible (a flux <= 0 somewhere)");
        return MLGPU_ERR_ARGUMENT;
    }
    total = F.max();
    return MLGPU_OK;
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
    mlgpu::PoolBlock qe_block;
    hipError_t e = qe_block.get(sizeof(mlgpu::FuncPoint) * pts.size());
    if (e != hipSuccess) return hip_fail(e, "quantum efficiency table");
    auto* qe = qe_block.as<mlgpu::FuncPoint>();
    e = hipMemcpyAsync(qe, pts.data(), sizeof(mlgpu::FuncPoint) * pts.size(), hipMemcpyHostToDevice, st);
    int bad = 0, too_small = 0;
    int64_t m = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&m, p->offsets + num_sensors, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && m > 0 && (!p->arrival_time || !p->wavelength || !p->truth_id)) {
        set_last_error("null pipeline array");
        return MLGPU_ERR_ARGUMENT;
    }
    if (e == hipSuccess)
        e = mlgpu::convert_pulses(*p, num_sensors, m, qe, n, c->dark_rate, c->crosstalk, c->exposure_time, c->seed,
                                  *out, num_pulses, &bad, &too_small, st);
    (void)hipStreamSynchronize(st);
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

int mlgpu_photon_stream(const mlgpu_pulses* p, const int16_t* slice, int32_t num_sensors, uint8_t* symbols,
                        int32_t* truth, int64_t capacity, int64_t* num_symbols, void* stream) {
    if (!p || !num_symbols || num_sensors < 0 || capacity < 0 || !p->offsets) {
        set_last_error("null argument or negative size");
        return MLGPU_ERR_ARGUMENT;
    }
    *num_symbols = 0;
    if (num_sensors == 0) return MLGPU_OK;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    int64_t count = 0;
    hipError_t e = hipMemcpyAsync(&count, p->offsets + num_sensors, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "photon stream");
    if (count > 0 && (!slice || (truth && !p->truth_id))) {
        set_last_error("null slice or truth array");
        return MLGPU_ERR_ARGUMENT;
    }
    if (!symbols && count + num_sensors > 1) {
        set_last_error("null symbols");
        return MLGPU_ERR_ARGUMENT;
    }
    int too_small = 0;
    e = mlgpu::photon_stream(*p, slice, num_sensors, count, symbols, truth, capacity, num_symbols, &too_small, st);
    if (e != hipSuccess) return hip_fail(e, "photon stream");
    if (too_small) { set_last_error("capacity below the number of symbols"); return MLGPU_ERR_ARGUMENT; }
    return MLGPU_OK;
}

int mlgpu_fact_photon_stream(const mlgpu_pulses* p, int32_t num_pixels, double trigger_time, uint8_t* raw,
                             int64_t capacity, int64_t* num_raw, void* stream) {
    if (!p || !num_raw || num_pixels < 0 || capacity < 0 || !p->offsets) {
        set_last_error("null argument or negative size");
        return MLGPU_ERR_ARGUMENT;
    }
    *num_raw = 0;
    if (num_pixels == 0) return MLGPU_OK;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    int64_t count = 0;
    hipError_t e = hipMemcpyAsync(&count, p->offsets + num_pixels, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "FACT photon stream");
    if (count > 0 && !p->arrival_time) { set_last_error("null pulse array"); return MLGPU_ERR_ARGUMENT; }
    if (!raw) { set_last_error("null raw"); return MLGPU_ERR_ARGUMENT; }
    mlgpu::PoolBlock slices;
    e = slices.get(static_cast<size_t>(count > 0 ? count : 1) * sizeof(int16_t));
    if (e != hipSuccess) return hip_fail(e, "FACT photon stream");
    // photon_stream.h:16: the slice duration is the float 0.5e-9
    const double slice_duration = static_cast<double>(0.5e-9f);
    e = mlgpu::fact_slices(*p, count, trigger_time, slice_duration, slices.as<int16_t>(), st);
    int too_small = 0;
    if (e == hipSuccess)
        e = mlgpu::photon_stream(*p, slices.as<int16_t>(), num_pixels, count, raw, nullptr, capacity, num_raw,
                                 &too_small, st, true);
    (void)hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "FACT photon stream");
    if (too_small) { set_last_error("capacity below the stream's length"); return MLGPU_ERR_ARGUMENT; }
    return MLGPU_OK;
}

int mlgpu_nsb_light_rate(const double* flux_x, const double* flux_y, int32_t flux_n, double max_fov_radius,
                         double max_aperture_radius, double* rate) {
    if (!rate) { set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    HostFunc cdf;
    double total = 0.0;
    const int rc = flux_tables(flux_x, flux_y, flux_n, cdf, total);
    if (rc != MLGPU_OK) return rc;
    // Light.cpp:15-27 with NightSkyBackground.h:11-12 and
    // get_solid_angle_for_opening_angle (merlict/tools.cpp:25-48)
    const double fov_radius = 1.20 * max_fov_radius;
    const double solid_angle = 2.0 * M_PI * (1.0 - std::cos(fov_radius));
    const double aperture_radius = 1.05 * max_aperture_radius;
    const double area = M_PI * aperture_radius * aperture_radius;
    *rate = total * area * solid_angle;
    return MLGPU_OK;
}

int mlgpu_inject_nsb(const mlgpu_pipelines* in, int32_t num_sensors, const float* efficiencies, const mlgpu_nsb* nsb,
                     const mlgpu_pipelines* out, int64_t* num_photons, double* exposure_start_time, void* stream) {
    if (!in || !nsb || !out || !num_photons || !exposure_start_time || num_sensors < 0 || !in->offsets ||
        !out->offsets) {
        set_last_error("null argument or negative size");
        return MLGPU_ERR_ARGUMENT;
    }
    *num_photons = 0;
    *exposure_start_time = 0.0;
    if (!(nsb->rate >= 0.0) || !(nsb->exposure_time >= 0.0)) {
        set_last_error("need rate >= 0 and exposure_time >= 0");
        return MLGPU_ERR_ARGUMENT;
    }
    if (num_sensors == 0) return MLGPU_OK;  // Injector.cpp:20-21
    if (!efficiencies) { set_last_error("null efficiencies"); return MLGPU_ERR_ARGUMENT; }
    if (out->offsets == in->offsets || (out->arrival_time && out->arrival_time == in->arrival_time)) {
        set_last_error("out must not share arrays with in");
        return MLGPU_ERR_ARGUMENT;
    }
    HostFunc cdf;
    double total = 0.0;
    const int rc = flux_tables(nsb->flux_x, nsb->flux_y, nsb->flux_n, cdf, total);
    if (rc != MLGPU_OK) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    int64_t m = 0;
    hipError_t e = hipMemcpyAsync(&m, in->offsets + num_sensors, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "night-sky background");
    if (m > INT32_MAX) { set_last_error("at most 2^31 - 1 photons"); return MLGPU_ERR_ARGUMENT; }
    if (m > 0 && (!in->arrival_time || !in->wavelength || !in->truth_id)) {
        set_last_error("null pipeline array");
        return MLGPU_ERR_ARGUMENT;
    }
    mlgpu::PoolBlock cdf_block;
    e = cdf_block.get(sizeof(mlgpu::FuncPoint) * cdf.p.size());
    if (e != hipSuccess) return hip_fail(e, "night-sky table");
    auto* dcdf = cdf_block.as<mlgpu::FuncPoint>();
    e = hipMemcpyAsync(dcdf, cdf.p.data(), sizeof(mlgpu::FuncPoint) * cdf.p.size(), hipMemcpyHostToDevice, st);
    int bad = 0, too_small = 0;
    if (e == hipSuccess)
        e = mlgpu::inject_nsb(*in, num_sensors, m, efficiencies, dcdf, static_cast<int32_t>(cdf.p.size()), total,
                              nsb->rate, nsb->exposure_time, nsb->seed, *out, num_photons, exposure_start_time, &bad,
                              &too_small, st);
    (void)hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "night-sky background");
    if (bad) {
        set_last_error("night-sky background: a negative sensor rate, more than 1e7 expected photons in one sensor, "
                       "or a wavelength draw outside the table");
        return MLGPU_ERR_ARGUMENT;
    }
    if (too_small) { set_last_error("capacity below the number of photons with the background"); return MLGPU_ERR_ARGUMENT; }
    return MLGPU_OK;
}

}  // extern "C"
