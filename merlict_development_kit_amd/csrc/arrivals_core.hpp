// Arrival-time binning shared by the histogram kernel and host builds of it
// (include/mlgpu.h, mlgpu_histogram).  The jitter is a Box-Muller normal from
// Philox4x32-10 block (0, 2, id_lo, id_hi): radius from words 0-1, angle
// from words 2-3, built from the series of sources_core.hpp so that host and
// device agree bit for bit.
#pragma once

#include "sources_core.hpp"

namespace mlgpu {

constexpr uint32_t kJitterStream = 2u;

MLGPU_HD inline double arrival_jitter(uint64_t seed, uint64_t id) {
    uint32_t w0, w1, w2, w3;
    philox4x32_10(0u, kJitterStream, static_cast<uint32_t>(id), static_cast<uint32_t>(id >> 32),
                  static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32), w0, w1, w2, w3);
    const double r = sqrt(-2.0 * log_series(1.0 - u53(w0, w1)));
    double s, c;
    sincos_turn(u53(w2, w3), s, c);
    return r * c;
}

// Histogram cell of photon i, or -1 when it is not counted.
MLGPU_HD inline int64_t arrival_cell(const mlgpu_histogram& h, const int32_t sensor, const uint8_t err,
                                     const double tof, const uint64_t id) {
    if (err != 0 || sensor < 0 || sensor >= h.num_sensors) return -1;
    const double t = h.jitter_sigma > 0.0 ? tof + h.jitter_sigma * arrival_jitter(h.seed, id) : tof;
    const double x = (t - h.t_start) / h.bin_width;
    if (!(x >= 0.0) || !(x < static_cast<double>(h.num_bins))) return -1;
    return static_cast<int64_t>(sensor) * h.num_bins + static_cast<int64_t>(x);
}

}  // namespace mlgpu
