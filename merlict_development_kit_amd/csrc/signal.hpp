// Launchers of the sensor signal chain (signal.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mlgpu.h"
#include "device_scene.hpp"

namespace mlgpu {

// Per-sensor pipelines of the detected photons of n results, sorted by
// (sensor, arrival time, photon index).  Synchronises.  *too_small when
// out.capacity is below the number of detected photons (*num_photons).
hipError_t build_pipelines(const mlgpu_results& r, const double* wavelength, const int32_t* truth, int64_t n,
                           int32_t num_sensors, const mlgpu_pipelines& out, int64_t* num_photons,
                           hipStream_t stream, int* too_small);

// Photo-electric conversion of every pipeline; qe (device) is the quantum
// efficiency table sorted by wavelength with merlict's pre-sort slopes.
// Synchronises; *bad_wavelength when a photon's wavelength is outside the
// table (nothing is written then), *too_small when out.capacity is below
// the number of pulses (*num_pulses).
hipError_t convert_pulses(const mlgpu_pipelines& p, int32_t num_sensors, int64_t num_photons, const FuncPoint* qe,
                          int32_t qe_n, double dark_rate, double crosstalk, double exposure_time, uint64_t seed,
                          const mlgpu_pulses& out, int64_t* num_pulses, int* bad_wavelength, int* too_small,
                          hipStream_t stream);

// Time slice of each of the count pulses (-1: outside the 100 slices).
hipError_t extract_pulses(const mlgpu_pulses& p, int64_t count, double duration, double std_dev, uint64_t seed,
                          int16_t* slice, hipStream_t stream);

// Night-sky-background injection into the m photons of `in` (ns sensors):
// cdf (device) is the inverse cumulative table of the flux, cdf_total the
// flux integral.  Synchronises; *num_photons receives m + the background
// photons, *start the exposure start; *too_small when out.capacity is below
// it, *bad when a wavelength draw falls outside the table or a sensor's
// rate is negative or above 1e7 / exposure_time (nothing is written then).
hipError_t inject_nsb(const mlgpu_pipelines& in, int32_t num_sensors, int64_t m, const float* efficiencies,
                      const FuncPoint* cdf, int32_t cdf_n, double cdf_total, double rate, double exposure_time,
                      uint64_t seed, const mlgpu_pipelines& out, int64_t* num_photons, double* start, int* bad,
                      int* too_small, hipStream_t stream);

// The photon-stream symbols of count pulses of num_sensors channels: kept
// pulses' slices with a marker between channels.  Synchronises;
// *num_symbols receives their number, *too_small when it exceeds capacity.
hipError_t photon_stream(const mlgpu_pulses& p, const int16_t* slice, int32_t num_sensors, int64_t count,
                         uint8_t* symbols, int32_t* truth, int64_t capacity, int64_t* num_symbols, int* too_small,
                         hipStream_t stream, bool trailing_marker = false);

// FACT's slices of count pulses: round((t - median + trigger_time) /
// slice_duration) within [30, 130), else -1.  Synchronises.
hipError_t fact_slices(const mlgpu_pulses& p, int64_t count, double trigger_time, double slice_duration,
                       int16_t* slice, hipStream_t stream);

}  // namespace mlgpu
