/*
 * oracle.h -- FP64 CPU restatement of merlict's photon propagation.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU propagator is
 * compared against; only tests, the smoke check and the CPU-baseline leg of
 * the benchmark may load it.  The product never links or calls it.
 *
 * It consumes the same scene description as the GPU library (include/mlgpu.h)
 * and restates, in plain C and double precision, the reference algorithm:
 *   tree build       merlict/Frame.cpp:125-147,190-255, merlict/Frames.cpp
 *   DFS first hit    merlict/RayAndFrame.cpp:8-173
 *   primitives       merlict/scenery/primitive/*.cpp, scenery/geometry/*.cpp
 *   propagator       merlict/PhotonAndFrame.cpp:6-124
 *   Fresnel          merlict/Fresnel.cpp:7-38
 *   Func1            merlict/function/Func1.cpp:22-130, Limits.cpp:37-46
 *   time of flight   merlict/Photon.cpp:44-56
 *   sensor record    merlict/sensor/Sensor.cpp:14-30, RayForPropagation.cpp:114-136
 * Random draws come from the Philox4x32-10 stream (Salmon et al., SC'11)
 * keyed by (seed, photon id), the same stream the GPU kernel consumes.
 */
#ifndef MLGPU_ORACLE_H_
#define MLGPU_ORACLE_H_

#include <stdint.h>
#include "../include/mlgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene oracle_scene;

oracle_scene* oracle_scene_create(const mlgpu_scene_desc* desc);
void oracle_scene_destroy(oracle_scene* s);
const char* oracle_last_error(void);

/* Number of frames after clustering (description frames + octant frames). */
int oracle_num_frames(const oracle_scene* s);

/* Pre-order walk of the initialised tree.  For position k writes the frame
 * index (octant frames have index >= n_frames of the description), the
 * number of children, frame2world (12 doubles, row major 3x4) and the
 * bounding sphere radius.  Returns the number of frames written. */
int oracle_tree_dump(const oracle_scene* s, int32_t* frame, int32_t* n_children,
                     double* frame2world, double* radius, int capacity);

/* First intersection of one ray with the frame tree below root_frame.
 * frame = -1 for no hit.  position in the root frame.  Returns the number
 * of bounding-sphere tests and leaf tests through the counters if given. */
int oracle_first_hit(const oracle_scene* s, int32_t root_frame,
                     const double support[3], const double direction[3],
                     int32_t* frame, double* distance, double position[3],
                     int64_t* sphere_tests, int64_t* leaf_tests);

typedef struct oracle_results {
    int32_t* final_frame;
    int32_t* sensor_id;
    uint8_t* final_interaction;
    uint8_t* num_interactions;
    uint8_t* error;
    double* time_of_flight;
    double* x;
    double* y;
    double* cx;
    double* cy;
    double* accumulated_distance;
} oracle_results;

/* Propagate photons (host arrays, double SoA).  One Philox generator per
 * photon, as merlict_multi_thread.cpp:12-26 uses one Mt19937 per photon.
 * num_threads <= 1 runs single threaded; otherwise contiguous chunks. */
/* Work counts of the reference's algorithm over one oracle_propagate call,
 * per event (the instrumented restatement behind the VALU roofline):
 * first-intersection queries, bounding-sphere tests, leaf intersections by
 * kind, Func1 lookups (and the sum over them of ceil(log2(points)), the
 * compares of their binary search), Fresnel evaluations. */
typedef struct oracle_work {
    int64_t queries;
    int64_t sphere_tests;
    int64_t leaf_cap, leaf_planar, leaf_cylinder, leaf_sphere;
    int64_t func_lookups, func_compares;
    int64_t fresnel;
    int64_t bounces;
} oracle_work;
/* the counts of the last oracle_propagate call on the calling thread */
void oracle_last_work(oracle_work* out);

int oracle_propagate(const oracle_scene* s, int32_t world_frame,
                     const mlgpu_photons* photons, int64_t n,
                     uint32_t max_num_interactions, uint64_t seed,
                     uint64_t photon_id_offset, oracle_results* out,
                     int num_threads, int64_t* sphere_tests,
                     int64_t* leaf_tests);

/* Philox4x32-10 block and the derived 24-bit uniform draw. */
void oracle_philox4x32_10(const uint32_t counter[4], const uint32_t key[2],
                          uint32_t out[4]);
double oracle_uniform(uint64_t seed, uint32_t stream, uint64_t photon_id,
                      uint32_t draw_index);

/* Func1 evaluation of description function f (-1..-3 = defaults
 * reflection, refraction, absorption).  Returns 0, or 1 when x is outside
 * the table. */
int oracle_func_evaluate(const oracle_scene* s, int32_t f, double x,
                         double* y);

/* Photon sources (mlgpu_source of include/mlgpu.h): photons
 * photon_id_offset .. + n - 1 into host arrays out[0..6] (support xyz,
 * direction xyz, wavelength).  The helpers are exposed for their tests. */
void oracle_generate_photons(const mlgpu_source* src, int64_t n, uint64_t seed,
                             uint64_t photon_id_offset, double* const out[7]);
double oracle_log_series(double x);
void oracle_sincos_turn(double u, double* s, double* c);

/* Arrival-time histograms (mlgpu_histogram of include/mlgpu.h) of host
 * result arrays: counts (host, num_sensors * num_bins) accumulate. */
void oracle_histogram_arrivals(const int32_t* sensor_id, const uint8_t* error,
                               const double* time_of_flight, int64_t n,
                               const mlgpu_histogram* hist, uint32_t* counts);
double oracle_arrival_jitter(uint64_t seed, uint64_t photon_id);

/* CORSIKA photon bunches to photons (merlict::EventIoPhotonFactory,
 * merlict_corsika/PhotonFactory.cpp:8-148; the weight draw is Philox stream
 * 5 of (seed, id_offset + bunch), uniform (w0 >> 8) 2^-24).  bunches: n x 8
 * float (x, y [cm], cx, cy, time [ns], emission height [cm], weight,
 * wavelength [nm]).  The photons of every bunch, in bunch order, go to
 * out[0..6] (support xyz, direction xyz, wavelength) and their bunch index
 * to truth (both up to capacity entries).  Returns the number of photons,
 * -1 for a weight outside [0, 16], -2 when capacity is too small. */
int64_t oracle_corsika_photons(const float* bunches, int64_t n, uint64_t seed, uint64_t id_offset,
                               double* const out[7], int32_t* truth, int64_t capacity);

/* Sensor signal chain (merlict_signal_processing): per-sensor pipelines of
 * the detected photons sorted by (arrival time, photon index)
 * (PipelinePhoton.cpp:25-66), returning their number; the photo-electric
 * converter (PhotoElectricConverter.cpp:10-95) with Philox stream 6 of
 * (seed, sensor), returning the number of pulses (-1: a wavelength outside
 * the table, -2: capacity); the pulse extraction (pulse_extraction.cpp:
 * 9-30) with Philox stream 7 of (seed, sensor), slice -1 when dropped. */
int64_t oracle_sensor_pipelines(const int32_t* sensor_id, const uint8_t* error, const double* time_of_flight,
                                const double* wavelength, const int32_t* truth_id, int64_t n, int32_t num_sensors,
                                int64_t* offsets, double* arrival_time, double* out_wavelength,
                                int32_t* out_truth);
int64_t oracle_photoelectric(const int64_t* offsets, const double* arrival_time, const double* wavelength,
                             const int32_t* truth_id, int32_t num_sensors, const double* qe_x, const double* qe_y,
                             int32_t qe_n, double dark_rate, double crosstalk, double exposure_time, uint64_t seed,
                             int64_t* out_offsets, double* out_time, int32_t* out_truth, int64_t capacity);
void oracle_extract_pulses(const int64_t* offsets, const double* arrival_time, int32_t num_sensors,
                           double duration, double std_dev, uint64_t seed, int16_t* slice);

/* Night-sky background (Injector.cpp:11-107): sensor s draws from Philox
   stream 9 of (seed, s); the merged pipelines (arrival times minus the
   exposure start, stored in *start); -1 bad table/rate/draw, -2 capacity. */
int64_t oracle_inject_nsb(const int64_t* offsets, const double* arrival_time, const double* wavelength,
                          const int32_t* truth_id, int32_t num_sensors, const float* eff, const double* flux_x,
                          const double* flux_y, int32_t flux_n, double rate, double exposure_time, uint64_t seed,
                          int64_t* out_offsets, double* out_time, double* out_wavelength, int32_t* out_truth,
                          int64_t capacity, double* start);
/* one_photon's calibration result (Calibrator.cpp:48-102) for n photons of
   `src` and their propagation results. */
void oracle_calibration_photons(const mlgpu_source* src, uint64_t seed, uint64_t photon_id_offset,
                                const int32_t* sensor_id, const uint8_t* error, const double* time_of_flight,
                                int64_t n, int32_t num_lixel, mlgpu_calibration_photon* out);
/* LixelStatisticsFiller over n records in order; -1 for a lixel id out of
   range. */
int oracle_lixel_statistics(const mlgpu_calibration_photon* rec, int64_t n, int64_t num_lixel, int64_t num_blocks,
                            int64_t photons_per_block, mlgpu_lixel_statistic* out);
/* merlict_visual's PinHoleCamera + Tracer (Tracer.cpp:21-145) per pixel on
   Philox stream 10 of (seed, pixel); image rows x cols x 3 floats; returns
   the pixels with a 533 nm lookup outside a table, -1 on bad arguments. */
int64_t oracle_render_pinhole(const oracle_scene* s, int32_t world, const mlgpu_pinhole_camera* cam,
                              const mlgpu_render_config* cfg, const float* colors, uint64_t seed, float* image);
/* FACT's photon stream of an event (fact_propagation.cpp:34-94); the length
   or -1 when capacity is too small. */
int64_t oracle_fact_photon_stream(const int64_t* offsets, const double* arrival_time, int32_t num_pixels,
                                  double trigger_time, uint8_t* raw, int64_t capacity);
/* night_sky_background::Light::rate (Light.cpp:9-28); -1 for a bad table */
double oracle_nsb_light_rate(const double* flux_x, const double* flux_y, int32_t flux_n, double max_fov_radius,
                             double max_aperture_radius);

#ifdef __cplusplus
}
#endif

#endif  /* MLGPU_ORACLE_H_ */
