/*
 * mlgpu.h -- C ABI of the MI355X photon propagator for merlict sceneries.
 *
 * Every entry point is plain C: pointers, sizes and status codes, no C++ or
 * torch types.  Device-side entry points take DEVICE pointers and a
 * hipStream_t passed as void* (stream-ordered, asynchronous).  Host-side
 * entry points (scene description, scene builders) work without a GPU.
 *
 * Units follow merlict: metres, seconds, wavelength in metres.
 */
#ifndef MLGPU_H_
#define MLGPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MLGPU_ABI_VERSION 2  /* 2: mlgpu_tuning.shade_lanes, the persistent schedule */

/* ---------------------------------------------------------------- status */
enum mlgpu_status {
    MLGPU_OK = 0,
    MLGPU_ERR_ARGUMENT = 1,    /* bad pointer / size / index              */
    MLGPU_ERR_SCENE = 2,       /* inconsistent scene description           */
    MLGPU_ERR_HIP = 3,         /* a HIP runtime call failed                */
    MLGPU_ERR_NO_DEVICE = 4,   /* no usable GPU                            */
    MLGPU_ERR_PHOTON = 5,      /* at least one photon carries an error flag */
    MLGPU_ERR_UNSUPPORTED = 6,
    MLGPU_ERR_FORMAT = 7       /* malformed input file                     */
};

/* Text of the last error raised on the calling thread. */
const char* mlgpu_last_error(void);
int mlgpu_abi_version(void);

/* ------------------------------------------------------ scene description
 * A flat list of frames in creation order, mother before child, frame 0 is
 * the root ("world").  It mirrors merlict's Frame tree before
 * init_tree_based_on_mother_child_relations(): every frame carries its pose
 * in its mother (position + Rot3), a surface kind with parameters, the eight
 * surface functions (merlict/SurfaceEntity.h:21-28; colours are not needed)
 * and an optional restriction frame (SurfaceEntity.h:13).
 * A frame with children is never intersected itself (RayAndFrame.cpp:141-147).
 */
enum mlgpu_kind {
    MLGPU_KIND_FRAME = 0,               /* no surface                       */
    MLGPU_KIND_SPHERE_CAP_HEX = 1,      /* p0 curvature R, p1 outer hex r   */
    MLGPU_KIND_SPHERE_CAP_CYL = 2,      /* p0 curvature R, p1 cap radius    */
    MLGPU_KIND_SPHERE_CAP_RECT = 3,     /* p0 R, p1 x width, p2 y width     */
    MLGPU_KIND_PLANE = 4,               /* p0 x width, p1 y width           */
    MLGPU_KIND_DISC = 5,                /* p0 radius                        */
    MLGPU_KIND_HEX_PLANE = 6,           /* p0 outer hex radius              */
    MLGPU_KIND_ANNULUS = 7,             /* p0 outer radius, p1 inner radius */
    MLGPU_KIND_PLANE_DUAL_SPHERICAL = 8,/* p0 x hight, p1 y width           */
    MLGPU_KIND_CYLINDER = 9,            /* p0 radius, p1 length (along z)   */
    MLGPU_KIND_TRIANGLE = 10,           /* p0..p5 Ax Ay Bx By Cx Cy         */
    MLGPU_KIND_SPHERE = 11,             /* p0 radius                        */
    MLGPU_KIND_ELLIPTICAL_CAP_HEX = 12, /* p0 focal length, p1 ratio of the
                                           x to the y curvature radius, p2
                                           hex rotation (kept, unused as in
                                           merlict), p3 outer hex radius    */
    MLGPU_NUM_KINDS = 13
};

enum mlgpu_rot_mode {
    MLGPU_ROT_XYZ = 0,   /* rot[0..2] = Tait-Bryan x, y, z angles (Rot3)   */
    MLGPU_ROT_AXIS = 1,  /* rot[0..2] = axis, rot[3] = angle              */
    MLGPU_ROT_AXIS_STORED = 2 /* rot[0..2] = the axis as a Rot3 holds it
                           (Rot3::rot_axis: already divided by its norm by
                           Rot3::set, Rot3.cpp:25-33), rot[3] = angle; only
                           HomTra3's own normalisation follows
                           (HomTra3.cpp:67-70) */
};

/* surface slots, same order as the SurfaceEntity members */
enum mlgpu_surface_slot {
    MLGPU_OUTER_REFLECTION = 0,
    MLGPU_INNER_REFLECTION = 1,
    MLGPU_OUTER_REFRACTION = 2,
    MLGPU_INNER_REFRACTION = 3,
    MLGPU_OUTER_ABSORPTION = 4,
    MLGPU_INNER_ABSORPTION = 5,
    MLGPU_NUM_SURFACE_SLOTS = 6
};

#define MLGPU_NUM_PARAMS 6
#define MLGPU_DEFAULT_FUNCTION (-1) /* slot uses merlict's DEFAULT_* table */
#define MLGPU_NO_FRAME (-1)
#define MLGPU_NO_SENSOR (-1)

typedef struct mlgpu_scene_desc {
    int32_t n_frames;
    const int32_t* mother;        /* [n]   -1 for the root, else < own index */
    const double* pos;            /* [n*3] position in mother               */
    const int32_t* rot_mode;      /* [n]   mlgpu_rot_mode                   */
    const double* rot;            /* [n*4]                                  */
    const int32_t* kind;          /* [n]   mlgpu_kind                       */
    const double* params;         /* [n*MLGPU_NUM_PARAMS]                   */
    const int32_t* surface;       /* [n*6] function index or -1 (default)   */
    const int32_t* allowed_frame; /* [n]   restriction frame or -1          */
    const int32_t* sensor_id;     /* [n]   sensor id or -1                  */
    int32_t n_functions;
    const int32_t* func_start;    /* [n_functions+1] offsets into func_x/y  */
    const double* func_x;         /* argument (wavelength), table order     */
    const double* func_y;         /* value                                  */
} mlgpu_scene_desc;

/* Owned, growable scene description produced by the scene builders. */
typedef struct mlgpu_desc mlgpu_desc;

void mlgpu_desc_destroy(mlgpu_desc* d);
/* View of an owned description; valid until the description is destroyed. */
int mlgpu_desc_view(const mlgpu_desc* d, mlgpu_scene_desc* out);
/* an owned copy of a description given as arrays (validated) */
int mlgpu_desc_create(const mlgpu_scene_desc* view, mlgpu_desc** out);
/* Frame::frame2world() of every frame after merlict's tree initialisation
 * (merlict/Frame.cpp:26-31,125-147): out (host) holds 12 doubles per frame,
 * the row-major 3 x 4 transform */
int mlgpu_desc_frame_poses(const mlgpu_scene_desc* desc, double* out);
/* Write / read the binary description file ("MLGPUSC1"). */
int mlgpu_desc_save(const mlgpu_desc* d, const char* path);
int mlgpu_desc_load(const char* path, mlgpu_desc** out);

/* Scene builders (host only).  Each returns a new owned description.
 * Lens test bench: merlict BiConvexLens (lens_maker.cpp:15-34 curvature for
 * focal length f, aperture radius r, index n) at z = 0 and a 19-cell hexagonal
 * HexPlane flower at z = -f (pitch p).  */
int mlgpu_build_lens_scene(double focal_length, double aperture_radius,
                           double refractive_index, double cell_pitch,
                           mlgpu_desc** out);
/* FACT telescope as merlict_fact/apps/fact_propagation.cpp:282-341 builds
 * it: segmented_imaging_reflector (30 facets, f = 4.889 m) with the app's
 * fact_mirror_reflectivity() table (fact_propagation.cpp:145-259) on the
 * facets' outer side, a shield disc, and the 1440 hexagonal pixels of the
 * app's pixel table at the focal plane, sensor id = FACT chid;
 * camera 1 adds a hexagonal PMMA light guide (15 mm) in front of each pixel
 * (entry hexagons as wide as the tightest pair of neighbouring pixels allows);
 * camera 0 has bare pixels; camera 2 replaces the camera by one absorbing
 * sensor disc (sensor 0, radius 0.33 m) at the guides' entrance plane, 15 mm
 * in front of the pixels: the same dish and shield, so a photon reaches the
 * entrance plane there exactly as it does in the camera-1 scene under the
 * same draws (the surrogate's focal-plane targets). */
int mlgpu_build_fact_scene(int camera, mlgpu_desc** out);
/* Plenoscope: segmented reflector (f = 106.5 m, 71 m aperture) and the
 * light-field sensor of light_field_sensor/Factory.cpp (8,443 lenslets,
 * 515,023 lixels for num_paxel_on_pixel_diagonal = 9). */
int mlgpu_build_plenoscope_scene(double focal_length,
                                 double max_outer_aperture_radius,
                                 double min_inner_aperture_radius,
                                 double facet_inner_hex_radius,
                                 double gap_between_facets,
                                 double max_FoV_diameter_deg,
                                 double hex_pixel_FoV_flat2flat_deg,
                                 int num_paxel_on_pixel_diagonal,
                                 double housing_overhead,
                                 double lens_refraction,
                                 double mirror_reflection,
                                 mlgpu_desc** out);

/* --------------------------------------------------------------- photons
 * Structure of arrays.  For mlgpu_propagate all pointers are device pointers.
 * Direction need not be normalised (Ray.cpp:16-20 normalises). */
typedef struct mlgpu_photons {
    const double* support_x;
    const double* support_y;
    const double* support_z;
    const double* direction_x;
    const double* direction_y;
    const double* direction_z;
    const double* wavelength;
} mlgpu_photons;

/* The same arrays, written by mlgpu_generate_photons. */
typedef struct mlgpu_photon_buffers {
    double* support_x;
    double* support_y;
    double* support_z;
    double* direction_x;
    double* direction_y;
    double* direction_z;
    double* wavelength;
} mlgpu_photon_buffers;

/* merlict::Interaction (RayForPropagation.h:14-28) */
enum mlgpu_interaction {
    MLGPU_PRODUCTION = 0,
    MLGPU_ABSORPTION_IN_VOID = 1,
    MLGPU_ABSORPTION_IN_MEDIUM = 2,
    MLGPU_ABSORPTION_ON_SURFACE = 3,
    MLGPU_FRESNEL_REFLECTION_ON_SURFACE = 4,
    MLGPU_REFLECTION_ON_SURFACE = 5,
    MLGPU_REFRACTION_TO_OUTSIDE = 6,
    MLGPU_REFRACTION_TO_INSIDE = 7
};

/* per-photon error flags */
enum mlgpu_photon_error {
    MLGPU_PHOTON_OK = 0,
    MLGPU_PHOTON_BAD_WAVELENGTH = 1,     /* lambda <= 0 (Photon.cpp:20-28)   */
    MLGPU_PHOTON_OUT_OF_TABLE = 2        /* Func1 range (Limits.cpp:37-46)   */
};

/* Results, one entry per photon (device pointers for mlgpu_propagate). */
typedef struct mlgpu_results {
    int32_t* final_frame;       /* frame index of final intersection, -1 void */
    int32_t* sensor_id;         /* sensor of the final frame, -1 none        */
    uint8_t* final_interaction; /* mlgpu_interaction                         */
    uint8_t* num_interactions;  /* history length incl. production           */
    uint8_t* error;             /* mlgpu_photon_error                        */
    double* time_of_flight;     /* Photon::time_of_flight (Photon.cpp:44-56) */
    float* x;                   /* final position, object frame              */
    float* y;
    float* cx;                  /* -incident direction, object frame         */
    float* cy;
    double* accumulated_distance; /* RayForPropagation::accumulated_distance
                                     (RayForPropagation.cpp:91-98); may be NULL */
} mlgpu_results;

typedef struct mlgpu_config {
    uint32_t max_num_interactions_per_photon; /* PropagationConfig.cpp:7 = 5 */
    uint64_t seed;             /* Philox key                                 */
    uint64_t photon_id_offset; /* Philox stream of photon i is offset + i    */
} mlgpu_config;

/* ------------------------------------------------------------- device scene */
typedef struct mlgpu_scene mlgpu_scene;

/* Build the traversal hierarchy on the host and upload it to the current
 * device.  Replaces init_tree_based_on_mother_child_relations()
 * (merlict/Frame.cpp:125-130) for the GPU.  MLGPU_ERR_SCENE when the scene
 * exceeds the device layout: a tree deeper than 28 levels or more than
 * 32,767 functions. */
int mlgpu_scene_create(const mlgpu_scene_desc* desc, mlgpu_scene** out);
void mlgpu_scene_destroy(mlgpu_scene* scene);
/* hierarchy statistics: nodes, leaves, device bytes */
int mlgpu_scene_stats(const mlgpu_scene* scene, int64_t* num_nodes,
                      int64_t* num_leaves, int64_t* device_bytes);

/* Kernel schedule of mlgpu_propagate:
 *   FUSED      one lane carries a photon through all its bounces;
 *   WAVEFRONT  per bounce level a trace pass and a shade kernel over a
 *              compacted queue of live photons (workspace owned by the scene,
 *              about 210 B per photon; calls sharing a scene are serialised
 *              on it);
 *   PERSISTENT a grid of resident waves; a lane carries one photon through
 *              all its bounces with its state in registers and, when the
 *              photon ends, takes the next one, so waves stay full; traces
 *              advance in quanta and the lanes of a wave shade together;
 *   AUTO       WAVEFRONT from 65,536 photons on. */
enum mlgpu_schedule {
    MLGPU_SCHEDULE_AUTO = 0,
    MLGPU_SCHEDULE_FUSED = 1,
    MLGPU_SCHEDULE_WAVEFRONT = 2,
    MLGPU_SCHEDULE_PERSISTENT = 3
};

/* Launch parameters of mlgpu_propagate; they never change results.  A new
 * scene starts with {AUTO, 32, 48, 0, 32}.
 *   schedule       an mlgpu_schedule
 *   trace_budget   WAVEFRONT: traversal steps a ray gets in the first trace
 *                  pass of a level; the unfinished rays resume, packed, in a
 *                  second pass.  <= 0: one pass.
 *                  PERSISTENT: traversal steps per quantum, between which a
 *                  wave checks whether enough lanes wait to shade; <= 0:
 *                  traces run to their end.
 *   blocks_per_cu  WAVEFRONT, PERSISTENT: workgroups per compute unit of the
 *                  grid-stride and persistent kernels, 1 to 64.
 *   tail_level     WAVEFRONT: from this bounce level on (>= 1) a photon's
 *                  remaining bounces run in one lane, its state kept in
 *                  registers; <= 0: every level is a trace pass and a shading
 *                  pass of its own.
 *   shade_lanes    PERSISTENT: lanes of a wave whose traces have ended before
 *                  the wave shades them (or all of its busy lanes, if fewer),
 *                  1 to 64. */
typedef struct mlgpu_tuning {
    int32_t schedule;
    int32_t trace_budget;
    int32_t blocks_per_cu;
    int32_t tail_level;
    int32_t shade_lanes;
} mlgpu_tuning;
int mlgpu_scene_set_tuning(mlgpu_scene* scene, const mlgpu_tuning* tuning);
int mlgpu_scene_get_tuning(const mlgpu_scene* scene, mlgpu_tuning* tuning);

/* In-library autotuner.  Times mlgpu_propagate of the given photons (device
 * arrays) under each candidate tuning: FUSED; WAVEFRONT with trace
 * budget in {0, 12, 16, 24, 32, 48} x blocks per CU in {8, 16, 24, 32, 48}
 * x tail level in {0, 1, 2}; PERSISTENT with quantum (trace_budget) in
 * {0, 8, 16, 32} x blocks per CU in {4, 8, 16} x shade lanes in
 * {1, 16, 32, 48};
 * one warm-up call and `repeats` (>= 1) timed calls each, between HIP events
 * on `stream`.  Keeps the fastest on the scene and returns it in *best and
 * its mean time per call in *best_ms (either may be NULL).  The results of
 * the trial calls go to a temporary device buffer; results never depend on
 * the tuning.  Synchronises `stream`.  No counterpart in the reference. */
int mlgpu_scene_autotune(mlgpu_scene* scene, int32_t world_frame,
                         const mlgpu_photons* photons, int64_t n,
                         const mlgpu_config* config, int32_t repeats,
                         mlgpu_tuning* best, double* best_ms, void* stream);

/* Propagate n photons, one Philox stream per photon.  Replaces
 * merlict::propagate_photons_in_frame_with_config (merlict/Photons.cpp:9-22)
 * and its multi-thread twin (merlict_multi_thread.cpp:28-55).  world_frame
 * is the frame the propagation starts in (usually 0).  Asynchronous on
 * `stream` (a hipStream_t, NULL = default stream). */
int mlgpu_propagate(const mlgpu_scene* scene, int32_t world_frame,
                    const mlgpu_photons* photons, int64_t n,
                    const mlgpu_config* config, mlgpu_results* results,
                    void* stream);

/* mlgpu_propagate without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The wavefront
 * workspace is a static device array of the library (about 58 MB, held on
 * every GPU the library runs on until the process ends) instead of the
 * scene's grown workspace; results equal mlgpu_propagate's bit for bit.
 * At most MLGPU_NOALLOC_MAX_PHOTONS photons per call (larger n: an argument
 * error).  Every call on one GPU shares that array: calls that overlap on
 * one GPU (other streams, other threads) race; calls on one stream are
 * ordered and safe. */
#define MLGPU_NOALLOC_MAX_PHOTONS 262144
int mlgpu_propagate_noalloc(const mlgpu_scene* scene, int32_t world_frame,
                            const mlgpu_photons* photons, int64_t n,
                            const mlgpu_config* config,
                            mlgpu_results* results, void* stream);

/* One entry of a photon's history after its production: the arguments of
 * the Intersection that RayForPropagation::push_back_intersection_and_
 * interaction appends (merlict/Intersection.h:24-29, RayForPropagation.h:
 * 47-50) and the interaction.  A photon absorbed in void has the void
 * intersection of RayAndFrame.cpp:160-167: frame -1, position the world
 * point 1e4 m along the ray, normal the ray direction, distance 1e4. */
typedef struct mlgpu_bounce {
    double position[3];   /* Intersection::position_in_object_frame        */
    double normal[3];     /* surface normal in the object frame            */
    double distance;      /* Intersection::distance_to_ray_support         */
    int32_t frame;        /* frame of the intersection, -1 void            */
    uint8_t interaction;  /* mlgpu_interaction                             */
    uint8_t from_outside; /* Intersection::from_outside_to_inside          */
    uint8_t reserved[2];
} mlgpu_bounce;

/* Device arrays of the histories of n photons. */
typedef struct mlgpu_history {
    mlgpu_bounce* bounces; /* [n * capacity]: photon i's entries from
                              i * capacity on, in order                     */
    int32_t capacity;      /* >= config->max_num_interactions_per_photon    */
    int32_t* num_bounces;  /* [n] entries written; a photon with an error
                              keeps those before the bounce that failed     */
    double* final_ray;     /* [n * 6] or NULL: support and direction of the
                              photon's last ray (world frame), as
                              Ray::support / Ray::direction read after
                              merlict's propagation                         */
} mlgpu_history;

/* mlgpu_propagate that also records every photon's full history, so a
 * caller can rebuild merlict::Photon objects bounce for bounce (the drop-in
 * of propagate_photons_in_frame_with_config needs them: Photon::
 * time_of_flight and Sensor::assign_photon read the history).  Same results
 * and Philox streams as mlgpu_propagate; one lane per photon.  Asynchronous
 * on `stream`. */
int mlgpu_propagate_history(const mlgpu_scene* scene, int32_t world_frame,
                            const mlgpu_photons* photons, int64_t n,
                            const mlgpu_config* config,
                            mlgpu_results* results,
                            const mlgpu_history* history, void* stream);

/* ------------------------------------------------------------ photon sources
 * Synthetic photons drawn on the GPU from Philox4x32-10 stream 1 of
 * (seed, photon id) -- propagation uses stream 0 -- with only +, -, *, /
 * and sqrt in IEEE double, so a CPU restatement yields the same photons bit
 * for bit.  Replaces the reference's sources: photon_source::
 * parallel_towards_z_from_xy_disc and point_like_towards_z_opening_angle_
 * num_photons (merlict/Photons.cpp:85-125), the calibration photons
 * (merlict_portal_plenoscope/calibration/Calibrator.cpp:37-102) and the pose
 * of transform_photons (merlict_json/json.cpp:512-519).
 *
 * Per photon: a support on the xy plane (the origin, or uniform on a disc by
 * the reference's rejection rule, Generator.cpp:22-34), a direction about +z,
 * a wavelength; the photon then starts start_distance back along its
 * direction, and the pose (rot row major, pos) maps support and direction:
 * s' = rot s + pos, d' = rot d. */
enum mlgpu_support {
    MLGPU_SUPPORT_POINT = 0,
    MLGPU_SUPPORT_DISC = 1
};
enum mlgpu_direction {
    MLGPU_DIRECTION_AXIS = 0,      /* +z */
    MLGPU_DIRECTION_TILT = 1,      /* tilt ~ U[0, angle], azimuth ~ U[0, 2 pi) */
    MLGPU_DIRECTION_GAUSSIAN = 2,  /* tan of the x and y tilts ~ N(0, angle) */
    MLGPU_DIRECTION_CAP = 3,       /* uniform on the cap zenith <= angle */
    MLGPU_DIRECTION_RING = 4       /* zenith ~ |N(angle, angle_width)|, azimuth ~ U[0, 2 pi):
                                      a Cherenkov ring of radius angle */
};
enum mlgpu_spectrum {
    MLGPU_SPECTRUM_INVERSE_SQUARE = 0,  /* pdf ~ lambda^-2 on [lambda_min, lambda_max] */
    MLGPU_SPECTRUM_UNIFORM = 1          /* uniform on [lambda_min, lambda_max] */
};
typedef struct mlgpu_source {
    int32_t support;        /* mlgpu_support */
    int32_t direction;      /* mlgpu_direction */
    double radius;          /* disc radius (m) */
    double angle;           /* rad; TILT, CAP and RING: [0, pi/2] */
    double lambda_min;      /* m; lambda_min == lambda_max: monochromatic, */
    double lambda_max;      /* else drawn by `spectrum` */
    double start_distance;  /* m */
    double rot[9];
    double pos[3];
    double angle_width;     /* rad; RING: the zenith's standard deviation */
    int32_t spectrum;       /* mlgpu_spectrum */
    int32_t pad;
} mlgpu_source;
int mlgpu_generate_photons(const mlgpu_source* source, int64_t n, uint64_t seed,
                           uint64_t photon_id_offset, const mlgpu_photon_buffers* out,
                           void* stream);

/* ------------------------------------------------------- sensor arrivals
 * Arrival-time histograms per sensor.  Sensor::assign_photon counts a photon
 * on the sensor of its final frame whatever its final interaction
 * (merlict/sensor/Sensor.cpp:14-30); here counts[s * num_bins + b] += 1 for
 * every photon without error flag whose sensor_id s is in [0, num_sensors)
 * and whose arrival time t = time_of_flight + jitter falls in bin
 * b = floor((t - t_start) / bin_width) < num_bins.  jitter_sigma > 0 adds
 * N(0, jitter_sigma) drawn from Philox stream 2 of
 * (seed, photon_id_offset + i).  Counts accumulate; clear them first. */
typedef struct mlgpu_histogram {
    uint32_t* counts;          /* device, num_sensors * num_bins */
    int32_t num_sensors;
    int32_t num_bins;
    double t_start;            /* s */
    double bin_width;          /* s */
    double jitter_sigma;       /* s */
    uint64_t seed;
    uint64_t photon_id_offset;
} mlgpu_histogram;
int mlgpu_histogram_arrivals(const mlgpu_results* results, int64_t n,
                             const mlgpu_histogram* hist, void* stream);

/* ------------------------------------------------------------ photon bunches
 * Many bunches of synthetic photons in one call, all on the device: bunch b
 * draws photons 0 .. num_photons - 1 of its own source under its own seed
 * (its source pose carries the bunch's pointing, as transform_photons does,
 * merlict_json/json.cpp:512-519), the bunches are propagated together
 * (photon i of their concatenation uses Philox stream 0 of (config seed,
 * config photon_id_offset + i)), and, when hist is not NULL, their arrivals
 * are binned into it (photon i as hist photon_id_offset + i).  results, when
 * not NULL, receives the per-photon results of the concatenation (device
 * arrays of the total length).  Replaces the per-event loop of
 * plenoscope_propagation.cpp:160-220 (read, transform, propagate, record).
 * The photons live in the scene's device memory; calls sharing a scene are
 * serialised on it. */
typedef struct mlgpu_bunch {
    mlgpu_source source;
    uint64_t seed;
    int64_t num_photons;
} mlgpu_bunch;
/* The photons of bunches, concatenated in bunch order into out (capacity
 * photons): photon j of bunch b is photon j of its source under its seed,
 * as mlgpu_generate_photons draws it with photon_id_offset 0.  The bunch
 * table is host memory.  Synchronises `stream`.  No counterpart in the
 * reference (its apps build each event's photons on the host). */
int mlgpu_generate_bunches(const mlgpu_bunch* bunches, int32_t num_bunches,
                           const mlgpu_photon_buffers* out, int64_t capacity,
                           void* stream);
int mlgpu_propagate_bunches(const mlgpu_scene* scene, int32_t world_frame,
                            const mlgpu_bunch* bunches, int32_t num_bunches,
                            const mlgpu_config* config,
                            const mlgpu_histogram* hist, mlgpu_results* results,
                            void* stream);

/* Bunch plans: the launch sequence of mlgpu_propagate_bunches (bunch table
 * up, photon draw, propagation with the scene's tuning at creation --
 * the PERSISTENT work-queue kernel when that is the scene's schedule,
 * WAVEFRONT otherwise --, histogram into hist when it is not NULL) captured once in a
 * hipGraph for batches of up to max_bunches bunches and max_photons photons,
 * and replayed per batch: the kernels read the batch's photon count, seed
 * and photon id offset from a device parameter block that each run uploads
 * with the bunch table from pinned memory, so batches of any size up to the
 * capacities replay the same graph.  A run gives the results (and histogram
 * counts, which accumulate) that mlgpu_propagate_bunches gives for the same
 * batch and config; its max_num_interactions_per_photon must be the plan's.
 * Runs on one plan are serialised (a run waits for the previous one before
 * rewriting the staging buffer).  hist->counts and the scene must outlive
 * the plan.  mlgpu_bunch_plan_results gives the device arrays holding the
 * last run's per-photon results (valid once the run has finished).  No
 * counterpart in the reference. */
typedef struct mlgpu_bunch_plan mlgpu_bunch_plan;
int mlgpu_bunch_plan_create(const mlgpu_scene* scene, int32_t world_frame,
                            int32_t max_bunches, int64_t max_photons,
                            const mlgpu_config* config,
                            const mlgpu_histogram* hist, mlgpu_bunch_plan** out);
int mlgpu_bunch_plan_run(mlgpu_bunch_plan* plan, const mlgpu_bunch* bunches,
                         int32_t num_bunches, const mlgpu_config* config,
                         void* stream);
int mlgpu_bunch_plan_results(const mlgpu_bunch_plan* plan, mlgpu_results* out);
void mlgpu_bunch_plan_destroy(mlgpu_bunch_plan* plan);

/* ------------------------------------------------------------- surrogate
 * A learned surrogate of the propagator: an MLP 5 -> 256 -> 256 -> 256 ->
 * 256 -> 57 from a photon's (support x, y, direction cx, cy, wavelength),
 * normalised, to the survival logit and an 8-component Gaussian mixture
 * (weight logit, mean x y t, log sigma x y t) over the sensor-plane
 * position and arrival time of a photon that reaches a sensor.  bf16
 * weights and activations on the matrix cores with fp32 accumulation,
 * fp32 master parameters and Adam.  The flat fp32 parameter vector is
 * W0 b0 W1 b1 ... W4 b4, each W row-major (out x in).  Features are n x 5
 * fp32, targets n x 4 fp32 (detected 0/1, x, y, t), device pointers.  No
 * counterpart in the reference: merlict only has the plain propagator,
 * which trains and judges this model. */
typedef struct mlgpu_surrogate mlgpu_surrogate;
int64_t mlgpu_surrogate_num_params(void);
int mlgpu_surrogate_create(uint64_t seed, mlgpu_surrogate** out);
void mlgpu_surrogate_destroy(mlgpu_surrogate* s);
int mlgpu_surrogate_get_params(const mlgpu_surrogate* s, float* dst, void* stream);
/* also resets the optimiser state */
int mlgpu_surrogate_set_params(mlgpu_surrogate* s, const float* src, void* stream);
/* Inference precision of mlgpu_surrogate_forward and
 * mlgpu_propagate_hybrid (training stays bf16 with fp32 masters):
 * MLGPU_SURROGATE_BF16 (the default) or MLGPU_SURROGATE_FP8, where layers
 * 1.. run on the fp8 matrix instructions (e4m3 weights, each row scaled by
 * a power of two to the e4m3 range; e4m3 ReLU activations clamped to 448;
 * fp32 accumulation) and layer 0 stays bf16.  The fp8 weight copies are
 * refreshed with the bf16 ones (create, set_params, train_step).  No
 * counterpart in the reference. */
#define MLGPU_SURROGATE_BF16 0
#define MLGPU_SURROGATE_FP8 1
int mlgpu_surrogate_set_inference(mlgpu_surrogate* s, int32_t precision);
int32_t mlgpu_surrogate_get_inference(const mlgpu_surrogate* s);
/* out: n x 57 fp32 */
int mlgpu_surrogate_forward(mlgpu_surrogate* s, const float* features, int64_t n,
                            float* out, void* stream);
/* mean loss over the batch (*loss, device) and its gradient (grads,
 * device, num_params; either may be NULL) */
int mlgpu_surrogate_gradients(mlgpu_surrogate* s, const float* features,
                              const float* targets, int64_t n, float* grads,
                              float* loss, void* stream);
/* one Adam step (beta 0.9 / 0.999) on the batch; *loss (device, may be
 * NULL) receives the summed loss before the step */
int mlgpu_surrogate_train_step(mlgpu_surrogate* s, const float* features,
                               const float* targets, int64_t n,
                               float learning_rate, float* loss, void* stream);

/* The surrogate's domain and units.  Features of a photon: x / r, y / r of
 * its support (r = support_radius), its direction's x and y over
 * sin(max_angle), and 2 (1/lambda - 1/lambda_max) / (1/lambda_min -
 * 1/lambda_max) - 1.  In the domain: support within r of the axis,
 * direction within max_angle of -z, lambda in [lambda_min, lambda_max].
 * Targets: the sensor-plane position (mlgpu_surrogate_prepare) relative
 * to the nominal image point image_scale (dx / dz, dy / dz) of the
 * photon's direction, over position_scale, and (time_of_flight -
 * time_offset) / time_scale. */
typedef struct mlgpu_surrogate_domain {
    double support_radius;
    double max_angle;
    double lambda_min;
    double lambda_max;
    double image_scale;
    double position_scale;
    double time_offset;
    double time_scale;
} mlgpu_surrogate_domain;

/* Features (n x 5) of photons and, when results is not NULL, targets
 * (n x 4: detected, x, y, t) of their propagation; frame_poses (device) is
 * mlgpu_desc_frame_poses of the scene.  The position target of a detected
 * photon is where it crosses the focal plane: the world position of its
 * result in focal_results, the same photons propagated under the same
 * config through a scene whose sensor is that plane (focal_poses: that
 * scene's frame poses), when that result is a detection; else, or when
 * focal_results is NULL, the world centre of the sensor frame it ends on.
 * All device pointers. */
int mlgpu_surrogate_prepare(const mlgpu_surrogate_domain* domain,
                            const mlgpu_photons* photons,
                            const mlgpu_results* results,
                            const double* frame_poses,
                            const mlgpu_results* focal_results,
                            const double* focal_poses, int64_t n,
                            float* features, float* targets, void* stream);

/* Sensor cells on a hexagonal lattice in the sensor plane.  Lattice site
 * (i, j) is origin + i a1 + j a2 with cell id
 * cell[(i + radius) * (2 radius + 1) + (j + radius)] (device; -1: none).
 * Without centres (centre NULL) a point belongs to the cell of its nearest
 * site.  With centres (device, 2 floats per cell id: the cell's own centre,
 * which may sit off its site, e.g. the FACT camera's pixel table) a point
 * belongs to the cell whose centre is nearest among the cells of its
 * nearest site and that site's six neighbours whose hexagon holds it
 * (inner radius inner_radius, flat sides facing a1 and a1 turned by 60 and
 * 120 degrees; inner_radius 0: every one of them); a point in no such
 * hexagon belongs to none.  Hexagons that do not overlap give cells with
 * gaps between them; hexagons a little larger than half the widest spacing
 * give the nearest-centre cells, closed inside the lattice and ending at
 * its rim. */
typedef struct mlgpu_hex_cells {
    double a1[2];
    double a2[2];
    double origin[2];
    int32_t radius;
    const int32_t* cell;
    const float* centre;
    double inner_radius;
} mlgpu_hex_cells;

/* Factor (> 0, default 1) on the training loss's gradient with respect to
 * the survival logit s, in mlgpu_surrogate_gradients and
 * mlgpu_surrogate_train_step; the loss value is unchanged.  The survival's
 * dependence on the wavelength (the mirror reflectivity, a few per cent)
 * is a small signal next to the mixture's position terms, and the shared
 * layers learn it only when its gradient is scaled up (10 in
 * surrogate.train).  A stationary point of the survival's own parameters
 * is the same for every factor.  No counterpart in the reference. */
int mlgpu_surrogate_set_survival_scale(mlgpu_surrogate* s, float scale);

/* The sensor cells of the surrogate's draws (mlgpu_propagate_hybrid: a
 * drawn position off the cells is a photon that misses the sensors).  The
 * model's probability of a detection on cell c is sigmoid(s) times the
 * mixture's mass on c.  While a region is set, the training loss of a
 * photon that was not detected is -log(1 - sigmoid(s) Z), Z the mixture's
 * mass on all cells, from a 32-point Gauss-Laguerre x angular rule per
 * component (rotated per photon and step); without a region it is
 * -log(1 - sigmoid(s)).  Domain and cells are copied; cells->cell (device)
 * must stay valid while set.  NULL cells clears the region. */
int mlgpu_surrogate_set_region(mlgpu_surrogate* s, const mlgpu_surrogate_domain* domain,
                               const mlgpu_hex_cells* cells);

/* The surrogate + fallback hybrid: every photon in the domain whose
 * mixture entropy bound (entropy of the weights plus their mean component
 * entropy, nats, normalised units) is at most max_entropy is drawn from the
 * surrogate (survival, component, position, time from Philox stream 3 of
 * (seed, config photon_id_offset + i)) and binned into hist at the cell
 * its position falls in, if any; every other photon is propagated by the
 * plain propagator and binned from its result, under its own photon id
 * (config photon_id_offset + i for propagation, hist photon_id_offset + i
 * for the jitter), so it gets exactly what mlgpu_propagate and
 * mlgpu_histogram_arrivals give photon i of the same batch, and the
 * histogram is the same from run to run.  A max_entropy of -inf sends every
 * photon to the propagator.  *num_fallback (host, may be NULL) receives the
 * number of propagated photons.  Synchronises the stream once per chunk of
 * 2^24 photons. */
typedef struct mlgpu_hybrid_config {
    double max_entropy;
    uint64_t seed;
} mlgpu_hybrid_config;
int mlgpu_propagate_hybrid(const mlgpu_scene* scene, mlgpu_surrogate* surrogate,
                           const mlgpu_surrogate_domain* domain,
                           const mlgpu_hex_cells* cells, int32_t world_frame,
                           const mlgpu_photons* photons, int64_t n,
                           const mlgpu_config* config,
                           const mlgpu_hybrid_config* hybrid,
                           const mlgpu_histogram* hist, int64_t* num_fallback,
                           void* stream);

/* CORSIKA EventIO runs on the host: replaces eventio::Run
 * (merlict_corsika/eventio.h:124-146, eventio.cpp:282-412), the reader whose
 * bunches the reference's apps turn into photons.  Open reads the run
 * header (the 273-float block, the input card, the telescope positions as
 * (x, y, z, r) records) and looks ahead to the first event's bunches.
 * has_events is 1 until the run end has been read (has_still_events_left);
 * next_event moves to the next event (its 273-float header, the number of
 * telescope offsets and of bunches), event_data copies its offsets
 * (toff, xoff, yoff, weight records) and bunches (n x 8 float32: x, y
 * [cm], cx, cy, time [ns], emission height [cm], weight, wavelength [nm];
 * compact bunches scaled as the reference does), run_end the run end
 * block once read.  Buffers are host memory.  A file the reader can not
 * follow is an MLGPU_ERR_FORMAT (the reference throws), next_event past the
 * run end an MLGPU_ERR_ARGUMENT. */
typedef struct mlgpu_eventio_run mlgpu_eventio_run;
int mlgpu_eventio_open(const char* path, mlgpu_eventio_run** out);
void mlgpu_eventio_close(mlgpu_eventio_run* run);
int mlgpu_eventio_run_header(const mlgpu_eventio_run* run, float* raw273, int32_t* num_telescopes,
                             int64_t* input_card_length);
int mlgpu_eventio_run_info(const mlgpu_eventio_run* run, float* telescope_positions, char* input_card);
int mlgpu_eventio_has_events(const mlgpu_eventio_run* run);
int mlgpu_eventio_next_event(mlgpu_eventio_run* run, float* header273, int32_t* num_offsets,
                             int64_t* num_bunches);
int mlgpu_eventio_event_data(const mlgpu_eventio_run* run, float* offsets, float* bunches);
int mlgpu_eventio_run_end(const mlgpu_eventio_run* run, float* raw273);

/* CORSIKA event tapes on the host: replaces event_tape::Run
 * (merlict_corsika/event_tape.h:23-36, event_tape.cpp:10-45), a tar of
 * ddddddddd/RUNH.float32, then per event ddddddddd/ddddddddd/EVTH.float32
 * and its bunch blocks ddddddddd/ddddddddd/ddddddddd.cer.x8.float32.  Open
 * reads the run header (a tape without one is an MLGPU_ERR_FORMAT) and the
 * first event header; next_event reads the event's bunches (n x 8
 * float32, host memory) and the next event header; has_events is 0 once
 * no event header follows (a header or block that fails the reference's
 * checks ends the run, as in the reference). */
typedef struct mlgpu_event_tape mlgpu_event_tape;
int mlgpu_event_tape_open(const char* path, mlgpu_event_tape** out);
void mlgpu_event_tape_close(mlgpu_event_tape* tape);
int mlgpu_event_tape_run_header(const mlgpu_event_tape* tape, float* raw273);
int mlgpu_event_tape_has_events(const mlgpu_event_tape* tape);
int mlgpu_event_tape_next_event(mlgpu_event_tape* tape, float* header273, int64_t* num_bunches);
int mlgpu_event_tape_event_bunches(const mlgpu_event_tape* tape, float* bunches);

/* CORSIKA photon bunches to photons.  Replaces merlict::EventIoPhotonFactory
 * (merlict_corsika/PhotonFactory.cpp:8-148) as the reference's apps apply it
 * to every bunch of an event (plenoscope_propagation.cpp:196-203).
 * bunches: n x 8 float (device, 16-byte aligned): x, y [cm], cx, cy (the
 * direction cosines of the motion), arrival time on the ground [ns],
 * emission height [cm], weight, wavelength [nm] (either sign).  A bunch of
 * weight w yields floor(w) photons plus one more when the uniform (w0 >> 8)
 * 2^-24 of Philox stream 5 of (seed, id_offset + bunch) is <= fmod(w, 1);
 * each starts 1e3 m + c t above its ground point (x, y in metres) against
 * its direction (cx, cy, -sqrt(1 - cx^2 - cy^2)), wavelength |lambda| nm.
 * The photons go to out in bunch order, their bunch index to truth_id (may
 * be NULL); *num_photons (host) receives their number.  A weight outside
 * [0, 16] (the reference's BadPhotonWeight) or a capacity below the number
 * of photons is an MLGPU_ERR_ARGUMENT (then nothing is written; the number
 * is still reported for a capacity error).  Synchronises `stream`. */
int mlgpu_corsika_photons(const float* bunches, int64_t n, uint64_t seed,
                          uint64_t id_offset, const mlgpu_photon_buffers* out,
                          int32_t* truth_id, int64_t capacity,
                          int64_t* num_photons, void* stream);

/* ------------------------------------------------------- signal chain
 * Per-sensor photon pipelines (merlict_signal_processing/PipelinePhoton.cpp:
 * 25-66): the detected photons (sensor_id in [0, num_sensors), no error
 * flag) of n results, grouped by sensor and sorted by arrival time (ties
 * by photon index): pipeline s is entries offsets[s] .. offsets[s + 1] - 1
 * of arrival_time, wavelength and truth_id (the photon's truth id, or its
 * index when truth_id is NULL).  All device arrays; offsets has
 * num_sensors + 1 entries.  *num_photons (host) receives the number of
 * detected photons; a capacity below it is an MLGPU_ERR_ARGUMENT.
 * Synchronises `stream`. */
typedef struct mlgpu_pipelines {
    int64_t* offsets;
    double* arrival_time;
    double* wavelength;
    int32_t* truth_id;
    int64_t capacity;
} mlgpu_pipelines;
int mlgpu_sensor_pipelines(const mlgpu_results* results, const double* wavelength,
                           const int32_t* truth_id, int64_t n, int32_t num_sensors,
                           const mlgpu_pipelines* out, int64_t* num_photons,
                           void* stream);

/* Photo-electric conversion of every pipeline (PhotoElectricConverter.cpp:
 * 10-95): a photon becomes a pulse when the quantum efficiency at its
 * wavelength (a Func1 table, host arrays qe_x, qe_y) is >= a uniform; each
 * pulse is followed by crosstalk pulses (truth -202) while crosstalk >= a
 * uniform; then accidental pulses (truth -201) at the times of a Poisson
 * process of dark_rate over [0, exposure_time).  Photon j of the pipelines
 * draws its uniforms from Philox stream 6 of (seed, j) and sensor s its
 * accidental pulses from stream 8 of (seed, s), each in the reference's
 * order.  Pulses of sensor s (its photons' pulses in pipeline order, then
 * its accidental pulses) are entries offsets[s] .. offsets[s + 1] - 1 of
 * out; *num_pulses (host)
 * receives their number.  A wavelength outside the table (the reference's
 * std::out_of_range), crosstalk >= 1 or a capacity below the number of
 * pulses is an MLGPU_ERR_ARGUMENT.  Synchronises `stream`. */
typedef struct mlgpu_pulses {
    int64_t* offsets;
    double* arrival_time;
    int32_t* truth_id;
    int64_t capacity;
} mlgpu_pulses;
typedef struct mlgpu_converter {
    const double* qe_x;
    const double* qe_y;
    int32_t qe_n;
    double dark_rate;
    double crosstalk;
    double exposure_time;
    uint64_t seed;
} mlgpu_converter;
int mlgpu_photoelectric_convert(const mlgpu_pipelines* pipelines, int32_t num_sensors,
                                const mlgpu_converter* converter, const mlgpu_pulses* out,
                                int64_t* num_pulses, void* stream);

/* Single-photon extraction (pulse_extraction.cpp:9-30): the time slice
 * round((t + N(0, arrival_time_std)) / time_slice_duration) of every pulse,
 * -1 when outside [0, 100); pulse k draws from Philox stream 7 of
 * (seed, k).  slice (device) has one entry per pulse.  Reads the pulse
 * count (offsets[num_sensors]) back, so it synchronises `stream`. */
int mlgpu_extract_pulses(const mlgpu_pulses* pulses, int32_t num_sensors,
                         double time_slice_duration, double arrival_time_std,
                         uint64_t seed, int16_t* slice, void* stream);

/* The photon stream of the extracted pulses (merlict_signal_processing/
 * PhotonStream.cpp:15-58): channel by channel the time slices of the kept
 * pulses (slice >= 0) in pulse order, a NEXT_CHANNEL_MARKER (255) after
 * every channel but the last, as PhotonStream::write stores them after its
 * 16-byte header; truth (device, may be NULL) receives the kept pulses'
 * truth ids (write_simulation_truth).  *num_symbols (host) = kept pulses +
 * channels - 1 (0 without either); a capacity below it is an
 * MLGPU_ERR_ARGUMENT with the count set.  Synchronises `stream`. */
int mlgpu_photon_stream(const mlgpu_pulses* pulses, const int16_t* slice, int32_t num_sensors,
                        uint8_t* symbols, int32_t* truth, int64_t capacity, int64_t* num_symbols,
                        void* stream);

/* FACT's photon stream of an event (merlict_fact/apps/fact_propagation.cpp:
 * 34-94, electric_pipelines_to_photon_stream on the converter's pulses):
 * every pulse's slice round((t - median + trigger_time) / 0.5 ns), with
 * the median the sorted arrival time at floor(pulses / 2) of the whole
 * event, kept within [30, 130) (photon_stream.h:29-34), pixel by pixel,
 * each pixel ended by NEXT_PIXEL_MARKER (255).  raw (device) receives the
 * bytes of the event's PhotonStream (photon_stream.h:284-308); *num_raw
 * (host) = kept pulses + num_pixels; a capacity below it is an
 * MLGPU_ERR_ARGUMENT with the count set.  Synchronises `stream`. */
int mlgpu_fact_photon_stream(const mlgpu_pulses* pulses, int32_t num_pixels, double trigger_time,
                             uint8_t* raw, int64_t capacity, int64_t* num_raw, void* stream);

/* Night-sky-background injection (merlict_portal_plenoscope/
 * night_sky_background/Injector.cpp:11-107) into the pipelines `in`: the
 * exposure [start, start + exposure_time) is centred on the mode of the
 * photons' arrival times (a histogram of 2 + sqrt(m) edges from the
 * earliest to the latest photon, Injector.cpp:23-66); sensor s receives
 * the photons of a Poisson process of rate * efficiencies[s] / num_sensors
 * over the exposure, drawn from Philox stream 9 of (seed, s) (arrival times
 * first, then one wavelength each from the flux's inverse cumulative
 * distribution, SamplesFromDistribution.cpp:8-25 with 4096-step
 * integral and inverse, function/tools.cpp:37-76), truth id -100; each
 * pipeline is merged by arrival time (a photon of `in` before a background
 * photon of the same time) and every arrival time becomes t - start.
 * efficiencies: device float per sensor; flux_x/flux_y: host table
 * (Func1).  `out` has its own arrays (not those of `in`); *num_photons and
 * *exposure_start_time (host) receive the merged count and start.  A
 * capacity below the count is an MLGPU_ERR_ARGUMENT with the count set.
 * Synchronises `stream`. */
typedef struct mlgpu_nsb {
    const double* flux_x;
    const double* flux_y;
    int32_t flux_n;
    double rate;
    double exposure_time;
    uint64_t seed;
} mlgpu_nsb;
int mlgpu_inject_nsb(const mlgpu_pipelines* in, int32_t num_sensors, const float* efficiencies,
                     const mlgpu_nsb* nsb, const mlgpu_pipelines* out, int64_t* num_photons,
                     double* exposure_start_time, void* stream);

/* Rate [1/s] of the night-sky light over the whole sensor, as
 * night_sky_background::Light computes it (Light.cpp:9-28): the integral of
 * flux vs wavelength (4096 steps) x pi (1.05 aperture_radius)^2 x the solid
 * angle of the cone of 1.2 max_fov_radius.  Host only. */
int mlgpu_nsb_light_rate(const double* flux_x, const double* flux_y, int32_t flux_n,
                         double max_fov_radius, double max_aperture_radius, double* rate);

/* --------------------------------------------- light-field calibration
 * One calibration photon's result, laid out as plenoscope::
 * CalibrationPhotonResult (merlict_portal_plenoscope/calibration/
 * CalibrationPhotonResult.h:9-19, 28 bytes): reached_sensor 0/1, the lixel,
 * the aperture point and the incident direction's x, y, and the time of
 * flight minus the travel to the aperture plane, all as float. */
typedef struct mlgpu_calibration_photon {
    uint32_t reached_sensor;
    uint32_t lixel_id;
    float x_pos_on_principal_aperture;
    float y_pos_on_principal_aperture;
    float x_tilt_vs_optical_axis;
    float y_tilt_vs_optical_axis;
    float time_of_flight;
} mlgpu_calibration_photon;

/* The results of n calibration photons (Calibrator.cpp:48-102, one_photon):
 * photon photon_id_offset + i was drawn by `source` (seed); its aperture
 * point is the posed support (rot (x, y, 0) + pos), its incident direction
 * the reverse
 * of its travel direction, and a photon whose result has no error flag and
 * a sensor_id in [0, num_lixel) reached that lixel, with time of flight
 * results.time_of_flight - start_distance / c.  out: n device records.
 * Asynchronous. */
int mlgpu_calibration_photons(const mlgpu_source* source, uint64_t seed, uint64_t photon_id_offset,
                              const mlgpu_results* results, int64_t n, int32_t num_lixel,
                              mlgpu_calibration_photon* out, void* stream);

/* Per-lixel statistics as plenoscope::calibration::LixelStatistic
 * (calibration/LixelStatistics.h:16-24, 48 bytes; the records of
 * lixel_statistics.bin). */
typedef struct mlgpu_lixel_statistic {
    float efficiency, efficiency_std;
    float cx_mean, cx_std;
    float cy_mean, cy_std;
    float x_mean, x_std;
    float y_mean, y_std;
    float time_delay_mean, time_delay_std;
} mlgpu_lixel_statistic;

/* LixelStatisticsFiller (calibration/LixelStatisticsFiller.cpp:8-62) on the
 * GPU: per lixel a count and the online statistics (OnlineStatistics.cpp:
 * 15-24: n, sum, Welford mean and m2 in double) of cx, cy, x, y and the time
 * delay.  fill_block adds the reached photons of a block of n device
 * records in record order (lixel by lixel, so each lixel's sums are those
 * of the reference bit for bit); statistics writes num_lixel records
 * (device): efficiency count / (num_blocks photons_per_block / num_lixel),
 * its Poisson error, each mean (sum / n) and standard deviation
 * (sqrt(m2 / n), NaN below 2 samples).  A record with reached_sensor and a
 * lixel_id >= num_lixel is an MLGPU_ERR_ARGUMENT (nothing is added). */
typedef struct mlgpu_lixel_filler mlgpu_lixel_filler;
int mlgpu_lixel_filler_create(int64_t num_lixel, int64_t num_blocks, int64_t photons_per_block,
                              mlgpu_lixel_filler** out);
int mlgpu_lixel_filler_fill_block(mlgpu_lixel_filler* filler, const mlgpu_calibration_photon* block,
                                  int64_t n, void* stream);
int mlgpu_lixel_filler_statistics(const mlgpu_lixel_filler* filler, mlgpu_lixel_statistic* out,
                                  void* stream);
void mlgpu_lixel_filler_destroy(mlgpu_lixel_filler* filler);

/* ------------------------------------------------------------- rendering
 * Images of a scene as merlict_visual renders them: a pin-hole camera
 * (PinHoleCamera.cpp:10-77: the camera frame from position and Tait-Bryan
 * rotation, the principal point (num_cols / 2) / tan(fov / 2) along the
 * optical axis, pixel (row, col) looking at principal point + x_cam
 * (row - num_rows / 2) + y_cam (col - num_cols / 2)) and the Tracer
 * (Tracer.cpp:21-145) per pixel: at the first hit a reflection when the
 * facing reflectivity at 533 nm is >= a uniform, else through a
 * transparent boundary by Fresnel (reflection when R > a uniform, else
 * refraction into the restriction frame as the propagator does), else the
 * facing colour of the surface, shaded against the global light when
 * global_illumination is on; no hit: the sky colour.  Pixel p draws its
 * uniforms from Philox stream 10 of (seed, p).  The reference recurses
 * without a bound; here a ray that has been reflected or refracted
 * max_interaction_depth times ends black. */
typedef struct mlgpu_pinhole_camera {
    double position[3];
    double rotation[3];      /* Tait-Bryan x, y, z angles (Rot3) */
    double field_of_view;    /* rad, in (0, pi] (CameraDevice.cpp:81-92) */
    int32_t num_cols;
    int32_t num_rows;
} mlgpu_pinhole_camera;
typedef struct mlgpu_render_config {
    int32_t global_illumination;    /* 0 / 1 */
    int32_t max_interaction_depth;  /* >= 1 (Config.cpp:7: 42) */
    double incoming_direction[3];   /* of the global light (Config.cpp:21) */
    float sky_color[3];             /* SkyDome background (Config.cpp:6) */
    float pad;
} mlgpu_render_config;
/* image (device): num_rows x num_cols x 3 floats, row-major (the Colors of
 * merlict's Image, 0..255); frame_colors (host): n_frames x 6 floats, the
 * outer then the inner colour of every frame, NULL for merlict's default
 * COLOR_DARK_GRAY (SurfaceEntity.cpp:42-43); n_frames must be the
 * description's.  *num_errors (host) receives the pixels whose ray met a
 * surface function that does not cover 533 nm (the reference throws; their
 * colour is NaN).  Synchronises `stream`. */
int mlgpu_render_pinhole(const mlgpu_scene* scene, int32_t world_frame, const mlgpu_pinhole_camera* camera,
                         const mlgpu_render_config* config, const float* frame_colors, int32_t n_frames,
                         uint64_t seed, float* image, int64_t* num_errors, void* stream);

/* Number of photons with a non-zero error flag in the last n results
 * (synchronises `stream`). */
int mlgpu_count_errors(const mlgpu_results* results, int64_t n, void* stream,
                       int64_t* num_errors);

/* First hit of probe rays.  Replaces
 * merlict::rays_first_intersection_with_frame (RayAndFrame.cpp:118-124). */
typedef struct mlgpu_hits {
    int32_t* frame;      /* frame index, -1 = no hit (void)        */
    double* distance;    /* distance to the ray support; 1e4 for void */
    double* position_x;  /* hit position in the root frame          */
    double* position_y;
    double* position_z;
} mlgpu_hits;

int mlgpu_first_hit(const mlgpu_scene* scene, int32_t root_frame,
                    const mlgpu_photons* rays, int64_t n, mlgpu_hits* hits,
                    void* stream);

#ifdef __cplusplus
}
#endif

#endif  /* MLGPU_H_ */
