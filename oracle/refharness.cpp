// refharness.cpp -- golden-vector harness around the REFERENCE merlict code.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference's own sources (never copied), it builds a reference Frame tree
// from a scene description file (include/mlgpu.h, "MLGPUSC1"), plugs a
// Philox4x32-10 stream into merlict::random::Generator (merlict/random/
// Generator.h:11-24) so the reference consumes the same draws as the GPU
// kernel and the oracle, and writes what the reference computes:
//
//   propagate  per-photon results of merlict::Propagator (PhotonAndFrame.cpp)
//   tree       pre-order dump of the initialised tree (Frame.cpp:125-147)
//   firsthit   rays_first_intersection_with_frame (RayAndFrame.cpp:118-124)
//   reflector  facet poses of segmented_imaging_reflector::Factory
//   lensmaker  lens_maker::get_curvature_radius
//   lightfield frames built by light_field_sensor::Factory
//   bench      timing of the reference CPU path (chunked threads, or the
//              reference's own merlict_multi_thread driver)
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "merlict/merlict.h"
#include "merlict/scenery/primitive/primitive.h"
#include "merlict/scenery/segmented_imaging_reflector/segmented_imaging_reflector.h"
#include "merlict_multi_thread/merlict_multi_thread.h"
#include "merlict_portal_plenoscope/light_field_sensor/LightFieldSensor.h"
#include "merlict_corsika/PhotonFactory.h"
#include "merlict_corsika/eventio.h"
#include "merlict_signal_processing/signal_processing.h"
#include "merlict_portal_plenoscope/night_sky_background/Injector.h"
#include "merlict_portal_plenoscope/night_sky_background/Light.h"
#include "merlict_portal_plenoscope/calibration/LixelStatisticsFiller.h"
#include "merlict_visual/PinHoleCamera.h"
#include "merlict_visual/Tracer.h"
#include "merlict_fact/apps/photon_stream.h"
namespace sp = signal_processing;

namespace ml = merlict;

#include "refscene.hpp"

struct Results {
    std::vector<int32_t> frame, sensor;
    std::vector<uint8_t> interaction, n_int, err;
    std::vector<double> tof, x, y, cx, cy, acc;
    void resize(size_t n) {
        frame.resize(n); sensor.resize(n); interaction.resize(n); n_int.resize(n); err.resize(n);
        tof.resize(n); x.resize(n); y.resize(n); cx.resize(n); cy.resize(n); acc.resize(n);
    }
};

template <typename T>
static void wr(std::ofstream& o, const std::vector<T>& v) {
    o.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}

static void one_photon(const RefScene& rs, const Photons& p, int64_t i, uint64_t seed, uint64_t id0,
                       const ml::PropagationConfig* cfg, Results* r) {
    try {
        ml::Photon ph(ml::Vec3(p.sx[i], p.sy[i], p.sz[i]), ml::Vec3(p.dx[i], p.dy[i], p.dz[i]), p.wl[i]);
        PhiloxGenerator gen(seed, id0 + static_cast<uint64_t>(i));
        ml::PropagationEnvironment env;
        env.root_frame = &rs.scenery.root;
        env.config = cfg;
        env.prng = &gen;
        ml::Propagator(&ph, env);
        const ml::Intersection& fi = ph.final_intersection();
        const ml::Frame* obj = fi.object();
        auto it = rs.index.find(obj);
        r->frame[i] = it != rs.index.end() ? it->second : -1;
        ml::sensor::FindSensorByFrame finder(obj, &rs.sensors.by_frame);
        r->sensor[i] = finder.is_absorbed_by_known_sensor ? static_cast<int32_t>(finder.final_sensor->id) : -1;
        r->interaction[i] = static_cast<uint8_t>(ph.final_interaction());
        r->n_int[i] = static_cast<uint8_t>(ph.num_interactions());
        r->err[i] = 0;
        // VOID_SURFACE_ENTITY is constructed before SurfaceEntity::DEFAULT_*
        // in SurfaceEntity.cpp, so its function pointers are null and
        // time_of_flight() would dereference them for photons ending in void.
        r->tof[i] = obj != &ml::VOID_SURFACE_ENTITY ? ph.time_of_flight() : NAN;
        r->x[i] = fi.position_in_object_frame().x;
        r->y[i] = fi.position_in_object_frame().y;
        const ml::Vec3 inc = ph.final_intersection_incident_direction_wrt_frame();
        r->cx[i] = -1.0 * inc.x;
        r->cy[i] = -1.0 * inc.y;
        r->acc[i] = ph.accumulated_distance();
    } catch (const std::invalid_argument&) {
        r->err[i] = 1;
    } catch (const std::out_of_range&) {
        r->err[i] = 2;
    }
}

static int cmd_propagate(int argc, char** argv) {
    if (argc < 8) return 2;
    Desc d = load_desc(argv[2]);
    Photons p = load_photons(argv[3]);
    const uint64_t seed = std::strtoull(argv[4], nullptr, 10);
    const unsigned max_int = static_cast<unsigned>(std::strtoul(argv[5], nullptr, 10));
    const uint64_t id0 = std::strtoull(argv[6], nullptr, 10);
    const int threads = argc > 8 ? std::atoi(argv[8]) : 1;
    RefScene rs;
    build(d, &rs);
    ml::PropagationConfig cfg;
    cfg.max_num_interactions_per_photon = max_int;
    Results r;
    r.resize(static_cast<size_t>(p.n));
    std::vector<std::thread> pool;
    const int64_t chunk = (p.n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        pool.emplace_back([&, t]() {
            for (int64_t i = t * chunk; i < std::min<int64_t>(p.n, (t + 1) * chunk); i++)
                one_photon(rs, p, i, seed, id0, &cfg, &r);
        });
    }
    for (auto& th : pool) th.join();
    std::ofstream o(argv[7], std::ios::binary);
    o.write("MLGPURS2", 8);
    o.write(reinterpret_cast<const char*>(&p.n), 8);
    wr(o, r.frame); wr(o, r.sensor); wr(o, r.interaction); wr(o, r.n_int); wr(o, r.err);
    wr(o, r.tof); wr(o, r.x); wr(o, r.y); wr(o, r.cx); wr(o, r.cy); wr(o, r.acc);
    return o ? 0 : 1;
}

static void dump_rec(const ml::Frame* f, const RefScene& rs, std::vector<int32_t>* idx, std::vector<int32_t>* nch,
                     std::vector<double>* t2w, std::vector<double>* rad) {
    auto it = rs.index.find(f);
    idx->push_back(it != rs.index.end() ? it->second : -1);
    nch->push_back(static_cast<int32_t>(f->get_children()->size()));
    const ml::HomTra3* T = f->frame2world();
    const ml::Vec3 cx = T->rot_x(), cy = T->rot_y(), cz = T->rot_z(), tr = T->translation();
    const double row[12] = {cx.x, cy.x, cz.x, tr.x, cx.y, cy.y, cz.y, tr.y, cx.z, cy.z, cz.z, tr.z};
    for (double v : row) t2w->push_back(v);
    rad->push_back(f->get_bounding_sphere_radius());
    for (const ml::Frame* c : *f->get_children()) dump_rec(c, rs, idx, nch, t2w, rad);
}

static int cmd_tree(int argc, char** argv) {
    if (argc < 4) return 2;
    Desc d = load_desc(argv[2]);
    RefScene rs;
    build(d, &rs);
    std::vector<int32_t> idx, nch;
    std::vector<double> t2w, rad;
    dump_rec(&rs.scenery.root, rs, &idx, &nch, &t2w, &rad);
    std::ofstream o(argv[3], std::ios::binary);
    const int64_t n = static_cast<int64_t>(idx.size());
    o.write("MLGPUTR1", 8);
    o.write(reinterpret_cast<const char*>(&n), 8);
    wr(o, idx); wr(o, nch); wr(o, t2w); wr(o, rad);
    return o ? 0 : 1;
}

static int cmd_firsthit(int argc, char** argv) {
    if (argc < 5) return 2;
    Desc d = load_desc(argv[2]);
    Photons p = load_photons(argv[3]);
    RefScene rs;
    build(d, &rs);
    std::vector<int32_t> frame(p.n);
    std::vector<double> dist(p.n), px(p.n), py(p.n), pz(p.n);
    for (int64_t i = 0; i < p.n; i++) {
        ml::Ray ray(ml::Vec3(p.sx[i], p.sy[i], p.sz[i]), ml::Vec3(p.dx[i], p.dy[i], p.dz[i]));
        ml::Intersection is = ml::rays_first_intersection_with_frame(&ray, &rs.scenery.root);
        auto it = rs.index.find(is.object());
        frame[i] = is.does_intersect() && it != rs.index.end() ? it->second : -1;
        dist[i] = is.distance_to_ray_support();
        const ml::Vec3 w = is.position_in_root_frame();
        px[i] = w.x; py[i] = w.y; pz[i] = w.z;
    }
    std::ofstream o(argv[4], std::ios::binary);
    o.write("MLGPUFH1", 8);
    o.write(reinterpret_cast<const char*>(&p.n), 8);
    wr(o, frame); wr(o, dist); wr(o, px); wr(o, py); wr(o, pz);
    return o ? 0 : 1;
}

static void print_rot(const ml::Rot3& r) {
    if (r.uses_xyz_angels())
        std::printf(" 0 %.17g %.17g %.17g 0", r.rot_x(), r.rot_y(), r.rot_z());
    else
        std::printf(" 1 %.17g %.17g %.17g %.17g", r.rot_axis().x, r.rot_axis().y, r.rot_axis().z, r.rot_angle());
}

static int cmd_reflector(int argc, char** argv) {
    if (argc < 9) return 2;
    ml::segmented_imaging_reflector::Config c;
    c.focal_length = std::atof(argv[2]);
    c.DaviesCotton_over_parabolic_mixing_factor = std::atof(argv[3]);
    c.max_outer_aperture_radius = std::atof(argv[4]);
    c.min_inner_aperture_radius = std::atof(argv[5]);
    c.outer_aperture_shape_hex = std::atoi(argv[6]);
    c.facet_inner_hex_radius = std::atof(argv[7]);
    c.gap_between_facets = std::atof(argv[8]);
    ml::SurfaceEntity refl;
    refl.set_name_pos_rot("reflector", ml::VEC3_ORIGIN, ml::ROT3_UNITY);
    ml::segmented_imaging_reflector::Factory fab(c);
    fab.add_to_SurfaceEntity(&refl);
    for (const ml::Frame* f : *refl.get_children()) {
        const ml::Vec3 p = f->position_in_mother();
        std::printf("%.17g %.17g %.17g", p.x, p.y, p.z);
        print_rot(f->rotation_in_mother());
        std::printf(" %.17g\n", f->get_bounding_sphere_radius());
    }
    return 0;
}

// corsika-extract <eventio file> <out>: every photon bunch of every event
// of a CORSIKA eventio file, as n x 8 float32 after an int64 count.
static int cmd_corsika_extract(int argc, char** argv) {
    if (argc != 4) return 2;
    eventio::Run run(argv[2]);
    std::vector<float> all;
    while (run.has_still_events_left()) {
        eventio::Event ev = run.next_event();
        for (const auto& b : ev.photons) all.insert(all.end(), b.begin(), b.end());
    }
    std::ofstream o(argv[3], std::ios::binary);
    const int64_t n = static_cast<int64_t>(all.size() / 8);
    o.write(reinterpret_cast<const char*>(&n), 8);
    o.write(reinterpret_cast<const char*>(all.data()), static_cast<std::streamsize>(all.size() * 4));
    return 0;
}

// corsika <bunches> <seed> <id0> <out>: the reference's EventIoPhotonFactory
// on each bunch (int64 count + n x 8 float32), one Philox stream 5 per bunch
// id0 + b; writes int64 n, then per bunch the uint8 status (0 ok, 1 bad
// weight) and the uint32 number of photons, then int64 m and m photons of
// 7 float64 (support, direction, wavelength) + int32 truth id.
static int cmd_corsika(int argc, char** argv) {
    if (argc != 6) return 2;
    std::ifstream in(argv[2], std::ios::binary);
    int64_t n = 0;
    in.read(reinterpret_cast<char*>(&n), 8);
    std::vector<float> b(static_cast<size_t>(n) * 8);
    in.read(reinterpret_cast<char*>(b.data()), static_cast<std::streamsize>(b.size() * 4));
    const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
    const uint64_t id0 = std::strtoull(argv[4], nullptr, 10);
    std::vector<uint8_t> status(static_cast<size_t>(n), 0);
    std::vector<uint32_t> count(static_cast<size_t>(n), 0);
    std::vector<double> ph;
    std::vector<int32_t> truth;
    for (int64_t i = 0; i < n; i++) {
        std::array<float, 8> c;
        for (int k = 0; k < 8; k++) c[k] = b[static_cast<size_t>(i) * 8 + k];
        PhiloxGenerator prng(seed, id0 + static_cast<uint64_t>(i), 5u);
        try {
            ml::EventIoPhotonFactory f(c, static_cast<unsigned>(i), &prng);
            while (f.has_still_photons_to_be_made()) {
                const ml::Photon p = f.make_photon();
                const ml::Vec3 s = p.support(), d = p.direction();
                ph.insert(ph.end(), {s.x, s.y, s.z, d.x, d.y, d.z, p.wavelength});
                truth.push_back(static_cast<int32_t>(p.simulation_truth_id));
                count[static_cast<size_t>(i)]++;
            }
        } catch (const ml::EventIoPhotonFactory::BadPhotonWeight&) {
            status[static_cast<size_t>(i)] = 1;
        }
    }
    std::ofstream o(argv[5], std::ios::binary);
    o.write(reinterpret_cast<const char*>(&n), 8);
    o.write(reinterpret_cast<const char*>(status.data()), static_cast<std::streamsize>(n));
    o.write(reinterpret_cast<const char*>(count.data()), static_cast<std::streamsize>(n * 4));
    const int64_t m = static_cast<int64_t>(truth.size());
    o.write(reinterpret_cast<const char*>(&m), 8);
    o.write(reinterpret_cast<const char*>(ph.data()), static_cast<std::streamsize>(ph.size() * 8));
    o.write(reinterpret_cast<const char*>(truth.data()), static_cast<std::streamsize>(m * 4));
    return 0;
}

// signal <in> <out>: the reference's signal chain per sensor.  in: int64 n,
// int32 num_sensors, n int32 sensor ids, n float64 arrival times, n float64
// wavelengths, n int32 truth ids, int32 m, m float64 qe x, m float64 qe y,
// float64 dark rate, crosstalk, exposure time, uint64 seed, float64 slice
// duration, arrival time std.  Pipelines are sorted by the reference
// (sort_photon_pipelines_arrival_time); sensor s converts with a Philox
// stream 6 and extracts with a Philox stream 7 of (seed, s).  out: per
// stage the int64 offsets (num_sensors + 1) and the entries: pipelines
// (time, wavelength, truth), pulses (time, truth), extracted (slice as
// int16, truth).
static int cmd_signal(int argc, char** argv) {
    if (argc != 4) return 2;
    std::ifstream in(argv[2], std::ios::binary);
    int64_t n = 0;
    int32_t ns = 0, m = 0;
    in.read(reinterpret_cast<char*>(&n), 8);
    in.read(reinterpret_cast<char*>(&ns), 4);
    std::vector<int32_t> sensor(n), truth(n);
    std::vector<double> t(n), lam(n);
    in.read(reinterpret_cast<char*>(sensor.data()), n * 4);
    in.read(reinterpret_cast<char*>(t.data()), n * 8);
    in.read(reinterpret_cast<char*>(lam.data()), n * 8);
    in.read(reinterpret_cast<char*>(truth.data()), n * 4);
    in.read(reinterpret_cast<char*>(&m), 4);
    std::vector<double> qx(m), qy(m);
    in.read(reinterpret_cast<char*>(qx.data()), m * 8);
    in.read(reinterpret_cast<char*>(qy.data()), m * 8);
    double dark = 0, xt = 0, expo = 0, dur = 0, sd = 0;
    uint64_t seed = 0;
    in.read(reinterpret_cast<char*>(&dark), 8);
    in.read(reinterpret_cast<char*>(&xt), 8);
    in.read(reinterpret_cast<char*>(&expo), 8);
    in.read(reinterpret_cast<char*>(&seed), 8);
    in.read(reinterpret_cast<char*>(&dur), 8);
    in.read(reinterpret_cast<char*>(&sd), 8);
    std::vector<std::vector<double>> table;
    for (int32_t i = 0; i < m; i++) table.push_back({qx[i], qy[i]});
    const ml::function::Func1 qe(table);
    sp::PhotoElectricConverter::Config cfg;
    cfg.quantum_efficiency_vs_wavelength = &qe;
    cfg.dark_rate = dark;
    cfg.probability_for_second_puls = xt;
    sp::PhotoElectricConverter::Converter conv(&cfg);
    std::vector<std::vector<sp::PipelinePhoton>> pipes(static_cast<size_t>(ns));
    for (int64_t i = 0; i < n; i++)
        if (sensor[i] >= 0 && sensor[i] < ns) pipes[sensor[i]].emplace_back(t[i], lam[i], truth[i]);
    std::vector<int64_t> po{0}, eo{0}, xo{0};
    std::vector<double> pt, pl, et;
    std::vector<int32_t> ptr, etr, xtr;
    std::vector<int16_t> xs;
    // The reference's converter and extraction, called on one photon (seed,
    // j, stream 6; exposure 0 adds no accidental), one sensor's empty
    // pipeline (seed, s, stream 8: its accidental pulses) and one pulse
    // (seed, k, stream 7) at a time.
    int64_t j = 0, k = 0;
    std::vector<std::vector<sp::ExtractedPulse>> channels(static_cast<size_t>(ns));
    for (int32_t s = 0; s < ns; s++) {
        sp::sort_photon_pipelines_arrival_time(&pipes[s]);
        for (const auto& ph : pipes[s]) { pt.push_back(ph.arrival_time); pl.push_back(ph.wavelength); ptr.push_back(ph.simulation_truth_id); }
        po.push_back(static_cast<int64_t>(pt.size()));
        std::vector<sp::ElectricPulse> pulses;
        for (const auto& ph : pipes[s]) {
            PhiloxGenerator g6(seed, static_cast<uint64_t>(j++), 6u);
            const std::vector<sp::PipelinePhoton> one{ph};
            for (const auto& pu : conv.get_pulse_pipeline_for_photon_pipeline(one, 0.0, &g6)) pulses.push_back(pu);
        }
        PhiloxGenerator g8(seed, static_cast<uint64_t>(s), 8u);
        for (const auto& pu : conv.get_pulse_pipeline_for_photon_pipeline({}, expo, &g8)) pulses.push_back(pu);
        for (const auto& pu : pulses) { et.push_back(pu.arrival_time); etr.push_back(pu.simulation_truth_id); }
        eo.push_back(static_cast<int64_t>(et.size()));
        for (const auto& pu : pulses) {
            PhiloxGenerator g7(seed, static_cast<uint64_t>(k++), 7u);
            const std::vector<sp::ElectricPulse> one{pu};
            for (const auto& x : sp::extract_pulses(one, dur, sd, &g7)) {
                xs.push_back(static_cast<int16_t>(x.arrival_time_slice));
                xtr.push_back(x.simulation_truth_id);
                channels[s].push_back(x);
            }
        }
        xo.push_back(static_cast<int64_t>(xs.size()));
    }
    std::ofstream o(argv[3], std::ios::binary);
    auto w = [&](const void* p, size_t b) { o.write(static_cast<const char*>(p), static_cast<std::streamsize>(b)); };
    w(po.data(), po.size() * 8); w(pt.data(), pt.size() * 8); w(pl.data(), pl.size() * 8); w(ptr.data(), ptr.size() * 4);
    w(eo.data(), eo.size() * 8); w(et.data(), et.size() * 8); w(etr.data(), etr.size() * 4);
    w(xo.data(), xo.size() * 8); w(xs.data(), xs.size() * 2); w(xtr.data(), xtr.size() * 4);
    // the photon stream and its truth as the reference writes them
    const std::string phs = std::string(argv[3]) + ".phs", tru = std::string(argv[3]) + ".truth";
    sp::PhotonStream::write(channels, static_cast<float>(dur), phs);
    sp::PhotonStream::write_simulation_truth(channels, tru);
    for (const std::string& path : {phs, tru}) {
        std::ifstream f(path, std::ios::binary);
        const std::vector<char> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        const int64_t nb = static_cast<int64_t>(bytes.size());
        w(&nb, 8);
        w(bytes.data(), bytes.size());
    }
    return 0;
}

// The Philox stream of sensor s (stream 9 of (seed, s)) for each sensor
// in turn, as inject_nsb_into_photon_pipeline consumes them: the
// exponential gaps of sensor s until their running sum reaches the
// exposure, then one uniform per background photon for its wavelength.
class NsbStreams : public ml::random::Generator {
    uint64_t seed_nsb_;
    std::vector<double> rate_;
    double exposure_;
    size_t s_ = 0;
    bool gaps_ = true;
    double sum_ = 0.0;
    int64_t draws_ = 0, left_ = 0;
    PhiloxGenerator g_;

    void begin() {
        g_ = PhiloxGenerator(seed_nsb_, static_cast<uint64_t>(s_), 9u);
        gaps_ = true;
        sum_ = 0.0;
        draws_ = 0;
        left_ = 0;
    }

 public:
    NsbStreams(uint64_t seed, std::vector<double> rate, double exposure)
        : seed_nsb_(seed), rate_(std::move(rate)), exposure_(exposure), g_(seed, 0, 9u) { begin(); }
    double uniform() override {
        while (!gaps_ && left_ == 0) { s_++; begin(); }
        const double u = g_.uniform();
        if (gaps_) {
            const double d = -log(u) / rate_.at(s_);
            sum_ = draws_ == 0 ? d : sum_ + d;
            draws_++;
            if (!(sum_ < exposure_)) { gaps_ = false; left_ = draws_ - 1; }
        } else {
            left_--;
        }
        return u;
    }
    uint64_t create_seed() override { return 0; }
    double normal(const double, const double) override { throw std::logic_error("no normal draws here"); }
};

// nsb <in> <out> <focal> <aperture_radius> <fov_deg> <pixel_fov_deg> <paxel>
static int cmd_nsb(int argc, char** argv) {
    if (argc != 9) return 2;
    plenoscope::light_field_sensor::Config cfg;
    static ml::function::Func1 refr({{200e-9, 1.49}, {1200e-9, 1.49}});
    static ml::function::Func1 refl({{200e-9, 0.0}, {1200e-9, 0.0}});
    cfg.expected_imaging_system_focal_length = std::atof(argv[4]);
    cfg.expected_imaging_system_max_aperture_radius = std::atof(argv[5]);
    cfg.max_FoV_diameter = ml::deg2rad(std::atof(argv[6]));
    cfg.pixel_FoV_hex_flat2flat = ml::deg2rad(std::atof(argv[7]));
    cfg.num_paxel_on_pixel_diagonal = static_cast<unsigned>(std::atoi(argv[8]));
    cfg.lens_refraction = &refr;
    cfg.bin_reflection = &refl;
    plenoscope::light_field_sensor::Geometry geo(cfg);
    std::ifstream in(argv[2], std::ios::binary);
    int32_t ns = 0, fn = 0;
    in.read(reinterpret_cast<char*>(&ns), 4);
    if (static_cast<unsigned>(ns) != geo.num_lixel()) {
        std::fprintf(stderr, "need %u sensors\n", geo.num_lixel());
        return 1;
    }
    std::vector<int64_t> off(ns + 1);
    in.read(reinterpret_cast<char*>(off.data()), (ns + 1) * 8);
    const int64_t m = off[ns];
    std::vector<double> t(m), lam(m);
    std::vector<int32_t> truth(m);
    std::vector<float> eff(ns);
    in.read(reinterpret_cast<char*>(t.data()), m * 8);
    in.read(reinterpret_cast<char*>(lam.data()), m * 8);
    in.read(reinterpret_cast<char*>(truth.data()), m * 4);
    in.read(reinterpret_cast<char*>(eff.data()), ns * 4);
    in.read(reinterpret_cast<char*>(&fn), 4);
    std::vector<double> fx(fn), fy(fn);
    in.read(reinterpret_cast<char*>(fx.data()), fn * 8);
    in.read(reinterpret_cast<char*>(fy.data()), fn * 8);
    double expo = 0;
    uint64_t seed = 0;
    in.read(reinterpret_cast<char*>(&expo), 8);
    in.read(reinterpret_cast<char*>(&seed), 8);
    std::vector<std::vector<double>> table;
    for (int32_t i = 0; i < fn; i++) table.push_back({fx[i], fy[i]});
    const ml::function::Func1 flux(table);
    const plenoscope::night_sky_background::Light light(&geo, &flux);
    std::vector<std::vector<sp::PipelinePhoton>> pipes(static_cast<size_t>(ns));
    for (int32_t s = 0; s < ns; s++)
        for (int64_t j = off[s]; j < off[s + 1]; j++) pipes[s].emplace_back(t[j], lam[j], truth[j]);
    std::vector<double> rate(static_cast<size_t>(ns));
    for (int32_t s = 0; s < ns; s++) rate[s] = light.rate * eff.at(s) / light.sensor_geometry->num_lixel();
    NsbStreams gen(seed, rate, expo);
    double start = 0.0;
    plenoscope::night_sky_background::inject_nsb_into_photon_pipeline(&pipes, expo, &eff, &light, &start, &gen);
    std::vector<int64_t> po{0};
    std::vector<double> pt, pl;
    std::vector<int32_t> ptr;
    for (const auto& pipe : pipes) {
        for (const auto& ph : pipe) { pt.push_back(ph.arrival_time); pl.push_back(ph.wavelength); ptr.push_back(ph.simulation_truth_id); }
        po.push_back(static_cast<int64_t>(pt.size()));
    }
    std::ofstream o(argv[3], std::ios::binary);
    auto w = [&](const void* p, size_t b) { o.write(static_cast<const char*>(p), static_cast<std::streamsize>(b)); };
    const double head[4] = {light.rate, start, geo.max_FoV_radius(), geo.expected_imaging_system_max_aperture_radius()};
    w(head, sizeof(head));
    w(po.data(), po.size() * 8); w(pt.data(), pt.size() * 8); w(pl.data(), pl.size() * 8); w(ptr.data(), ptr.size() * 4);
    return 0;
}

// lixelstats <in> <out.bin>: LixelStatisticsFiller over blocks of
// CalibrationPhotonResult records, written by calibration::write
static int cmd_lixelstats(int argc, char** argv) {
    if (argc != 4) return 2;
    std::ifstream in(argv[2], std::ios::binary);
    int64_t num_lixel = 0, num_blocks = 0, per_block = 0, n = 0;
    in.read(reinterpret_cast<char*>(&num_lixel), 8);
    in.read(reinterpret_cast<char*>(&num_blocks), 8);
    in.read(reinterpret_cast<char*>(&per_block), 8);
    in.read(reinterpret_cast<char*>(&n), 8);
    static_assert(sizeof(plenoscope::CalibrationPhotonResult) == 28, "record size");
    std::vector<plenoscope::CalibrationPhotonResult> rec(static_cast<size_t>(n));
    in.read(reinterpret_cast<char*>(rec.data()), n * 28);
    plenoscope::calibration::LixelStatisticsFiller filler(num_lixel, num_blocks, per_block);
    for (int64_t b = 0; b * per_block < n; b++) {
        const int64_t lo = b * per_block, hi = std::min(n, lo + per_block);
        const std::vector<plenoscope::CalibrationPhotonResult> block(rec.begin() + lo, rec.begin() + hi);
        filler.fill_in_block(block);
    }
    plenoscope::calibration::write(filler.get_lixel_statistics(), argv[3]);
    return 0;
}

// The Tracer's generator: an Mt19937 whose uniforms come from Philox stream
// `stream` of (seed, id).
class PhiloxMt : public ml::random::Mt19937 {
    PhiloxGenerator g_;

 public:
    PhiloxMt(uint64_t seed, uint64_t id, uint32_t stream) : ml::random::Mt19937(0), g_(seed, id, stream) {}
    double uniform() override { return g_.uniform(); }
};

// render <scene> <in> <out>: the reference's PinHoleCamera rays and Tracer
// colours per pixel, pixel p on Philox stream 10 of (seed, p)
static int cmd_render(int argc, char** argv) {
    if (argc != 5) return 2;
    Desc d = load_desc(argv[2]);
    RefScene rs;
    build(d, &rs);
    std::ifstream in(argv[3], std::ios::binary);
    double pos[3], rot[3], fov, light[3];
    int32_t cols = 0, rows = 0, gi = 0, depth = 0, nf = 0;
    float sky[3];
    uint64_t seed = 0;
    in.read(reinterpret_cast<char*>(pos), 24);
    in.read(reinterpret_cast<char*>(rot), 24);
    in.read(reinterpret_cast<char*>(&fov), 8);
    in.read(reinterpret_cast<char*>(&cols), 4);
    in.read(reinterpret_cast<char*>(&rows), 4);
    in.read(reinterpret_cast<char*>(&gi), 4);
    in.read(reinterpret_cast<char*>(&depth), 4);
    in.read(reinterpret_cast<char*>(light), 24);
    in.read(reinterpret_cast<char*>(sky), 12);
    in.read(reinterpret_cast<char*>(&nf), 4);
    std::vector<float> col(static_cast<size_t>(nf) * 6);
    in.read(reinterpret_cast<char*>(col.data()), nf * 24);
    in.read(reinterpret_cast<char*>(&seed), 8);
    if (nf != d.n) { std::fprintf(stderr, "need %d frame colours\n", d.n); return 1; }
    std::vector<ml::Color> palette;
    palette.reserve(static_cast<size_t>(nf) * 2);
    for (int32_t i = 0; i < nf; i++) {
        auto* s = dynamic_cast<ml::SurfaceEntity*>(rs.frames[i]);
        palette.emplace_back(col[6 * i], col[6 * i + 1], col[6 * i + 2]);
        palette.emplace_back(col[6 * i + 3], col[6 * i + 4], col[6 * i + 5]);
        if (s) {
            s->outer_color = &palette[2 * i];
            s->inner_color = &palette[2 * i + 1];
        }
    }
    ml::visual::Config cfg;
    cfg.global_illumination.on = gi != 0;
    cfg.global_illumination.incoming_direction = ml::Vec3(light[0], light[1], light[2]);
    cfg.sky_dome = ml::visual::SkyDome(ml::Color(sky[0], sky[1], sky[2]));
    ml::visual::PinHoleCamera cam("cam", static_cast<unsigned>(cols), static_cast<unsigned>(rows));
    cam.set_pos_rot_fov(ml::Vec3(pos[0], pos[1], pos[2]), ml::Rot3(rot[0], rot[1], rot[2]), fov);
    std::vector<float> img(static_cast<size_t>(cols) * rows * 3);
    const auto t0 = std::chrono::steady_clock::now();
    for (int64_t p = 0; p < static_cast<int64_t>(cols) * rows; p++) {
        const int row = static_cast<int>(p / cols), c = static_cast<int>(p % cols);
        ml::visual::CameraRay ray = cam.get_ray_for_pixel_in_row_and_col(row, c);
        PhiloxMt gen(seed, static_cast<uint64_t>(p), 10u);
        ml::visual::Tracer tracer(&ray, &rs.scenery.root, &cfg, &gen);
        img[3 * p] = tracer.color.r;
        img[3 * p + 1] = tracer.color.g;
        img[3 * p + 2] = tracer.color.b;
    }
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("{\"seconds\": %.6f}\n", sec);
    std::ofstream o(argv[4], std::ios::binary);
    o.write(reinterpret_cast<const char*>(img.data()), static_cast<std::streamsize>(img.size() * 4));
    return 0;
}

// factphs <file>: the reference's reader of FACT simulation events; one
// line per event: valid, pass version, type, run, event, reuse, zd, az,
// number_of_photons, raw bytes
static int cmd_factphs(int argc, char** argv) {
    if (argc != 3) return 2;
    std::ifstream in(argv[2], std::ios::binary);
    in.seekg(0, std::ios::end);
    const std::streamoff size = in.tellg();
    in.seekg(0, std::ios::beg);
    while (in.tellg() < size) {
        photon_stream::SimulationEvent e = photon_stream::read_SimulationEvent_from_file(in);
        std::printf("%d %u %u %u %u %u %.9g %.9g %u", e.descriptor.is_valid() ? 1 : 0, e.descriptor.pass_version,
                    e.descriptor.event_type, e.id.run, e.id.event, e.id.reuse, e.pointing.zd, e.pointing.az,
                    e.photon_stream.number_of_photons());
        for (uint8_t b : e.photon_stream.raw) std::printf(" %u", b);
        std::printf("\n");
    }
    return 0;
}

static int cmd_lensmaker(int argc, char** argv) {
    if (argc < 5) return 2;
    ml::lens_maker::Config c;
    c.focal_length = std::atof(argv[2]);
    c.aperture_radius = std::atof(argv[3]);
    c.refractive_index = std::atof(argv[4]);
    std::printf("%.17g\n", ml::lens_maker::get_curvature_radius(c));
    return 0;
}

static void lf_rec(const ml::Frame* f, int depth) {
    const ml::Vec3 p = f->position_in_mother();
    int kind = 0;
    if (dynamic_cast<const ml::SphereCapWithHexagonalBound*>(f)) kind = 1;
    else if (dynamic_cast<const ml::Plane*>(f)) kind = 4;
    else if (dynamic_cast<const ml::Disc*>(f)) kind = 5;
    else if (dynamic_cast<const ml::HexPlane*>(f)) kind = 6;
    else if (dynamic_cast<const ml::Annulus*>(f)) kind = 7;
    else if (dynamic_cast<const ml::PlaneDualSphericalBound*>(f)) kind = 8;
    else if (dynamic_cast<const ml::Cylinder*>(f)) kind = 9;
    std::printf("%d %d %zu %.17g %.17g %.17g", depth, kind, f->get_children()->size(), p.x, p.y, p.z);
    print_rot(f->rotation_in_mother());
    std::printf(" %.17g\n", f->get_bounding_sphere_radius());
    for (const ml::Frame* c : *f->get_children()) lf_rec(c, depth + 1);
}

// light_field_sensor::Factory on the plenoscope example configuration
static int cmd_lightfield(int argc, char** argv) {
    if (argc < 9) return 2;
    plenoscope::light_field_sensor::Config cfg;
    static ml::function::Func1 refr({{200e-9, std::atof(argv[7])}, {1200e-9, std::atof(argv[7])}});
    static ml::function::Func1 refl({{200e-9, std::atof(argv[8])}, {1200e-9, std::atof(argv[8])}});
    cfg.expected_imaging_system_focal_length = std::atof(argv[2]);
    cfg.expected_imaging_system_max_aperture_radius = std::atof(argv[3]);
    cfg.max_FoV_diameter = ml::deg2rad(std::atof(argv[4]));
    cfg.pixel_FoV_hex_flat2flat = ml::deg2rad(std::atof(argv[5]));
    cfg.num_paxel_on_pixel_diagonal = static_cast<unsigned>(std::atoi(argv[6]));
    cfg.housing_overhead = argc > 9 ? std::atof(argv[9]) : 1.1;
    cfg.lens_refraction = &refr;
    cfg.bin_reflection = &refl;
    plenoscope::light_field_sensor::Geometry geo(cfg);
    ml::Scenery scenery;
    ml::Frame* lfs = scenery.root.add<ml::Frame>();
    lfs->set_name_pos_rot("light_field_sensor", ml::VEC3_ORIGIN, ml::ROT3_UNITY);
    plenoscope::light_field_sensor::Factory fab(&geo);
    fab.add_light_field_sensor_to_frame_in_scenery(lfs, &scenery);
    std::printf("# lens_R %.17g lens_f %.17g num_lixel %u\n", geo.pixel_lens_curvature_radius(),
                geo.pixel_lens_focal_length(), geo.num_lixel());
    lf_rec(lfs, 0);
    return 0;
}

// bench <scene> <photons> <threads> <mode: chunked|multi_thread> <seed> [results]
// (chunked: one Philox stream per photon, ids 0 .. n-1, max 5 interactions;
// with a results path the per-photon results are written as by propagate)
static int cmd_bench(int argc, char** argv) {
    if (argc < 7) return 2;
    Desc d = load_desc(argv[2]);
    Photons p = load_photons(argv[3]);
    const int threads = std::max(1, std::atoi(argv[4]));
    const std::string mode = argv[5];
    const uint64_t seed = std::strtoull(argv[6], nullptr, 10);
    RefScene rs;
    build(d, &rs);
    ml::PropagationConfig cfg;
    Results r;
    r.resize(static_cast<size_t>(p.n));
    const auto t0 = std::chrono::steady_clock::now();
    if (mode == "multi_thread") {
        std::vector<ml::Photon> photons;
        photons.reserve(static_cast<size_t>(p.n));
        for (int64_t i = 0; i < p.n; i++)
            photons.emplace_back(ml::Vec3(p.sx[i], p.sy[i], p.sz[i]), ml::Vec3(p.dx[i], p.dy[i], p.dz[i]), p.wl[i]);
        ml::random::Mt19937 prng(seed);
        const auto t1 = std::chrono::steady_clock::now();
        ml::propagate_photons_in_frame_with_config_multi_thread(&photons, &rs.scenery.root, &cfg, &prng);
        const auto t2 = std::chrono::steady_clock::now();
        const double s = std::chrono::duration<double>(t2 - t1).count();
        std::printf("{\"photons\": %lld, \"seconds\": %.6f, \"threads\": %u, \"mode\": \"multi_thread\"}\n",
                    static_cast<long long>(p.n), s, std::thread::hardware_concurrency());
        return 0;
    }
    std::vector<std::thread> pool;
    const int64_t chunk = (p.n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        pool.emplace_back([&, t]() {
            for (int64_t i = t * chunk; i < std::min<int64_t>(p.n, (t + 1) * chunk); i++)
                one_photon(rs, p, i, seed, 0, &cfg, &r);
        });
    }
    for (auto& th : pool) th.join();
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int64_t detected = 0;
    for (int64_t i = 0; i < p.n; i++) detected += r.sensor[i] >= 0;
    std::printf("{\"photons\": %lld, \"seconds\": %.6f, \"threads\": %d, \"mode\": \"chunked\", \"detected\": %lld}\n",
                static_cast<long long>(p.n), s, threads, static_cast<long long>(detected));
    if (argc > 7) {
        std::ofstream o(argv[7], std::ios::binary);
        o.write("MLGPURS2", 8);
        o.write(reinterpret_cast<const char*>(&p.n), 8);
        wr(o, r.frame); wr(o, r.sensor); wr(o, r.interaction); wr(o, r.n_int); wr(o, r.err);
        wr(o, r.tof); wr(o, r.x); wr(o, r.y); wr(o, r.cx); wr(o, r.cy); wr(o, r.acc);
        if (!o) return 1;
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: merlict_golden propagate|tree|firsthit|reflector|lensmaker|lightfield|bench|corsika-extract|corsika|signal ...\n");
        return 2;
    }
    try {
        const std::string c = argv[1];
        int rc = 2;
        if (c == "propagate") rc = cmd_propagate(argc, argv);
        else if (c == "tree") rc = cmd_tree(argc, argv);
        else if (c == "firsthit") rc = cmd_firsthit(argc, argv);
        else if (c == "reflector") rc = cmd_reflector(argc, argv);
        else if (c == "lensmaker") rc = cmd_lensmaker(argc, argv);
        else if (c == "lightfield") rc = cmd_lightfield(argc, argv);
        else if (c == "bench") rc = cmd_bench(argc, argv);
        else if (c == "corsika-extract") rc = cmd_corsika_extract(argc, argv);
        else if (c == "corsika") rc = cmd_corsika(argc, argv);
        else if (c == "signal") rc = cmd_signal(argc, argv);
        else if (c == "nsb") rc = cmd_nsb(argc, argv);
        else if (c == "lixelstats") rc = cmd_lixelstats(argc, argv);
        else if (c == "render") rc = cmd_render(argc, argv);
        else if (c == "factphs") rc = cmd_factphs(argc, argv);
        if (rc == 2) std::fprintf(stderr, "bad arguments for %s\n", c.c_str());
        return rc;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
