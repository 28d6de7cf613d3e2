// refscene.hpp -- reference scene, photon and Philox plumbing shared by the
// golden-vector harness (refharness.cpp) and the drop-in check of the GPU
// binding (facade_check.cpp).
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference's own headers: builds a reference Frame tree from a scene
// description file (include/mlgpu.h, "MLGPUSC1"), reads photon files
// ("MLGPUPH1"), and plugs a Philox4x32-10 stream into
// merlict::random::Generator (merlict/random/Generator.h:11-24) so the
// reference consumes the same draws as the GPU kernel and the oracle.
#ifndef MLGPU_ORACLE_REFSCENE_HPP_
#define MLGPU_ORACLE_REFSCENE_HPP_

#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "merlict/merlict.h"
#include "merlict/scenery/primitive/primitive.h"

namespace ml = merlict;

// ------------------------------------------------------------ Philox stream
static void philox(uint32_t c[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
        const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = static_cast<uint32_t>(p1); c2 = n2; c3 = static_cast<uint32_t>(p0);
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

class PhiloxGenerator : public ml::random::Generator {
    uint32_t key_[2];
    uint32_t ctr_[4];
    uint32_t buf_[4];
    uint32_t next_ = 0;

 public:
    PhiloxGenerator(uint64_t seed, uint64_t photon_id, uint32_t stream = 0) {
        key_[0] = static_cast<uint32_t>(seed);
        key_[1] = static_cast<uint32_t>(seed >> 32);
        ctr_[0] = 0; ctr_[1] = stream;
        ctr_[2] = static_cast<uint32_t>(photon_id);
        ctr_[3] = static_cast<uint32_t>(photon_id >> 32);
        seed_ = seed;
    }
    double uniform() override {
        if ((next_ & 3u) == 0u) {
            ctr_[0] = next_ >> 2;
            philox(ctr_, key_, buf_);
        }
        const uint32_t w = buf_[next_ & 3u];
        next_++;
        return static_cast<double>(w >> 8) * (1.0 / 16777216.0);
    }
    uint64_t create_seed() override {
        const uint64_t a = static_cast<uint64_t>(uniform() * 16777216.0);
        const uint64_t b = static_cast<uint64_t>(uniform() * 16777216.0);
        return (a << 24) | b;
    }
    double normal(const double mean, const double std_dev) override {
        const double u1 = 1.0 - uniform(), u2 = uniform();
        return mean + std_dev * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// ----------------------------------------------------------- description file
struct Desc {
    int32_t n = 0, nf = 0, np = 0;
    std::vector<int32_t> mother, rot_mode, kind, surface, allowed, sensor, func_start;
    std::vector<double> pos, rot, params, fx, fy;
};

template <typename T>
static void rd(std::ifstream& in, std::vector<T>* v, size_t n) {
    v->resize(n);
    in.read(reinterpret_cast<char*>(v->data()), static_cast<std::streamsize>(n * sizeof(T)));
}

static Desc load_desc(const char* path) {
    std::ifstream in(path, std::ios::binary);
    char magic[8];
    int32_t hdr[4];
    in.read(magic, 8);
    in.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    if (!in || std::memcmp(magic, "MLGPUSC1", 8) != 0) throw std::runtime_error("bad scene file");
    Desc d;
    d.n = hdr[0]; d.nf = hdr[1]; d.np = hdr[2];
    const size_t n = static_cast<size_t>(d.n);
    rd(in, &d.mother, n); rd(in, &d.pos, 3 * n); rd(in, &d.rot_mode, n); rd(in, &d.rot, 4 * n);
    rd(in, &d.kind, n); rd(in, &d.params, 6 * n); rd(in, &d.surface, 6 * n);
    rd(in, &d.allowed, n); rd(in, &d.sensor, n);
    rd(in, &d.func_start, static_cast<size_t>(d.nf) + 1);
    rd(in, &d.fx, static_cast<size_t>(d.np)); rd(in, &d.fy, static_cast<size_t>(d.np));
    if (!in) throw std::runtime_error("truncated scene file");
    return d;
}

struct RefScene {
    ml::Scenery scenery;
    std::vector<ml::Frame*> frames;
    std::map<const ml::Frame*, int32_t> index;
    std::vector<ml::function::Func1*> funcs;
    std::vector<ml::sensor::Sensor*> sensor_list;
    ml::sensor::Sensors sensors;
    ~RefScene() { for (auto* f : funcs) delete f; for (auto* s : sensor_list) delete s; }
};

template <typename T>
static T* add_surface(ml::Frame* mother, const Desc& d, int32_t i, RefScene* rs) {
    T* s = mother->add<T>();
    const int32_t* sf = &d.surface[6 * i];
    if (sf[0] >= 0) s->outer_reflection = rs->funcs[sf[0]];
    if (sf[1] >= 0) s->inner_reflection = rs->funcs[sf[1]];
    if (sf[2] >= 0) s->outer_refraction = rs->funcs[sf[2]];
    if (sf[3] >= 0) s->inner_refraction = rs->funcs[sf[3]];
    if (sf[4] >= 0) s->outer_absorption = rs->funcs[sf[4]];
    if (sf[5] >= 0) s->inner_absorption = rs->funcs[sf[5]];
    return s;
}

static void build(const Desc& d, RefScene* rs) {
    for (int32_t f = 0; f < d.nf; f++) {
        std::vector<std::vector<double>> xy;
        for (int32_t k = d.func_start[f]; k < d.func_start[f + 1]; k++) xy.push_back({d.fx[k], d.fy[k]});
        rs->funcs.push_back(new ml::function::Func1(xy));
    }
    std::vector<int> has_children(d.n, 0);
    for (int32_t i = 1; i < d.n; i++) has_children[d.mother[i]] = 1;
    rs->frames.resize(d.n);
    rs->scenery.root.set_name_pos_rot("root", ml::Vec3(d.pos[0], d.pos[1], d.pos[2]),
                                      d.rot_mode[0] == 0 ? ml::Rot3(d.rot[0], d.rot[1], d.rot[2])
                                                         : ml::Rot3(ml::Vec3(d.rot[0], d.rot[1], d.rot[2]), d.rot[3]));
    rs->frames[0] = &rs->scenery.root;
    for (int32_t i = 1; i < d.n; i++) {
        ml::Frame* m = rs->frames[d.mother[i]];
        const double* p = &d.params[6 * i];
        ml::Frame* f = nullptr;
        switch (d.kind[i]) {
        case 0: f = has_children[i] ? static_cast<ml::Frame*>(add_surface<ml::SurfaceEntity>(m, d, i, rs))
                                    : m->add<ml::Frame>();
                break;
        case 1: { auto* s = add_surface<ml::SphereCapWithHexagonalBound>(m, d, i, rs); f = s; break; }
        case 2: { auto* s = add_surface<ml::SphereCapWithCylinderBound>(m, d, i, rs); f = s; break; }
        case 3: { auto* s = add_surface<ml::SphereCapWithRectangularBound>(m, d, i, rs); f = s; break; }
        case 4: { auto* s = add_surface<ml::Plane>(m, d, i, rs); f = s; break; }
        case 5: { auto* s = add_surface<ml::Disc>(m, d, i, rs); f = s; break; }
        case 6: { auto* s = add_surface<ml::HexPlane>(m, d, i, rs); f = s; break; }
        case 7: { auto* s = add_surface<ml::Annulus>(m, d, i, rs); f = s; break; }
        case 8: { auto* s = add_surface<ml::PlaneDualSphericalBound>(m, d, i, rs); f = s; break; }
        case 9: { auto* s = add_surface<ml::Cylinder>(m, d, i, rs); f = s; break; }
        case 10: { auto* s = add_surface<ml::Triangle>(m, d, i, rs); f = s; break; }
        case 11: { auto* s = add_surface<ml::Sphere>(m, d, i, rs); f = s; break; }
        case 12: { auto* s = add_surface<ml::EllipticalCapWithHexagonalBound>(m, d, i, rs); f = s; break; }
        default: throw std::runtime_error("unknown kind");
        }
        const ml::Rot3 rot = d.rot_mode[i] == 0 ? ml::Rot3(d.rot[4 * i], d.rot[4 * i + 1], d.rot[4 * i + 2])
                                                : ml::Rot3(ml::Vec3(d.rot[4 * i], d.rot[4 * i + 1], d.rot[4 * i + 2]),
                                                           d.rot[4 * i + 3]);
        f->set_name_pos_rot("f" + std::to_string(i), ml::Vec3(d.pos[3 * i], d.pos[3 * i + 1], d.pos[3 * i + 2]), rot);
        switch (d.kind[i]) {
        case 1: static_cast<ml::SphereCapWithHexagonalBound*>(f)->set_curvature_radius_and_outer_hex_radius(p[0], p[1]); break;
        case 2: static_cast<ml::SphereCapWithCylinderBound*>(f)->set_curvature_radius_and_outer_radius(p[0], p[1]); break;
        case 3: static_cast<ml::SphereCapWithRectangularBound*>(f)->set_curvature_radius_and_x_y_width(p[0], p[1], p[2]); break;
        case 4: static_cast<ml::Plane*>(f)->set_x_y_width(p[0], p[1]); break;
        case 5: static_cast<ml::Disc*>(f)->set_radius(p[0]); break;
        case 6: static_cast<ml::HexPlane*>(f)->set_outer_hex_radius(p[0]); break;
        case 7: static_cast<ml::Annulus*>(f)->set_outer_inner_radius(p[0], p[1]); break;
        case 8: static_cast<ml::PlaneDualSphericalBound*>(f)->set_x_hight_and_y_width(p[0], p[1]); break;
        case 9: static_cast<ml::Cylinder*>(f)->set_radius_and_length(p[0], p[1]); break;
        case 10: static_cast<ml::Triangle*>(f)->set_corners_in_xy_plane(p[0], p[1], p[2], p[3], p[4], p[5]); break;
        case 11: static_cast<ml::Sphere*>(f)->set_radius(p[0]); break;
        case 12:
            static_cast<ml::EllipticalCapWithHexagonalBound*>(f)->set_focalLength_radiiRatio_hexAngel_hexRadius(
                p[0], p[1], p[2], p[3]);
            break;
        default: break;
        }
        rs->frames[i] = f;
    }
    for (int32_t i = 0; i < d.n; i++) {
        rs->index[rs->frames[i]] = i;
        if (d.allowed[i] >= 0) {
            auto* s = dynamic_cast<ml::SurfaceEntity*>(rs->frames[i]);
            if (!s) throw std::runtime_error("restriction on a non-surface frame");
            s->set_allowed_frames_to_propagate_to(rs->frames[d.allowed[i]]);
        }
        if (d.sensor[i] >= 0) rs->sensor_list.push_back(new ml::sensor::Sensor(d.sensor[i], rs->frames[i]));
    }
    rs->sensors.init(rs->sensor_list);
    rs->scenery.root.init_tree_based_on_mother_child_relations();
}

// ---------------------------------------------------------------- photons
struct Photons {
    int64_t n = 0;
    std::vector<double> sx, sy, sz, dx, dy, dz, wl;
};

static Photons load_photons(const char* path) {
    std::ifstream in(path, std::ios::binary);
    char magic[8];
    Photons p;
    in.read(magic, 8);
    in.read(reinterpret_cast<char*>(&p.n), 8);
    if (!in || std::memcmp(magic, "MLGPUPH1", 8) != 0) throw std::runtime_error("bad photon file");
    const size_t n = static_cast<size_t>(p.n);
    rd(in, &p.sx, n); rd(in, &p.sy, n); rd(in, &p.sz, n);
    rd(in, &p.dx, n); rd(in, &p.dy, n); rd(in, &p.dz, n); rd(in, &p.wl, n);
    if (!in) throw std::runtime_error("truncated photon file");
    return p;
}

#endif  // MLGPU_ORACLE_REFSCENE_HPP_
