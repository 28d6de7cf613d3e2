// Host-side scene description for the MI355X photon propagator.
//
// mlgpu_desc is the owned form of mlgpu_scene_desc (include/mlgpu.h): a flat
// list of frames in creation order, like the merlict Frame tree before
// Frame::init_tree_based_on_mother_child_relations (merlict/Frame.cpp:125).
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mlgpu.h"

namespace mlgpu {

struct Vec3d {
    double x, y, z;
};

struct Surface {
    // function index per mlgpu_surface_slot, -1 = merlict default table
    int32_t f[MLGPU_NUM_SURFACE_SLOTS] = {-1, -1, -1, -1, -1, -1};
    Surface inside_out() const {
        // SurfaceEntity::adopt_surface_inside_out (merlict/SurfaceEntity.cpp:89-100)
        Surface s;
        s.f[MLGPU_OUTER_REFLECTION] = f[MLGPU_INNER_REFLECTION];
        s.f[MLGPU_INNER_REFLECTION] = f[MLGPU_OUTER_REFLECTION];
        s.f[MLGPU_OUTER_REFRACTION] = f[MLGPU_INNER_REFRACTION];
        s.f[MLGPU_INNER_REFRACTION] = f[MLGPU_OUTER_REFRACTION];
        s.f[MLGPU_OUTER_ABSORPTION] = f[MLGPU_INNER_ABSORPTION];
        s.f[MLGPU_INNER_ABSORPTION] = f[MLGPU_OUTER_ABSORPTION];
        return s;
    }
};

struct Rot {
    int32_t mode = MLGPU_ROT_XYZ;
    double v[4] = {0, 0, 0, 0};
    static Rot xyz(double rx, double ry, double rz) {
        Rot r;
        r.mode = MLGPU_ROT_XYZ;
        r.v[0] = rx; r.v[1] = ry; r.v[2] = rz; r.v[3] = 0.0;
        return r;
    }
    static Rot axis(Vec3d a, double angle) {
        Rot r;
        r.mode = MLGPU_ROT_AXIS;
        r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = angle;
        return r;
    }
};

}  // namespace mlgpu

struct mlgpu_desc {
    std::vector<int32_t> mother;
    std::vector<double> pos;
    std::vector<int32_t> rot_mode;
    std::vector<double> rot;
    std::vector<int32_t> kind;
    std::vector<double> params;
    std::vector<int32_t> surface;
    std::vector<int32_t> allowed_frame;
    std::vector<int32_t> sensor_id;
    std::vector<int32_t> func_start{0};
    std::vector<double> func_x;
    std::vector<double> func_y;

    int32_t n_frames() const { return static_cast<int32_t>(mother.size()); }
    int32_t n_functions() const { return static_cast<int32_t>(func_start.size()) - 1; }

    int32_t add_frame(int32_t m, mlgpu::Vec3d p, mlgpu::Rot r, int32_t k,
                      std::initializer_list<double> prm = {},
                      const mlgpu::Surface& s = mlgpu::Surface(),
                      int32_t allowed = MLGPU_NO_FRAME,
                      int32_t sensor = MLGPU_NO_SENSOR) {
        const int32_t idx = n_frames();
        if (idx == 0 && m != -1) throw std::invalid_argument("frame 0 must be the root");
        if (idx > 0 && (m < 0 || m >= idx)) throw std::invalid_argument("mother must precede child");
        mother.push_back(m);
        pos.push_back(p.x); pos.push_back(p.y); pos.push_back(p.z);
        rot_mode.push_back(r.mode);
        for (int i = 0; i < 4; i++) rot.push_back(r.v[i]);
        kind.push_back(k);
        double pp[MLGPU_NUM_PARAMS] = {0, 0, 0, 0, 0, 0};
        int i = 0;
        for (double v : prm) {
            if (i >= MLGPU_NUM_PARAMS) throw std::invalid_argument("too many params");
            pp[i++] = v;
        }
        for (int j = 0; j < MLGPU_NUM_PARAMS; j++) params.push_back(pp[j]);
        for (int j = 0; j < MLGPU_NUM_SURFACE_SLOTS; j++) surface.push_back(s.f[j]);
        allowed_frame.push_back(allowed);
        sensor_id.push_back(sensor);
        return idx;
    }

    int32_t add_function(const std::vector<double>& x, const std::vector<double>& y) {
        if (x.size() != y.size() || x.empty()) throw std::invalid_argument("bad function table");
        for (size_t i = 0; i < x.size(); i++) { func_x.push_back(x[i]); func_y.push_back(y[i]); }
        func_start.push_back(static_cast<int32_t>(func_x.size()));
        return n_functions() - 1;
    }

    mlgpu_scene_desc view() const {
        mlgpu_scene_desc d;
        d.n_frames = n_frames();
        d.mother = mother.data();
        d.pos = pos.data();
        d.rot_mode = rot_mode.data();
        d.rot = rot.data();
        d.kind = kind.data();
        d.params = params.data();
        d.surface = surface.data();
        d.allowed_frame = allowed_frame.data();
        d.sensor_id = sensor_id.data();
        d.n_functions = n_functions();
        d.func_start = func_start.data();
        d.func_x = func_x.data();
        d.func_y = func_y.data();
        return d;
    }
};

namespace mlgpu {

void set_last_error(const std::string& msg);
// Check a description for index and parameter consistency; throws
// std::invalid_argument with the first problem found.
void validate_desc(const mlgpu_scene_desc& d);

// mlgpu_propagate and mlgpu_histogram_arrivals for a batch gathered from a
// larger one: entry i is photon ids[i] of that batch, so it draws (and is
// jittered) on the Philox stream of photon id offset + ids[i].  Null ids:
// the public entry points.
int propagate_gathered(const mlgpu_scene* s, int32_t world_frame, const mlgpu_photons* photons, int64_t n,
                       const mlgpu_config* cfg, mlgpu_results* r, const int32_t* ids, void* stream);
int histogram_gathered(const mlgpu_results* r, int64_t n, const mlgpu_histogram* h, const int32_t* ids,
                       void* stream);

}  // namespace mlgpu
