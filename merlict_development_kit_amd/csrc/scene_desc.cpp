// Owned scene descriptions: validation, binary file I/O and the C ABI around
// them (include/mlgpu.h).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>

#include "scene.hpp"

namespace mlgpu {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

static bool finite3(const double* v) {
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

void validate_desc(const mlgpu_scene_desc& d) {
    if (d.n_frames < 1) throw std::invalid_argument("scene has no frames");
    if (!d.mother || !d.pos || !d.rot_mode || !d.rot || !d.kind || !d.params || !d.surface)
        throw std::invalid_argument("scene description has a null array");
    if (d.n_functions < 0 || (d.n_functions > 0 && (!d.func_start || !d.func_x || !d.func_y)))
        throw std::invalid_argument("bad function pool");
    if (d.mother[0] != -1) throw std::invalid_argument("frame 0 must be the root (mother -1)");
    for (int32_t i = 0; i < d.n_frames; i++) {
        if (i > 0 && (d.mother[i] < 0 || d.mother[i] >= i))
            throw std::invalid_argument("frame " + std::to_string(i) + ": mother must precede child");
        if (d.kind[i] < 0 || d.kind[i] >= MLGPU_NUM_KINDS)
            throw std::invalid_argument("frame " + std::to_string(i) + ": unknown kind");
        if (d.rot_mode[i] != MLGPU_ROT_XYZ && d.rot_mode[i] != MLGPU_ROT_AXIS &&
            d.rot_mode[i] != MLGPU_ROT_AXIS_STORED)
            throw std::invalid_argument("frame " + std::to_string(i) + ": unknown rotation mode");
        if (!finite3(d.pos + 3 * i)) throw std::invalid_argument("frame " + std::to_string(i) + ": position not finite");
        for (int s = 0; s < MLGPU_NUM_SURFACE_SLOTS; s++) {
            const int32_t f = d.surface[MLGPU_NUM_SURFACE_SLOTS * i + s];
            if (f < -1 || f >= d.n_functions)
                throw std::invalid_argument("frame " + std::to_string(i) + ": function index out of range");
        }
        if (d.allowed_frame) {
            const int32_t a = d.allowed_frame[i];
            if (a < -1 || a >= d.n_frames)
                throw std::invalid_argument("frame " + std::to_string(i) + ": allowed frame out of range");
        }
        const double* p = d.params + MLGPU_NUM_PARAMS * i;
        switch (d.kind[i]) {
        case MLGPU_KIND_SPHERE_CAP_HEX:
        case MLGPU_KIND_SPHERE_CAP_CYL:
            if (!(p[0] > 0.0) || !(p[1] > 0.0))
                throw std::invalid_argument("frame " + std::to_string(i) + ": cap radii must be > 0");
            break;
        case MLGPU_KIND_SPHERE_CAP_RECT:
            if (!(p[1] > 0.0) || !(p[2] > 0.0))
                throw std::invalid_argument("frame " + std::to_string(i) + ": widths must be > 0");
            break;
        case MLGPU_KIND_PLANE:
            if (!(p[0] > 0.0) || !(p[1] > 0.0))
                throw std::invalid_argument("frame " + std::to_string(i) + ": widths must be > 0");
            break;
        case MLGPU_KIND_DISC:
        case MLGPU_KIND_HEX_PLANE:
        case MLGPU_KIND_SPHERE:
            if (!(p[0] > 0.0)) throw std::invalid_argument("frame " + std::to_string(i) + ": radius must be > 0");
            break;
        case MLGPU_KIND_ANNULUS:
            if (!(p[0] > 0.0) || !(p[1] > 0.0))
                throw std::invalid_argument("frame " + std::to_string(i) + ": radii must be > 0");
            break;
        case MLGPU_KIND_PLANE_DUAL_SPHERICAL:
            if (!(p[0] > 0.0) || !(p[1] > 0.0))
                throw std::invalid_argument("frame " + std::to_string(i) + ": hight/width must be > 0");
            break;
        case MLGPU_KIND_CYLINDER:
            if (!(p[0] > 0.0) || !(p[1] > 0.0))
                throw std::invalid_argument("frame " + std::to_string(i) + ": cylinder radius/length must be > 0");
            break;
        case MLGPU_KIND_ELLIPTICAL_CAP_HEX:
            if (!(p[0] > 0.0) || !(p[1] > 0.0) || !(p[3] > 0.0) || !std::isfinite(p[2]))
                throw std::invalid_argument("frame " + std::to_string(i) +
                                            ": focal length, radii ratio and outer hex radius must be > 0");
            break;
        default: break;
        }
    }
    for (int32_t f = 0; f < d.n_functions; f++) {
        const int32_t a = d.func_start[f], b = d.func_start[f + 1];
        if (b <= a) throw std::invalid_argument("function " + std::to_string(f) + " is empty");
    }
}

}  // namespace mlgpu

namespace {
const char kMagic[8] = {'M', 'L', 'G', 'P', 'U', 'S', 'C', '1'};

template <typename T>
void put(std::ofstream& o, const std::vector<T>& v) {
    o.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}
template <typename T>
void get(std::ifstream& i, std::vector<T>* v, size_t n) {
    v->resize(n);
    i.read(reinterpret_cast<char*>(v->data()), static_cast<std::streamsize>(n * sizeof(T)));
}
}  // namespace

extern "C" {

const char* mlgpu_last_error(void) { return mlgpu::g_last_error.c_str(); }

int mlgpu_abi_version(void) { return MLGPU_ABI_VERSION; }

void mlgpu_desc_destroy(mlgpu_desc* d) { delete d; }

int mlgpu_desc_view(const mlgpu_desc* d, mlgpu_scene_desc* out) {
    if (!d || !out) { mlgpu::set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    *out = d->view();
    return MLGPU_OK;
}

int mlgpu_desc_create(const mlgpu_scene_desc* v, mlgpu_desc** out) {
    if (!v || !out) { mlgpu::set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    *out = nullptr;
    try {
        mlgpu::validate_desc(*v);
        auto* d = new mlgpu_desc();
        const size_t n = static_cast<size_t>(v->n_frames);
        d->mother.assign(v->mother, v->mother + n);
        d->pos.assign(v->pos, v->pos + 3 * n);
        d->rot_mode.assign(v->rot_mode, v->rot_mode + n);
        d->rot.assign(v->rot, v->rot + 4 * n);
        d->kind.assign(v->kind, v->kind + n);
        d->params.assign(v->params, v->params + MLGPU_NUM_PARAMS * n);
        d->surface.assign(v->surface, v->surface + MLGPU_NUM_SURFACE_SLOTS * n);
        if (v->allowed_frame) d->allowed_frame.assign(v->allowed_frame, v->allowed_frame + n);
        else d->allowed_frame.assign(n, MLGPU_NO_FRAME);
        if (v->sensor_id) d->sensor_id.assign(v->sensor_id, v->sensor_id + n);
        else d->sensor_id.assign(n, MLGPU_NO_SENSOR);
        const size_t nf = static_cast<size_t>(v->n_functions);
        if (nf > 0) {
            d->func_start.assign(v->func_start, v->func_start + nf + 1);
            const size_t np = static_cast<size_t>(v->func_start[nf]);
            d->func_x.assign(v->func_x, v->func_x + np);
            d->func_y.assign(v->func_y, v->func_y + np);
        }
        *out = d;
        return MLGPU_OK;
    } catch (const std::exception& e) {
        mlgpu::set_last_error(e.what());
        return MLGPU_ERR_SCENE;
    }
}

int mlgpu_desc_save(const mlgpu_desc* d, const char* path) {
    if (!d || !path) { mlgpu::set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    std::ofstream o(path, std::ios::binary);
    if (!o) { mlgpu::set_last_error(std::string("cannot write ") + path); return MLGPU_ERR_ARGUMENT; }
    const int32_t hdr[4] = {d->n_frames(), d->n_functions(), static_cast<int32_t>(d->func_x.size()), 0};
    o.write(kMagic, 8);
    o.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
    put(o, d->mother); put(o, d->pos); put(o, d->rot_mode); put(o, d->rot);
    put(o, d->kind); put(o, d->params); put(o, d->surface);
    put(o, d->allowed_frame); put(o, d->sensor_id);
    put(o, d->func_start); put(o, d->func_x); put(o, d->func_y);
    if (!o) { mlgpu::set_last_error("write failed"); return MLGPU_ERR_ARGUMENT; }
    return MLGPU_OK;
}

int mlgpu_desc_load(const char* path, mlgpu_desc** out) {
    if (!path || !out) { mlgpu::set_last_error("null argument"); return MLGPU_ERR_ARGUMENT; }
    std::ifstream in(path, std::ios::binary);
    if (!in) { mlgpu::set_last_error(std::string("cannot read ") + path); return MLGPU_ERR_ARGUMENT; }
    char magic[8];
    int32_t hdr[4];
    in.read(magic, 8);
    in.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    if (!in || std::memcmp(magic, kMagic, 8) != 0 || hdr[0] < 1 || hdr[1] < 0 || hdr[2] < 0) {
        mlgpu::set_last_error("not a scene description file");
        return MLGPU_ERR_SCENE;
    }
    auto* d = new mlgpu_desc();
    const size_t n = static_cast<size_t>(hdr[0]);
    get(in, &d->mother, n); get(in, &d->pos, 3 * n); get(in, &d->rot_mode, n); get(in, &d->rot, 4 * n);
    get(in, &d->kind, n); get(in, &d->params, MLGPU_NUM_PARAMS * n);
    get(in, &d->surface, MLGPU_NUM_SURFACE_SLOTS * n);
    get(in, &d->allowed_frame, n); get(in, &d->sensor_id, n);
    get(in, &d->func_start, static_cast<size_t>(hdr[1]) + 1);
    get(in, &d->func_x, static_cast<size_t>(hdr[2]));
    get(in, &d->func_y, static_cast<size_t>(hdr[2]));
    if (!in) { delete d; mlgpu::set_last_error("truncated scene description file"); return MLGPU_ERR_SCENE; }
    try {
        mlgpu::validate_desc(d->view());
    } catch (const std::exception& e) {
        delete d;
        mlgpu::set_last_error(e.what());
        return MLGPU_ERR_SCENE;
    }
    *out = d;
    return MLGPU_OK;
}

}  // extern "C"
