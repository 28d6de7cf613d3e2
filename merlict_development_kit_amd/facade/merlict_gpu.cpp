// merlict_gpu.cpp -- merlict's batch photon propagation through libmlgpu.so.
//
// Three steps per call:
//   1. flatten: the initialised Frame tree of world->root() in pre-order
//      (each frame's children in merlict's order) into an mlgpu scene
//      description -- pose in mother, primitive and its parameters, the six
//      surface functions, the propagation restriction;
//   2. propagate: photons to HBM, mlgpu_propagate_history, histories back;
//   3. rebuild: every bounce pushed onto its Photon as merlict's Propagator
//      pushes it (merlict/PhotonAndFrame.cpp:50-124), the ray of the last leg
//      set as Ray::set_support_and_direction leaves it.
#include "merlict_gpu.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>
#include <typeinfo>
#include <unordered_map>

#include "merlict/SurfaceEntity.h"
#include "merlict/scenery/primitive/primitive.h"

#include "mlgpu.h"

namespace merlict {
namespace gpu {
namespace {

// ------------------------------------------------------------ member access
// merlict keeps the parameters of its primitives protected or private and has
// no getter for several of them; an explicit instantiation may name any
// member ([temp.explicit]), which hands out a pointer to it.
template <typename Tag, typename Tag::type M>
struct Expose {
    friend typename Tag::type member(Tag) { return M; }
};

#define MLGPU_EXPOSE(NAME, CLASS, TYPE, MEMBER)  \
    struct NAME {                                \
        typedef TYPE CLASS::*type;               \
        friend type member(NAME);                \
    };                                           \
    template struct Expose<NAME, &CLASS::MEMBER>

MLGPU_EXPOSE(CapHexCurvature, SphereCapWithHexagonalBound, double, curvature_radius);
MLGPU_EXPOSE(CapHexRadius, SphereCapWithHexagonalBound, double, outer_hex_radius);
MLGPU_EXPOSE(CapCylCurvature, SphereCapWithCylinderBound, double, curvature_radius);
MLGPU_EXPOSE(CapCylRadius, SphereCapWithCylinderBound, double, cap_radius);
MLGPU_EXPOSE(CapRectCurvature, SphereCapWithRectangularBound, double, curvature_radius);
MLGPU_EXPOSE(CapRectBounds, SphereCapWithRectangularBound, RectangularPrismZ, rect_bounds);
MLGPU_EXPOSE(PlaneBounds, Plane, RectangularPrismZ, RectBounds);
MLGPU_EXPOSE(DiscBounds, Disc, CylinderPrismZ, cylinder_bounds);
MLGPU_EXPOSE(HexPlaneBounds, HexPlane, HexagonalPrismZ, hex_bounds);
MLGPU_EXPOSE(AnnulusOuter, Annulus, CylinderPrismZ, outer_bound);
MLGPU_EXPOSE(AnnulusInner, Annulus, CylinderPrismZ, inner_bound);
MLGPU_EXPOSE(DualBounds, PlaneDualSphericalBound, DualSphericalPrismZ, dual_sphere_bounds);
MLGPU_EXPOSE(CylinderRadius, Cylinder, double, Radius);
MLGPU_EXPOSE(CylinderLength, Cylinder, double, Length);
MLGPU_EXPOSE(TriangleA, Triangle, Vec2, A);
MLGPU_EXPOSE(TriangleB, Triangle, Vec2, B);
MLGPU_EXPOSE(TriangleC, Triangle, Vec2, C);
MLGPU_EXPOSE(SphereRadius, Sphere, double, radius);
MLGPU_EXPOSE(EllipFocal, EllipticalCapWithHexagonalBound, double, focal_length);
MLGPU_EXPOSE(EllipRatio, EllipticalCapWithHexagonalBound, double, ratio_x_to_y_curvature_radius);
MLGPU_EXPOSE(EllipRotation, EllipticalCapWithHexagonalBound, double, hex_bound_rotation);
MLGPU_EXPOSE(EllipHexRadius, EllipticalCapWithHexagonalBound, double, outer_hex_radius);
MLGPU_EXPOSE(RaySupport, Ray, Vec3, support_);
MLGPU_EXPOSE(RayDirection, Ray, Vec3, direction_);

#undef MLGPU_EXPOSE

[[noreturn]] void fail_hip(hipError_t e, const char* what) {
    std::stringstream s;
    s << "merlict::gpu: " << what << ": " << hipGetErrorString(e);
    throw std::runtime_error(s.str());
}

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) fail_hip(e, what);
}

void check_mlgpu(int rc, const char* what) {
    if (rc == MLGPU_OK) return;
    std::stringstream s;
    s << "merlict::gpu: " << what << ": " << mlgpu_last_error();
    if (rc == MLGPU_ERR_ARGUMENT || rc == MLGPU_ERR_SCENE || rc == MLGPU_ERR_UNSUPPORTED)
        throw std::invalid_argument(s.str());
    throw std::runtime_error(s.str());
}

// ------------------------------------------------------------ flattening
struct Flat {
    std::vector<const Frame*> frames;
    std::unordered_map<const Frame*, int32_t> index;
    std::vector<const SurfaceEntity*> entity;  // the intersected object of frame i
    std::vector<int32_t> mother, rot_mode, kind, surface, allowed, sensor;
    std::vector<double> pos, rot, params;
    std::unordered_map<const function::Func1*, int32_t> function;
    std::vector<int32_t> func_start{0};
    std::vector<double> fx, fy;

    mlgpu_scene_desc view() const {
        mlgpu_scene_desc d;
        d.n_frames = static_cast<int32_t>(frames.size());
        d.mother = mother.data();
        d.pos = pos.data();
        d.rot_mode = rot_mode.data();
        d.rot = rot.data();
        d.kind = kind.data();
        d.params = params.data();
        d.surface = surface.data();
        d.allowed_frame = allowed.data();
        d.sensor_id = sensor.data();
        d.n_functions = static_cast<int32_t>(func_start.size()) - 1;
        d.func_start = func_start.data();
        d.func_x = fx.data();
        d.func_y = fy.data();
        return d;
    }
};

// merlict's DEFAULT_* table of a surface slot (SurfaceEntity.cpp:10-27,44-51)
const function::Func1* default_of_slot(int slot) {
    if (slot <= MLGPU_INNER_REFLECTION) return SurfaceEntity::DEFAULT_REFLECTION;
    if (slot <= MLGPU_INNER_REFRACTION) return SurfaceEntity::DEFAULT_REFRACTION;
    return SurfaceEntity::DEFAULT_ABSORPTION;
}

int32_t function_index(Flat* fl, const function::Func1* f, int slot, const Frame* owner) {
    if (f == default_of_slot(slot)) return MLGPU_DEFAULT_FUNCTION;
    if (!f) throw std::invalid_argument("merlict::gpu: frame '" + owner->path_in_tree() + "' has a null surface function");
    auto it = fl->function.find(f);
    if (it != fl->function.end()) return it->second;
    const int32_t k = static_cast<int32_t>(fl->func_start.size()) - 1;
    for (const function::Point& p : f->func) {
        fl->fx.push_back(p.x);
        fl->fy.push_back(p.y);
    }
    fl->func_start.push_back(static_cast<int32_t>(fl->fx.size()));
    fl->function[f] = k;
    return k;
}

// kind and parameters of a frame, from its exact class
int32_t primitive_of(const Frame* f, double* p) {
    const std::type_info& t = typeid(*f);
    if (t == typeid(SphereCapWithHexagonalBound)) {
        const auto& s = static_cast<const SphereCapWithHexagonalBound&>(*f);
        p[0] = s.*member(CapHexCurvature());
        p[1] = s.*member(CapHexRadius());
        return MLGPU_KIND_SPHERE_CAP_HEX;
    }
    if (t == typeid(SphereCapWithCylinderBound)) {
        const auto& s = static_cast<const SphereCapWithCylinderBound&>(*f);
        p[0] = s.*member(CapCylCurvature());
        p[1] = s.*member(CapCylRadius());
        return MLGPU_KIND_SPHERE_CAP_CYL;
    }
    if (t == typeid(SphereCapWithRectangularBound)) {
        const auto& s = static_cast<const SphereCapWithRectangularBound&>(*f);
        const RectangularPrismZ& b = s.*member(CapRectBounds());
        p[0] = s.*member(CapRectCurvature());
        p[1] = 2.0 * b.get_half_x_width();  // half = 0.5 * width (RectangularPrismZ.cpp:19)
        p[2] = 2.0 * b.get_half_y_width();
        return MLGPU_KIND_SPHERE_CAP_RECT;
    }
    if (t == typeid(Plane)) {
        const RectangularPrismZ& b = static_cast<const Plane&>(*f).*member(PlaneBounds());
        p[0] = 2.0 * b.get_half_x_width();
        p[1] = 2.0 * b.get_half_y_width();
        return MLGPU_KIND_PLANE;
    }
    if (t == typeid(Disc)) {
        // sqrt(r * r) is r again in IEEE arithmetic (CylinderPrismZ keeps r^2)
        p[0] = (static_cast<const Disc&>(*f).*member(DiscBounds())).get_radius();
        return MLGPU_KIND_DISC;
    }
    if (t == typeid(HexPlane)) {
        p[0] = (static_cast<const HexPlane&>(*f).*member(HexPlaneBounds())).get_outer_radius();
        return MLGPU_KIND_HEX_PLANE;
    }
    if (t == typeid(Annulus)) {
        const auto& s = static_cast<const Annulus&>(*f);
        p[0] = (s.*member(AnnulusOuter())).get_radius();
        p[1] = (s.*member(AnnulusInner())).get_radius();
        return MLGPU_KIND_ANNULUS;
    }
    if (t == typeid(PlaneDualSphericalBound)) {
        const DualSphericalPrismZ& b = static_cast<const PlaneDualSphericalBound&>(*f).*member(DualBounds());
        p[0] = b.get_hight();
        p[1] = b.get_width();
        return MLGPU_KIND_PLANE_DUAL_SPHERICAL;
    }
    if (t == typeid(Cylinder)) {
        const auto& s = static_cast<const Cylinder&>(*f);
        p[0] = s.*member(CylinderRadius());
        p[1] = s.*member(CylinderLength());
        return MLGPU_KIND_CYLINDER;
    }
    if (t == typeid(Triangle)) {
        const auto& s = static_cast<const Triangle&>(*f);
        const Vec2 a = s.*member(TriangleA()), b = s.*member(TriangleB()), c = s.*member(TriangleC());
        p[0] = a.x; p[1] = a.y; p[2] = b.x; p[3] = b.y; p[4] = c.x; p[5] = c.y;
        return MLGPU_KIND_TRIANGLE;
    }
    if (t == typeid(Sphere)) {
        p[0] = static_cast<const Sphere&>(*f).*member(SphereRadius());
        return MLGPU_KIND_SPHERE;
    }
    if (t == typeid(EllipticalCapWithHexagonalBound)) {
        // the stored outer radius is already restricted to the z curvature
        // radius; restricting it again changes nothing
        const auto& s = static_cast<const EllipticalCapWithHexagonalBound&>(*f);
        p[0] = s.*member(EllipFocal());
        p[1] = s.*member(EllipRatio());
        p[2] = s.*member(EllipRotation());
        p[3] = s.*member(EllipHexRadius());
        return MLGPU_KIND_ELLIPTICAL_CAP_HEX;
    }
    // frames that are never intersected themselves: plain frames, surface
    // entities without a geometry of their own, and merlict's composites,
    // which only add primitive children (BiConvexLens, BiConvexLensHexBound,
    // RectangularBox)
    if (t == typeid(Frame) || t == typeid(SurfaceEntity) || t == typeid(BiConvexLens) ||
        t == typeid(BiConvexLensHexBound) || t == typeid(RectangularBox))
        return MLGPU_KIND_FRAME;
    throw std::invalid_argument("merlict::gpu: frame '" + f->path_in_tree() + "' is a " + t.name() +
                                ", a surface the GPU propagator does not model");
}

void add_frame(Flat* fl, const Frame* f, int32_t mother) {
    const int32_t i = static_cast<int32_t>(fl->frames.size());
    fl->frames.push_back(f);
    fl->index[f] = i;
    fl->mother.push_back(mother);
    const Vec3 pos = f->position_in_mother();
    fl->pos.insert(fl->pos.end(), {pos.x, pos.y, pos.z});
    const Rot3 r = f->rotation_in_mother();
    if (r.uses_xyz_angels()) {
        fl->rot_mode.push_back(MLGPU_ROT_XYZ);
        fl->rot.insert(fl->rot.end(), {r.rot_x(), r.rot_y(), r.rot_z(), 0.0});
    } else {
        const Vec3 ax = r.rot_axis();
        fl->rot_mode.push_back(MLGPU_ROT_AXIS_STORED);
        fl->rot.insert(fl->rot.end(), {ax.x, ax.y, ax.z, r.rot_angle()});
    }
    double p[MLGPU_NUM_PARAMS] = {0, 0, 0, 0, 0, 0};
    const int32_t kind = f->has_children() ? MLGPU_KIND_FRAME : primitive_of(f, p);
    if (f->has_children()) primitive_of(f, p);  // rejects unknown classes all the same
    fl->kind.push_back(kind);
    fl->params.insert(fl->params.end(), p, p + MLGPU_NUM_PARAMS);
    const auto* se = dynamic_cast<const SurfaceEntity*>(f);
    fl->entity.push_back(se);
    const function::Func1* slots[MLGPU_NUM_SURFACE_SLOTS] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (se) {
        slots[0] = se->outer_reflection;
        slots[1] = se->inner_reflection;
        slots[2] = se->outer_refraction;
        slots[3] = se->inner_refraction;
        slots[4] = se->outer_absorption;
        slots[5] = se->inner_absorption;
    }
    for (int s = 0; s < MLGPU_NUM_SURFACE_SLOTS; s++)
        fl->surface.push_back(se ? function_index(fl, slots[s], s, f) : MLGPU_DEFAULT_FUNCTION);
    fl->allowed.push_back(MLGPU_NO_FRAME);  // resolved once every frame has its index
    fl->sensor.push_back(MLGPU_NO_SENSOR);  // sensors stay with merlict's sensor::Sensors
}

Flat flatten(const Frame* root) {
    Flat fl;
    // pre-order, children in merlict's order: mother before child, and the
    // library's depth-first candidate order equals RayAndFrame.cpp:137-149's
    std::vector<std::pair<const Frame*, int32_t>> todo{{root, -1}};
    while (!todo.empty()) {
        const auto [f, mother] = todo.back();
        todo.pop_back();
        add_frame(&fl, f, mother);
        const int32_t me = static_cast<int32_t>(fl.frames.size()) - 1;
        const std::vector<Frame*>& kids = *f->get_children();
        if (kids.size() > 16) {
            // merlict leaves children it cannot cluster (positions too close)
            // in place; the library would cluster such a frame once more
            throw std::invalid_argument("merlict::gpu: frame '" + f->path_in_tree() +
                                        "' keeps more than 16 children after clustering");
        }
        for (auto it = kids.rbegin(); it != kids.rend(); ++it) todo.push_back({*it, me});
    }
    for (size_t i = 0; i < fl.frames.size(); i++) {
        const SurfaceEntity* se = fl.entity[i];
        if (!se || !se->has_restrictions_on_frames_to_propagate_to()) continue;
        auto it = fl.index.find(se->allowed_frame_to_propagate_to());
        if (it == fl.index.end())
            throw std::invalid_argument("merlict::gpu: frame '" + se->path_in_tree() +
                                        "' restricts propagation to a frame outside its tree");
        fl.allowed[i] = it->second;
    }
    return fl;
}

// ------------------------------------------------------------ device buffers
class DeviceBuffer {
 public:
    DeviceBuffer() = default;
    ~DeviceBuffer() { if (p_) (void)hipFree(p_); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void reserve(size_t bytes) {
        if (bytes <= size_) return;
        if (p_) check_hip(hipFree(p_), "hipFree");
        p_ = nullptr;
        size_ = 0;
        check_hip(hipMalloc(&p_, bytes), "hipMalloc");
        size_ = bytes;
    }
    void* get() const { return p_; }

 private:
    void* p_ = nullptr;
    size_t size_ = 0;
};

// carve typed arrays out of one allocation, 256-byte aligned (a null base
// only measures)
struct Carve {
    char* base;
    size_t off = 0;
    template <typename T>
    T* take(size_t n) {
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += (n * sizeof(T) + 255) / 256 * 256;
        return p;
    }
};

// the device arrays of one batch of n photons with history capacity cap
struct Batch {
    double* in;
    mlgpu_results r;
    mlgpu_bounce* bounce;
    int32_t* num_bounces;
    double* ray;
    Batch(Carve* c, size_t n, size_t cap) {
        in = c->take<double>(7 * n);
        r.final_frame = c->take<int32_t>(n);
        r.sensor_id = c->take<int32_t>(n);
        r.final_interaction = c->take<uint8_t>(n);
        r.num_interactions = c->take<uint8_t>(n);
        r.error = c->take<uint8_t>(n);
        r.time_of_flight = c->take<double>(n);
        r.x = c->take<float>(n);
        r.y = c->take<float>(n);
        r.cx = c->take<float>(n);
        r.cy = c->take<float>(n);
        r.accumulated_distance = c->take<double>(n);
        bounce = c->take<mlgpu_bounce>(n * cap);
        num_bounces = c->take<int32_t>(n);
        ray = c->take<double>(6 * n);
    }
};

}  // namespace

// ------------------------------------------------------------ Scene
struct Scene::Impl {
    Flat flat;
    mlgpu_scene* scene = nullptr;
    int32_t world = 0;
    int64_t batch = int64_t{1} << 20;
    mutable DeviceBuffer mem;
    ~Impl() { if (scene) mlgpu_scene_destroy(scene); }
};

Scene::Scene(const Frame* world) : impl_(new Impl) {
    if (!world) throw std::invalid_argument("merlict::gpu: null world frame");
    impl_->flat = flatten(world->root());
    impl_->world = impl_->flat.index.at(world);
    const mlgpu_scene_desc d = impl_->flat.view();
    check_mlgpu(mlgpu_scene_create(&d, &impl_->scene), "scene");
}

Scene::~Scene() = default;

int32_t Scene::frame_index(const Frame* frame) const {
    auto it = impl_->flat.index.find(frame);
    return it == impl_->flat.index.end() ? -1 : it->second;
}

int32_t Scene::num_frames() const { return static_cast<int32_t>(impl_->flat.frames.size()); }

void Scene::set_batch_size(int64_t n) {
    if (n < 1) throw std::invalid_argument("merlict::gpu: batch size must be positive");
    impl_->batch = n;
}

void Scene::propagate(std::vector<Photon>* photons, const PropagationConfig* settings,
                      random::Generator* prng) const {
    if (!prng) throw std::invalid_argument("merlict::gpu: null generator");
    propagate_with_seed(photons, settings, prng->create_seed());
}

void Scene::propagate_with_seed(std::vector<Photon>* photons, const PropagationConfig* settings,
                                uint64_t seed) const {
    if (!photons || !settings) throw std::invalid_argument("merlict::gpu: null argument");
    const int64_t total = static_cast<int64_t>(photons->size());
    if (total == 0) return;
    const uint32_t max_int = settings->max_num_interactions_per_photon;
    const size_t cap = max_int > 0 ? max_int : 1;
    const int64_t batch = std::min<int64_t>(impl_->batch, total);
    const size_t nb = static_cast<size_t>(batch);
    Carve measure{nullptr};
    (void)Batch(&measure, nb, cap);
    impl_->mem.reserve(measure.off);
    Carve c{static_cast<char*>(impl_->mem.get())};
    const Batch dev(&c, nb, cap);

    std::vector<double> in(7 * nb), ray(6 * nb);
    std::vector<uint8_t> err(nb);
    std::vector<int32_t> nbounce(nb);
    std::vector<mlgpu_bounce> bounce(nb * cap);
    const Flat& fl = impl_->flat;

    for (int64_t start = 0; start < total; start += batch) {
        const int64_t n = std::min<int64_t>(batch, total - start);
        const size_t un = static_cast<size_t>(n);
        for (size_t k = 0; k < un; k++) {
            const Photon& ph = (*photons)[static_cast<size_t>(start) + k];
            const Vec3 s = ph.support(), d = ph.direction();
            in[0 * un + k] = s.x; in[1 * un + k] = s.y; in[2 * un + k] = s.z;
            in[3 * un + k] = d.x; in[4 * un + k] = d.y; in[5 * un + k] = d.z;
            in[6 * un + k] = ph.wavelength;
        }
        // the photons packed as 7 arrays of n (not of the batch capacity)
        double* const d_in = dev.in;
        check_hip(hipMemcpy(d_in, in.data(), 7 * un * sizeof(double), hipMemcpyHostToDevice), "photons to device");
        mlgpu_photons p{d_in, d_in + un, d_in + 2 * un, d_in + 3 * un, d_in + 4 * un, d_in + 5 * un, d_in + 6 * un};
        mlgpu_results r = dev.r;
        mlgpu_history h{dev.bounce, static_cast<int32_t>(cap), dev.num_bounces, dev.ray};
        mlgpu_config cfg{max_int, seed, static_cast<uint64_t>(start)};
        check_mlgpu(mlgpu_propagate_history(impl_->scene, impl_->world, &p, n, &cfg, &r, &h, nullptr), "propagation");
        check_hip(hipMemcpy(err.data(), r.error, un, hipMemcpyDeviceToHost), "errors to host");
        check_hip(hipMemcpy(nbounce.data(), dev.num_bounces, un * sizeof(int32_t), hipMemcpyDeviceToHost),
                  "history to host");
        check_hip(hipMemcpy(bounce.data(), dev.bounce, un * cap * sizeof(mlgpu_bounce), hipMemcpyDeviceToHost),
                  "history to host");
        check_hip(hipMemcpy(ray.data(), dev.ray, 6 * un * sizeof(double), hipMemcpyDeviceToHost), "rays to host");

        for (size_t k = 0; k < un; k++) {
            Photon& ph = (*photons)[static_cast<size_t>(start) + k];
            const mlgpu_bounce* b = bounce.data() + k * cap;
            const int32_t m = nbounce[k];
            for (int32_t j = 0; j < m; j++) {
                const mlgpu_bounce& e = b[j];
                const SurfaceEntity* obj = &VOID_SURFACE_ENTITY;
                if (e.frame >= 0) {
                    obj = fl.entity[static_cast<size_t>(e.frame)];
                    if (!obj) throw std::runtime_error("merlict::gpu: a bounce on a frame without a surface");
                }
                const Vec3 nrm(e.normal[0], e.normal[1], e.normal[2]);
                // Intersection keeps only the sign of normal . incident
                // (Intersection.cpp:133-139)
                const Vec3 inc = e.from_outside ? nrm * -1.0 : nrm;
                ph.push_back_intersection_and_interaction(
                    Intersection(obj, Vec3(e.position[0], e.position[1], e.position[2]), nrm, e.distance, inc),
                    static_cast<Interaction>(e.interaction));
            }
            if (err[k] == MLGPU_PHOTON_OUT_OF_TABLE) {
                std::stringstream s;
                s << "merlict::gpu: photon " << start + static_cast<int64_t>(k) << " (wavelength "
                  << ph.wavelength << " m) meets a function table that does not cover its wavelength";
                throw std::out_of_range(s.str());
            }
            if (err[k] != MLGPU_PHOTON_OK) {
                std::stringstream s;
                s << "merlict::gpu: photon " << start + static_cast<int64_t>(k) << ": error " << int(err[k]);
                throw std::invalid_argument(s.str());
            }
            const double* q = ray.data() + 6 * k;
            ph.*member(RaySupport()) = Vec3(q[0], q[1], q[2]);
            ph.*member(RayDirection()) = Vec3(q[3], q[4], q[5]);
        }
    }
}

void propagate_photons_in_frame_with_config(std::vector<Photon>* photons, const Frame* world,
                                            const PropagationConfig* settings, random::Generator* prng) {
    if (!photons || !world || !settings || !prng) throw std::invalid_argument("merlict::gpu: null argument");
    if (photons->empty()) return;
    Scene scene(world);
    scene.propagate(photons, settings, prng);
}

void propagate_photons_in_frame_with_config_multi_thread(std::vector<Photon>* photons, const Frame* world,
                                                         const PropagationConfig* settings,
                                                         random::Generator* prng) {
    propagate_photons_in_frame_with_config(photons, world, settings, prng);
}

}  // namespace gpu
}  // namespace merlict
