// Image rendering on the GPU: merlict_visual's pin-hole camera
// (merlict_visual/PinHoleCamera.cpp:10-77) and Tracer (Tracer.cpp:21-145),
// one lane per pixel, on the same first-hit traversal as the propagator.
//
// The Tracer's recursion (trace_back -> reflection / refraction ->
// trace_back) becomes a loop over one ray per lane; its decisions use the
// propagator's leaf solves, normals, mirror and Fresnel arithmetic, so a
// pixel's ray follows the reference's path bit for bit under the same
// uniforms.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "propagate_core.hpp"
#include "stream_rng.hpp"

namespace mlgpu {
namespace {

constexpr int kRenderBlock = 256;
constexpr uint32_t kRenderStream = 10;
constexpr double kRenderWavelength = 533e-9;  // merlict_visual/Tracer.h:20
constexpr double kMaxDarkening = 0.750;       // Tracer.cpp:106

// Vec3::mirror (merlict/Vec3.cpp:59-113): v reflected at the plane of normal n
__device__ __forceinline__ D3 mirror(const D3& n, const D3& v) {
    const double x = n.x, y = n.y, z = n.z;
    return D3{(1. - 2. * x * x) * v.x + -2. * x * y * v.y + -2. * x * z * v.z,
              -2. * x * y * v.x + (1. - 2. * y * y) * v.y + -2. * y * z * v.z,
              -2. * x * z * v.x + -2. * y * z * v.y + (1. - 2. * z * z) * v.z};
}

struct Rgb {
    float r, g, b;
};

__global__ __launch_bounds__(kRenderBlock) void render_kernel(const DeviceScene sc, const int32_t world_scope,
                                                              const RenderCamera cam, const RenderParams prm,
                                                              const float* __restrict__ colors, const uint64_t seed,
                                                              float* __restrict__ image,
                                                              unsigned long long* __restrict__ errors) {
    __shared__ int32_t stk_mem[kMaxTreeDepth * kRenderBlock];
    const LaneStack stk{stk_mem + threadIdx.x, kRenderBlock};
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= static_cast<int64_t>(cam.rows) * cam.cols) return;
    const int32_t row = static_cast<int32_t>(p / cam.cols), col = static_cast<int32_t>(p % cam.cols);
    // PinHoleCamera::get_intersection_of_ray_on_image_sensor_for_pixel
    const double ri = static_cast<double>(row - cam.rows / 2), ci = static_cast<double>(col - cam.cols / 2);
    const D3 a{cam.pp[0] + cam.srow[0] * ri, cam.pp[1] + cam.srow[1] * ri, cam.pp[2] + cam.srow[2] * ri};
    D3 d = normalized(D3{a.x + cam.scol[0] * ci, a.y + cam.scol[1] * ci, a.z + cam.scol[2] * ci});
    D3 s{cam.pos[0], cam.pos[1], cam.pos[2]};
    int32_t scope = world_scope;
    StreamRng rng;
    rng.init(seed, static_cast<uint64_t>(p), kRenderStream);
    Rgb out{0.0f, 0.0f, 0.0f};
    bool bad = false;
    int32_t bounces = 0;
    while (true) {
        Closest c;
        first_hit(sc, scope, s, d, c, stk);
        if (c.leaf < 0) {
            // trace_back_to_sky_dome: the background colour
            out = Rgb{prm.sky[0], prm.sky[1], prm.sky[2]};
            break;
        }
        const LeafGeo& g = sc.geo[c.leaf];
        const LeafPhys ph = sc.phys[c.leaf];
        const D3 ld = orient_inv(g, d);
        const D3 hp = at(position_inv(g, s), ld, c.t);
        const D3 n = leaf_normal(g, hp);
        const bool outside_in = dot(n, ld) < 0.0;
        const D3 hit_world = position(g, hp);
        double refl;
        if (!func_eval(sc, ph.surf[outside_in ? MLGPU_OUTER_REFLECTION : MLGPU_INNER_REFLECTION], kRenderWavelength,
                       refl)) {
            bad = true;
            break;
        }
        bool reflect = false;
        D3 refr{0.0, 0.0, 0.0};
        int32_t new_scope = scope;
        if (refl >= rng.uniform()) {
            reflect = true;  // trace_back_after_reflection
        } else {
            const bool transparent = ph.surf[MLGPU_OUTER_REFRACTION] != sc.default_refraction ||
                                     ph.surf[MLGPU_INNER_REFRACTION] != sc.default_refraction;
            if (!transparent) {
                // trace_back_to_surface
                const float* fc = colors + 6 * ph.frame + (outside_in ? 0 : 3);
                out = Rgb{fc[0], fc[1], fc[2]};
                if (prm.gi) {
                    // shadow_of_sky_light (Tracer.cpp:104-125)
                    const D3 L{prm.light[0], prm.light[1], prm.light[2]};
                    const D3 spec = orient(g, mirror(n, orient_inv(g, L)));
                    double darkening = dot(spec, d);
                    if (darkening < 0.0) darkening = 0.0;
                    if (darkening > kMaxDarkening) darkening = kMaxDarkening;
                    // surface_iluminated_by_global_light_source (:127-145)
                    Closest sh;
                    first_hit(sc, scope, hit_world, normalized(L), sh, stk);
                    const D3 nr = orient(g, n);
                    double pr = dot(nr, L);
                    if (dot(nr, d) > 0.0) pr = pr * -1.0;
                    const bool lit = !(sh.leaf >= 0 || pr < 0.0);
                    if (!lit) darkening = darkening * 0.25;
                    // Color::reflection_mix with COLOR_BLACK (merlict/Color.cpp:35-39)
                    const double m = kMaxDarkening - darkening;
                    out.r = static_cast<float>((1. - m) * out.r + m * 0.0f);
                    out.g = static_cast<float>((1. - m) * out.g + m * 0.0f);
                    out.b = static_cast<float>((1. - m) * out.b + m * 0.0f);
                }
                break;
            }
            // trace_back_to_fresnel_interaction (Fresnel.cpp:7-38)
            double n_from, n_to;
            const int32_t f_from = ph.surf[outside_in ? MLGPU_OUTER_REFRACTION : MLGPU_INNER_REFRACTION];
            const int32_t f_to = ph.surf[outside_in ? MLGPU_INNER_REFRACTION : MLGPU_OUTER_REFRACTION];
            if (!func_eval(sc, f_from, kRenderWavelength, n_from) || !func_eval(sc, f_to, kRenderWavelength, n_to)) {
                bad = true;
                break;
            }
            const D3 nrm = outside_in ? n : D3{n.x * -1.0, n.y * -1.0, n.z * -1.0};
            const double cosI = dot(nrm, ld) * (-1.0);
            const double eta = n_from / n_to;
            const double sinT2 = eta * eta * (1.0 - cosI * cosI);
            const double cosT = sqrt(1.0 - sinT2);
            double rf;
            if (sinT2 > 1.0) {
                rf = 1.0;
            } else {
                const double ro = (n_from * cosI - n_to * cosT) / (n_from * cosI + n_to * cosT);
                const double rp = (n_to * cosI - n_from * cosT) / (n_to * cosI + n_from * cosT);
                rf = (ro * ro + rp * rp) / 2.0;
            }
            if (rf > rng.uniform()) {
                reflect = true;
            } else {
                // trace_back_beyond_boundary_layer (Tracer.cpp:82-97)
                const bool to_default = (outside_in ? ph.surf[MLGPU_INNER_REFRACTION]
                                                    : ph.surf[MLGPU_OUTER_REFRACTION]) == sc.default_refraction;
                new_scope = (ph.allowed_scope >= 0 && !to_default) ? ph.allowed_scope : ph.root_scope;
                const double k = eta * cosI - cosT;
                refr = D3{ld.x * eta + nrm.x * k, ld.y * eta + nrm.y * k, ld.z * eta + nrm.z * k};
            }
        }
        if (bounces == prm.max_depth) {
            out = Rgb{0.0f, 0.0f, 0.0f};
            break;
        }
        bounces++;
        d = normalized(orient(g, reflect ? mirror(n, ld) : refr));
        s = hit_world;
        scope = new_scope;
    }
    if (bad) {
        const float nan = __builtin_nanf("");
        out = Rgb{nan, nan, nan};
        atomicAdd(errors, 1ull);
    }
    float* px = image + 3 * p;
    px[0] = out.r;
    px[1] = out.g;
    px[2] = out.b;
}

}  // namespace

hipError_t launch_render(const DeviceScene& sc, int32_t world_scope, const RenderCamera& cam,
                         const RenderParams& prm, const float* colors, uint64_t seed, float* image,
                         unsigned long long* errors, hipStream_t stream) {
    const int64_t n = static_cast<int64_t>(cam.rows) * cam.cols;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(render_kernel, dim3(static_cast<uint32_t>((n + kRenderBlock - 1) / kRenderBlock)),
                       dim3(kRenderBlock), 0, stream, sc, world_scope, cam, prm, colors, seed, image, errors);
    return hipGetLastError();
}

}  // namespace mlgpu
