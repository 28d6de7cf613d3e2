// HIP kernels of the photon propagator for CDNA4 (gfx950).
//
// One lane per photon.  Per bounce:
//   1. traversal: FP32 bounding-sphere tests over the depth-first node array
//      of the current root scope, stackless via skip indices, pruned by the
//      best hit found so far (merlict/RayAndFrame.cpp:137-149 visits every
//      candidate; the first minimum is the same);
//   2. leaf tests: FP64 in the object frame, with the ray support moved
//      along the ray to the point closest to the leaf origin so the solve
//      works on small numbers (merlict's equations, scenery/geometry/*.cpp);
//   3. physics: medium absorption, surface reflection, Fresnel, refraction
//      and root-frame switching exactly as merlict/PhotonAndFrame.cpp:23-124,
//      with one Philox4x32-10 stream per photon.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "kernels.hpp"

namespace mlgpu {
namespace {

constexpr double kC = 299792458.0;        // merlict/PhysicalConstants.h:7
constexpr double kVoidDistance = 1e4;     // RayAndFrame.cpp:161-167
constexpr double kSelfHit2 = 1e-14;       // Vec3.cpp:121-127

// ------------------------------------------------------------------ Philox
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t& o0, uint32_t& o1,
                                              uint32_t& o2, uint32_t& o3) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    o0 = c0; o1 = c1; o2 = c2; o3 = c3;
}

// Draw k of photon `id`: word k%4 of block k/4 of counter (k/4, 0, id).
struct Rng {
    uint32_t k0, k1, id_lo, id_hi;
    uint32_t b0, b1, b2, b3;
    uint32_t next;
    __device__ __forceinline__ double uniform() {
        if ((next & 3u) == 0u) philox4x32_10(next >> 2, 0u, id_lo, id_hi, k0, k1, b0, b1, b2, b3);
        const uint32_t sel = next & 3u;
        const uint32_t w = sel == 0u ? b0 : (sel == 1u ? b1 : (sel == 2u ? b2 : b3));
        next++;
        return static_cast<double>(w >> 8) * (1.0 / 16777216.0);
    }
};

// ------------------------------------------------------------------- Func1
// Func1::evaluate with Limits::assert_contains (merlict/function/Func1.cpp:66-71,
// Limits.cpp:37-46); false when x is outside [lower, upper).
__device__ __forceinline__ bool func_eval(const DeviceScene& sc, int32_t f, double x, double& y) {
    const FuncHead h = sc.funcs[f];
    if (!(x >= h.lower) || !(x < h.upper)) return false;
    if (h.flags & FUNC_DEFAULT_ABSORPTION) { y = INFINITY; return true; }
    if (h.flags & FUNC_DEFAULT_REFRACTION) { y = 1.0; return true; }
    if (h.flags & FUNC_DEFAULT_REFLECTION) { y = 0.0; return true; }
    int32_t lo = 0, hi = h.n;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (sc.points[h.start + mid].x > x) hi = mid; else lo = mid + 1;
    }
    const FuncPoint p0 = sc.points[h.start + lo - 1];
    y = p0.slope * x + p0.y - p0.slope * p0.x;
    return true;
}

// ----------------------------------------------------------------- leaves
struct LeafHit {
    double t;                 // distance from the ray support
    double px, py, pz;        // object-frame position
    double dx, dy, dz;        // object-frame direction
};

__device__ __forceinline__ bool hex_inside(double inner, double x, double y) {
    // HexagonalPrismZ::is_inside (merlict/scenery/geometry/HexagonalPrismZ.cpp:14-29)
    const double sv = 0.86602540378443871;   // sin(2 pi / 3)
    const double cv = -0.49999999999999978;  // cos(2 pi / 3)
    const double u = y;
    const double v = x * sv + y * cv;
    const double w = x * -sv + y * cv;
    return u < inner && u > -inner && v < inner && v > -inner && w < inner && w > -inner;
}

__device__ __forceinline__ bool cap_bound_inside(int32_t kind, double b1, double b2, double x, double y) {
    if (kind == MLGPU_KIND_SPHERE_CAP_HEX) return hex_inside(b1, x, y);
    if (kind == MLGPU_KIND_SPHERE_CAP_CYL) return x * x + y * y < b1;
    return fabs(x) < b1 && fabs(y) < b2;
}

// Intersection of a ray (world support s, unit direction d) with one leaf.
// Returns true and the hit when merlict's calculate_intersection_with of
// that primitive would emit an intersection.
__device__ bool intersect_leaf(const LeafGeo* __restrict__ gp, double sx, double sy, double sz, double dx, double dy,
                               double dz, LeafHit& h) {
    const LeafGeo& g = *gp;
    // move the support to the point of the ray closest to the leaf origin
    const double wx = g.tx - sx, wy = g.ty - sy, wz = g.tz - sz;
    const double t0 = wx * dx + wy * dy + wz * dz;
    const double ox = t0 * dx - wx, oy = t0 * dy - wy, oz = t0 * dz - wz;
    const double* R = g.rot;
    const double sxl = R[0] * ox + R[3] * oy + R[6] * oz;
    const double syl = R[1] * ox + R[4] * oy + R[7] * oz;
    const double szl = R[2] * ox + R[5] * oy + R[8] * oz;
    const double dxl = R[0] * dx + R[3] * dy + R[6] * dz;
    const double dyl = R[1] * dx + R[4] * dy + R[7] * dz;
    const double dzl = R[2] * dx + R[5] * dy + R[8] * dz;
    double tl;  // parameter relative to the moved support
    const int32_t kind = g.kind;
    if (kind <= MLGPU_KIND_SPHERE_CAP_RECT) {
        // SphericalCapRayIntersectionEquation.cpp:6-40 and
        // SurfaceWithOuterPrismBound::add_causeal_intersection (…Bound.cpp:7-70)
        const double Rc = g.p[0];
        const double d2 = dxl * dxl + dyl * dyl + dzl * dzl;
        const double P = 2.0 * (sxl * dxl + syl * dyl + szl * dzl - Rc * dzl) / d2;
        const double Q = (sxl * sxl + syl * syl + szl * szl - 2.0 * (Rc * szl)) / d2;
        const double p2 = 0.5 * P;
        const double inner = p2 * p2 - Q;
        if (!(inner >= 0.0)) return false;
        const double sq = sqrt(inner);
        const double mi = -p2 - sq, pl = -p2 + sq;
        const double pxp = sxl + dxl * pl, pyp = syl + dyl * pl, pzp = szl + dzl * pl;
        const double pxm = sxl + dxl * mi, pym = syl + dyl * mi, pzm = szl + dzl * mi;
        const double b1 = g.p[1], b2 = g.p[2];
        const bool p_in = cap_bound_inside(kind, b1, b2, pxp, pyp);
        const bool m_in = cap_bound_inside(kind, b1, b2, pxm, pym);
        const double rb2 = g.p[7] * g.p[7];
        const bool p_c = (t0 + pl) > 0.0 && pxp * pxp + pyp * pyp + pzp * pzp <= rb2;
        const bool m_c = (t0 + mi) > 0.0 && pxm * pxm + pym * pym + pzm * pzm <= rb2;
        bool ok = false;
        if (p_in && m_in) {
            if (p_c && m_c) { tl = (mi >= pl) ? pl : mi; ok = true; }
            else if (p_c) { tl = pl; ok = true; }
            else if (m_c) { tl = mi; ok = true; }
        } else if (p_in && p_c) { tl = pl; ok = true; }
        else if (m_in && m_c) { tl = mi; ok = true; }
        if (!ok) return false;
    } else if (kind == MLGPU_KIND_CYLINDER) {
        // ZaxisCylinderRayIntersectionEquation.cpp:6-57, Cylinder.cpp:121-150
        const double dnom = dxl * dxl + dyl * dyl;
        const double P = 2.0 * (sxl * dxl + syl * dyl) / dnom;
        const double Q = (sxl * sxl + syl * syl - g.p[0] * g.p[0]) / dnom;
        const double p2 = 0.5 * P;
        const double inner = p2 * p2 - Q;
        if (!(inner >= 0.0)) return false;
        const double sq = sqrt(inner);
        const double tp = t0 + (-p2 + sq), tm = t0 + (-p2 - sq);
        bool causal = false;
        if (tm < 0.0 && tp >= 0.0) { causal = true; tl = -p2 + sq; }
        if (tp >= 0.0 && tm >= 0.0) { causal = true; tl = -p2 - sq; }
        if (!causal) return false;
        if (!(fabs(szl + dzl * tl) <= g.p[1])) return false;
    } else if (kind == MLGPU_KIND_SPHERE) {
        // Sphere.cpp:62-105
        const double d2 = dxl * dxl + dyl * dyl + dzl * dzl;
        const double P = 2.0 * ((sxl * dxl + syl * dyl + szl * dzl) / d2);
        const double Q = (sxl * sxl + syl * syl + szl * szl) / d2 - g.p[0] * g.p[0];
        const double p2 = 0.5 * P;
        const double inner = p2 * p2 - Q;
        if (!(inner >= 0.0)) return false;
        const double sq = sqrt(inner);
        const double vP = t0 + (-p2 + sq), vM = t0 + (-p2 - sq);
        if (vP >= 0.0 && vM >= 0.0) tl = -p2 - sq;
        else if (!(vP < 0.0 && vM < 0.0)) tl = -p2 + sq;
        else return false;
    } else {
        // XyPlaneRayIntersectionEquation.cpp:6-41 and the planar primitives
        if (dzl == 0.0 && (dxl != 0.0 || dyl != 0.0)) return false;
        tl = -szl / dzl;
        if (!((t0 + tl) > 0.0)) return false;
        const double x = sxl + dxl * tl, y = syl + dyl * tl;
        bool in;
        switch (kind) {
        case MLGPU_KIND_PLANE: in = fabs(x) < g.p[1] && fabs(y) < g.p[2]; break;
        case MLGPU_KIND_DISC: in = x * x + y * y < g.p[1]; break;
        case MLGPU_KIND_HEX_PLANE: in = hex_inside(g.p[1], x, y); break;
        case MLGPU_KIND_ANNULUS: {
            const double r2 = x * x + y * y;
            in = r2 < g.p[1] && !(r2 < g.p[2]);
            break;
        }
        case MLGPU_KIND_PLANE_DUAL_SPHERICAL: {
            // DualSphericalPrismZ::is_inside (DualSphericalPrismZ.cpp:21-27)
            const double ax = g.p[1] - x, bx = -g.p[1] - x;
            in = ax * ax + y * y <= g.p[2] && bx * bx + y * y <= g.p[2];
            break;
        }
        case MLGPU_KIND_TRIANGLE: {
            // Triangle::is_inside_triangle (Triangle.cpp:92-102)
            const double Ax = g.p[0], Ay = g.p[1], Bx = g.p[2], By = g.p[3], Cx = g.p[4], Cy = g.p[5];
            const bool bA = ((x - Bx) * (Ay - By) - (Ax - Bx) * (y - By)) < 0.0;
            const bool bB = ((x - Cx) * (By - Cy) - (Bx - Cx) * (y - Cy)) < 0.0;
            const bool bC = ((x - Ax) * (Cy - Ay) - (Cx - Ax) * (y - Ay)) < 0.0;
            in = (bA == bB) && (bB == bC);
            break;
        }
        default: in = false;
        }
        if (!in) return false;
    }
    const double t = t0 + tl;
    // Vec3 != with |p - s|^2 > 1e-14 (the drop of self hits)
    if (!(t * t * (dxl * dxl + dyl * dyl + dzl * dzl) > kSelfHit2)) return false;
    h.t = t;
    h.px = sxl + dxl * tl;
    h.py = syl + dyl * tl;
    h.pz = szl + dzl * tl;
    h.dx = dxl; h.dy = dyl; h.dz = dzl;
    return true;
}

__device__ __forceinline__ void leaf_normal(const LeafGeo* __restrict__ gp, const LeafHit& h, double& nx, double& ny, double& nz) {
    const LeafGeo& g = *gp;
    const int32_t kind = g.kind;
    if (kind <= MLGPU_KIND_SPHERE_CAP_RECT) {
        // SphericalCapRayIntersectionEquation.cpp:52-59
        nx = 0.0 - h.px; ny = 0.0 - h.py; nz = g.p[0] - h.pz;
        const double n = sqrt(nx * nx + ny * ny + nz * nz);
        nx /= n; ny /= n; nz /= n;
    } else if (kind == MLGPU_KIND_CYLINDER) {
        const double n = sqrt(h.px * h.px + h.py * h.py);
        nx = h.px / n; ny = h.py / n; nz = 0.0;
    } else if (kind == MLGPU_KIND_SPHERE) {
        const double n = sqrt(h.px * h.px + h.py * h.py + h.pz * h.pz);
        nx = h.px / n; ny = h.py / n; nz = h.pz / n;
    } else {
        nx = 0.0; ny = 0.0; nz = 1.0;
    }
}

// --------------------------------------------------------------- traversal
struct Closest {
    int32_t leaf;
    LeafHit hit;
};

// First intersection below scope node `scope` (rays_first_intersection_with_frame).
__device__ __forceinline__ void first_hit(const DeviceScene& sc, int32_t scope, double sx, double sy, double sz,
                                          double dx, double dy, double dz, Closest& best) {
    best.leaf = -1;
    best.hit.t = INFINITY;
    const float ox = static_cast<float>(sx), oy = static_cast<float>(sy), oz = static_cast<float>(sz);
    const float fdx = static_cast<float>(dx), fdy = static_cast<float>(dy), fdz = static_cast<float>(dz);
    float best_f = INFINITY;
    const int32_t end = sc.nodes[scope].skip;
    int32_t i = scope;
    while (i < end) {
        const Node nd = sc.nodes[i];
        const float wx = nd.cx - ox, wy = nd.cy - oy, wz = nd.cz - oz;
        const float alpha = wx * fdx + wy * fdy + wz * fdz;
        const float qx = wx - alpha * fdx, qy = wy - alpha * fdy, qz = wz - alpha * fdz;
        const float dist2 = qx * qx + qy * qy + qz * qz;
        const float r2 = nd.r * nd.r;
        // ray_intersects_frames_bounding_sphere (RayAndFrame.cpp:16-107)
        bool hit = dist2 <= r2;
        if (hit && alpha < 0.0f) hit = wx * wx + wy * wy + wz * wz <= r2;
        // prune spheres entered only behind the best hit
        if (hit && alpha > best_f) {
            const float e = alpha - best_f;
            hit = e * e <= r2 - dist2;
        }
        if (!hit) { i = nd.skip; continue; }
        if (nd.leaf >= 0) {
            LeafHit h;
            if (intersect_leaf(sc.geo + nd.leaf, sx, sy, sz, dx, dy, dz, h) && h.t < best.hit.t) {
                best.leaf = nd.leaf;
                best.hit = h;
                best_f = static_cast<float>(h.t) * (1.0f + 1e-6f) + sc.eps;
            }
        }
        i++;
    }
}

// ------------------------------------------------------------------ kernels
__global__ __launch_bounds__(256) void propagate_kernel(const DeviceScene sc, const int32_t world_node,
                                                        const mlgpu_photons in, const int64_t n,
                                                        const uint32_t max_interactions, const uint64_t seed,
                                                        const uint64_t id_offset, const mlgpu_results out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sx = in.support_x[i], sy = in.support_y[i], sz = in.support_z[i];
    double dx = in.direction_x[i], dy = in.direction_y[i], dz = in.direction_z[i];
    {
        const double nn = sqrt(dx * dx + dy * dy + dz * dz);  // Ray.cpp:16-20
        dx /= nn; dy /= nn; dz /= nn;
    }
    const double lambda = in.wavelength[i];
    const uint64_t id = id_offset + static_cast<uint64_t>(i);
    Rng rng;
    rng.k0 = static_cast<uint32_t>(seed);
    rng.k1 = static_cast<uint32_t>(seed >> 32);
    rng.id_lo = static_cast<uint32_t>(id);
    rng.id_hi = static_cast<uint32_t>(id >> 32);
    rng.b0 = rng.b1 = rng.b2 = rng.b3 = 0u;
    rng.next = 0u;

    uint8_t err = lambda > 0.0 ? MLGPU_PHOTON_OK : MLGPU_PHOTON_BAD_WAVELENGTH;
    int32_t root = world_node;
    uint32_t n_hist = 1;
    uint8_t last_interaction = MLGPU_PRODUCTION;
    int32_t last_frame = -1;
    int32_t last_sensor = -1;
    double tof = 0.0;
    double out_x = sx, out_y = sy, out_cx = 0.0, out_cy = 0.0;

    while (err == MLGPU_PHOTON_OK && n_hist <= max_interactions) {
        Closest c;
        first_hit(sc, root, sx, sy, sz, dx, dy, dz, c);
        if (c.leaf < 0) {
            // no hit: absorbed in void 1e4 m downstream
            double nv;
            if (!func_eval(sc, sc.default_refraction, lambda, nv)) { err = MLGPU_PHOTON_OUT_OF_TABLE; break; }
            tof += nv * kVoidDistance / kC;
            n_hist++;
            last_interaction = MLGPU_ABSORPTION_IN_VOID;
            last_frame = -1;
            last_sensor = -1;
            out_x = sx + dx * kVoidDistance;
            out_y = sy + dy * kVoidDistance;
            out_cx = -dx;
            out_cy = -dy;
            break;
        }
        const LeafGeo& g = sc.geo[c.leaf];
        const LeafPhys ph = sc.phys[c.leaf];
        double nx, ny, nz;
        leaf_normal(&g, c.hit, nx, ny, nz);
        // Intersection::ray_runs_from_outside_to_inside (Intersection.cpp:133-139)
        const bool outside_in = (c.hit.dx * nx + c.hit.dy * ny + c.hit.dz * nz) < 0.0;
        const double* R = g.rot;
        const double hwx = R[0] * c.hit.px + R[1] * c.hit.py + R[2] * c.hit.pz + g.tx;
        const double hwy = R[3] * c.hit.px + R[4] * c.hit.py + R[5] * c.hit.pz + g.ty;
        const double hwz = R[6] * c.hit.px + R[7] * c.hit.py + R[8] * c.hit.pz + g.tz;

        uint8_t interaction;
        bool stop = true;
        double ndx = dx, ndy = dy, ndz = dz;
        int32_t new_root = root;
        // absorbed_in_medium_before_reaching_surface (PhotonAndFrame.cpp:32-38)
        double ell;
        if (!func_eval(sc, ph.surf[outside_in ? MLGPU_OUTER_ABSORPTION : MLGPU_INNER_ABSORPTION], lambda, ell)) {
            err = MLGPU_PHOTON_OUT_OF_TABLE;
            break;
        }
        const double survival = exp(-c.hit.t / ell);
        if (rng.uniform() > survival) {
            interaction = MLGPU_ABSORPTION_IN_VOID;
        } else {
            double refl;
            if (!func_eval(sc, ph.surf[outside_in ? MLGPU_OUTER_REFLECTION : MLGPU_INNER_REFLECTION], lambda, refl)) {
                err = MLGPU_PHOTON_OUT_OF_TABLE;
                break;
            }
            bool mirror = false;
            if (refl >= rng.uniform()) {
                interaction = MLGPU_REFLECTION_ON_SURFACE;
                mirror = true;
            } else {
                const bool transparent = ph.surf[MLGPU_OUTER_REFRACTION] != sc.default_refraction ||
                                         ph.surf[MLGPU_INNER_REFRACTION] != sc.default_refraction;
                if (!transparent) {
                    interaction = MLGPU_ABSORPTION_ON_SURFACE;
                } else {
                    // Fresnel (merlict/Fresnel.cpp:7-38) in the object frame
                    double n_from, n_to;
                    const int32_t f_from = ph.surf[outside_in ? MLGPU_OUTER_REFRACTION : MLGPU_INNER_REFRACTION];
                    const int32_t f_to = ph.surf[outside_in ? MLGPU_INNER_REFRACTION : MLGPU_OUTER_REFRACTION];
                    if (!func_eval(sc, f_from, lambda, n_from) || !func_eval(sc, f_to, lambda, n_to)) {
                        err = MLGPU_PHOTON_OUT_OF_TABLE;
                        break;
                    }
                    const double fx = outside_in ? nx : -nx, fy = outside_in ? ny : -ny, fz = outside_in ? nz : -nz;
                    const double cosI = -(fx * c.hit.dx + fy * c.hit.dy + fz * c.hit.dz);
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
                        interaction = MLGPU_FRESNEL_REFLECTION_ON_SURFACE;
                        mirror = true;
                    } else {
                        // pass_the_boundary_layer (PhotonAndFrame.cpp:89-118)
                        interaction = outside_in ? MLGPU_REFRACTION_TO_INSIDE : MLGPU_REFRACTION_TO_OUTSIDE;
                        const bool to_default = (outside_in ? ph.surf[MLGPU_INNER_REFRACTION]
                                                            : ph.surf[MLGPU_OUTER_REFRACTION]) == sc.default_refraction;
                        new_root = (ph.allowed_node >= 0 && !to_default) ? ph.allowed_node : ph.root_node;
                        const double k = eta * cosI - cosT;
                        const double rx = c.hit.dx * eta + fx * k;
                        const double ry = c.hit.dy * eta + fy * k;
                        const double rz = c.hit.dz * eta + fz * k;
                        ndx = R[0] * rx + R[1] * ry + R[2] * rz;
                        ndy = R[3] * rx + R[4] * ry + R[5] * rz;
                        ndz = R[6] * rx + R[7] * ry + R[8] * rz;
                        stop = false;
                    }
                }
            }
            if (mirror) {
                // mirror matrix I - 2 n n^T in the object frame (Vec3.cpp:59-113)
                const double dn = nx * c.hit.dx + ny * c.hit.dy + nz * c.hit.dz;
                const double mx = c.hit.dx - 2.0 * dn * nx;
                const double my = c.hit.dy - 2.0 * dn * ny;
                const double mz = c.hit.dz - 2.0 * dn * nz;
                ndx = R[0] * mx + R[1] * my + R[2] * mz;
                ndy = R[3] * mx + R[4] * my + R[5] * mz;
                ndz = R[6] * mx + R[7] * my + R[8] * mz;
                stop = false;
            }
        }
        // push_back_intersection_and_interaction + Photon::time_of_flight
        {
            double n_from;
            if (!func_eval(sc, ph.surf[outside_in ? MLGPU_OUTER_REFRACTION : MLGPU_INNER_REFRACTION], lambda, n_from)) {
                err = MLGPU_PHOTON_OUT_OF_TABLE;
                break;
            }
            tof += n_from * c.hit.t / kC;
        }
        n_hist++;
        last_interaction = interaction;
        last_frame = ph.frame;
        last_sensor = ph.sensor;
        out_x = c.hit.px;
        out_y = c.hit.py;
        out_cx = -c.hit.dx;
        out_cy = -c.hit.dy;
        if (stop) break;
        sx = hwx; sy = hwy; sz = hwz;
        const double nn = sqrt(ndx * ndx + ndy * ndy + ndz * ndz);
        dx = ndx / nn; dy = ndy / nn; dz = ndz / nn;
        root = new_root;
    }
    out.final_frame[i] = last_frame;
    out.sensor_id[i] = last_sensor;
    out.final_interaction[i] = last_interaction;
    out.num_interactions[i] = static_cast<uint8_t>(n_hist);
    out.error[i] = err;
    out.time_of_flight[i] = tof;
    out.x[i] = static_cast<float>(out_x);
    out.y[i] = static_cast<float>(out_y);
    out.cx[i] = static_cast<float>(out_cx);
    out.cy[i] = static_cast<float>(out_cy);
}

__global__ __launch_bounds__(256) void first_hit_kernel(const DeviceScene sc, const int32_t root_node,
                                                        const mlgpu_photons rays, const int64_t n,
                                                        const mlgpu_hits out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double sx = rays.support_x[i], sy = rays.support_y[i], sz = rays.support_z[i];
    double dx = rays.direction_x[i], dy = rays.direction_y[i], dz = rays.direction_z[i];
    const double nn = sqrt(dx * dx + dy * dy + dz * dz);
    dx /= nn; dy /= nn; dz /= nn;
    Closest c;
    first_hit(sc, root_node, sx, sy, sz, dx, dy, dz, c);
    if (c.leaf < 0) {
        out.frame[i] = -1;
        out.distance[i] = kVoidDistance;
        out.position_x[i] = sx + dx * kVoidDistance;
        out.position_y[i] = sy + dy * kVoidDistance;
        out.position_z[i] = sz + dz * kVoidDistance;
        return;
    }
    const LeafGeo& g = sc.geo[c.leaf];
    out.frame[i] = sc.phys[c.leaf].frame;
    out.distance[i] = c.hit.t;
    const double* R = g.rot;
    out.position_x[i] = R[0] * c.hit.px + R[1] * c.hit.py + R[2] * c.hit.pz + g.tx;
    out.position_y[i] = R[3] * c.hit.px + R[4] * c.hit.py + R[5] * c.hit.pz + g.ty;
    out.position_z[i] = R[6] * c.hit.px + R[7] * c.hit.py + R[8] * c.hit.pz + g.tz;
}

__global__ void count_errors_kernel(const uint8_t* err, int64_t n, unsigned long long* counter) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool bad = i < n && err[i] != 0;
    const unsigned long long ballot = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && ballot) atomicAdd(counter, static_cast<unsigned long long>(__popcll(ballot)));
}

}  // namespace

hipError_t launch_propagate(const DeviceScene& sc, int32_t world_node, const mlgpu_photons& in, int64_t n,
                            uint32_t max_interactions, uint64_t seed, uint64_t id_offset,
                            const mlgpu_results& out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int block = 256;
    const int64_t grid = (n + block - 1) / block;
    hipLaunchKernelGGL(propagate_kernel, dim3(static_cast<uint32_t>(grid)), dim3(block), 0, stream, sc, world_node,
                       in, n, max_interactions, seed, id_offset, out);
    return hipGetLastError();
}

hipError_t launch_first_hit(const DeviceScene& sc, int32_t root_node, const mlgpu_photons& rays, int64_t n,
                            const mlgpu_hits& out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int block = 256;
    const int64_t grid = (n + block - 1) / block;
    hipLaunchKernelGGL(first_hit_kernel, dim3(static_cast<uint32_t>(grid)), dim3(block), 0, stream, sc, root_node,
                       rays, n, out);
    return hipGetLastError();
}

hipError_t launch_count_errors(const uint8_t* err, int64_t n, unsigned long long* counter, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int block = 256;
    const int64_t grid = (n + block - 1) / block;
    hipLaunchKernelGGL(count_errors_kernel, dim3(static_cast<uint32_t>(grid)), dim3(block), 0, stream, err, n,
                       counter);
    return hipGetLastError();
}

}  // namespace mlgpu
