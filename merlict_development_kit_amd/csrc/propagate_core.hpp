// Per-photon propagation shared by the HIP kernels (propagate.hip).
//
// Everything here is __host__ __device__ so a host build of the same code
// can be stepped through on a CPU; the library itself only runs it on the GPU.
//
// Precision split:
//   * traversal: FP32 slab tests against axis-aligned boxes (the bulk of the
//     arithmetic), made conservative by inflating every box by the float
//     rounding bound, so the candidate set always contains merlict's;
//   * leaf solves and physics: FP64 in merlict's own operation order, with
//     floating-point contraction disabled for this code (the library is built
//     with -ffp-contract=off), so every decision that merlict makes from a
//     rounding-level quantity -- e.g. which root of a cap quadratic is the
//     self-intersection after a reflection -- is made the same way.
#pragma once

#include <cmath>
#include <cstdint>

#include "../../include/mlgpu.h"
#include "device_scene.hpp"

#if defined(__HIPCC__) || defined(__HIP__)
#define MLGPU_HD __host__ __device__
#else
#define MLGPU_HD
#endif

namespace mlgpu {

constexpr double kC = 299792458.0;     // merlict/PhysicalConstants.h:7
constexpr double kVoidDistance = 1e4;  // RayAndFrame.cpp:161-167
constexpr double kSelfHit2 = 1e-14;    // Vec3.cpp:121-127

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
MLGPU_HD inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                   uint32_t& o0, uint32_t& o1, uint32_t& o2, uint32_t& o3) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
        const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = static_cast<uint32_t>(p1); c2 = n2; c3 = static_cast<uint32_t>(p0);
    }
    o0 = c0; o1 = c1; o2 = c2; o3 = c3;
}

// Draw k of photon `id`: word k%4 of block k/4 of counter (k/4, 0, id_lo, id_hi)
// under key (seed_lo, seed_hi); uniform = (word >> 8) * 2^-24.
struct Rng {
    uint32_t k0, k1, id_lo, id_hi;
    uint32_t b0, b1, b2, b3;
    uint32_t next;   // index of the next draw
    bool have;       // b0..b3 hold block next / 4
    MLGPU_HD inline void init(uint64_t seed, uint64_t id, uint32_t draw) {
        k0 = static_cast<uint32_t>(seed);
        k1 = static_cast<uint32_t>(seed >> 32);
        id_lo = static_cast<uint32_t>(id);
        id_hi = static_cast<uint32_t>(id >> 32);
        b0 = b1 = b2 = b3 = 0u;
        next = draw;
        have = false;
    }
    MLGPU_HD inline double uniform() {
        if ((next & 3u) == 0u || !have) {
            philox4x32_10(next >> 2, 0u, id_lo, id_hi, k0, k1, b0, b1, b2, b3);
            have = true;
        }
        const uint32_t sel = next & 3u;
        const uint32_t w = sel == 0u ? b0 : (sel == 1u ? b1 : (sel == 2u ? b2 : b3));
        next++;
        return static_cast<double>(w >> 8) * (1.0 / 16777216.0);
    }
};

// ------------------------------------------------------------------- Func1
// Func1::evaluate with Limits::assert_contains (merlict/function/Func1.cpp:66-71,
// Limits.cpp:37-46); false when x is outside [lower, upper).  The segment is
// the one upper_bound picks (x_k <= x < x_k+1, Func1.cpp:118-130), found from
// the bucket index: the bucket's segment, moved down or up until it holds x
// (a step at most when the rounding of the bucket put x next door).
MLGPU_HD inline bool func_eval(const DeviceScene& sc, int32_t f, double x, double& y) {
    const FuncHead h = sc.funcs[f];
    if (!(x >= h.lower) || !(x < h.upper)) return false;
    if (h.flags & FUNC_DEFAULT_ABSORPTION) { y = HUGE_VAL; return true; }
    if (h.flags & FUNC_DEFAULT_REFRACTION) { y = 1.0; return true; }
    if (h.flags & FUNC_DEFAULT_REFLECTION) { y = 0.0; return true; }
    int32_t b = static_cast<int32_t>((x - h.lower) * h.scale);
    b = b < 0 ? 0 : (b >= h.n_buckets ? h.n_buckets - 1 : b);
    int32_t k = sc.buckets[h.bucket_start + b];
    const FuncPoint* pt = sc.points + h.start;
    while (k > 0 && pt[k].x > x) k--;
    while (k + 2 < h.n && pt[k + 1].x <= x) k++;
    const FuncPoint p0 = pt[k];
    y = p0.slope * x + p0.y - p0.slope * p0.x;
    return true;
}

// func_eval of N functions at the same x with their table reads
// interleaved: the heads, then the buckets, then the points of all N are
// read together, so the lookups wait on three rounds of memory instead of
// three per function.  ok[i] is false where x lies outside function i's
// range; y[i] is then unset.  Each y[i] is the value func_eval gives.
template <int N>
MLGPU_HD inline void func_eval_n(const DeviceScene& sc, const int32_t (&f)[N], const double x, double (&y)[N],
                                 bool (&ok)[N]) {
    FuncHead h[N];
#pragma unroll
    for (int i = 0; i < N; i++) h[i] = sc.funcs[f[i]];
    bool table[N];
    int32_t k[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        ok[i] = x >= h[i].lower && x < h[i].upper;
        table[i] = ok[i] && (h[i].flags & (FUNC_DEFAULT_ABSORPTION | FUNC_DEFAULT_REFRACTION |
                                           FUNC_DEFAULT_REFLECTION)) == 0;
        k[i] = 0;
        if (table[i]) {
            int32_t b = static_cast<int32_t>((x - h[i].lower) * h[i].scale);
            b = b < 0 ? 0 : (b >= h[i].n_buckets ? h[i].n_buckets - 1 : b);
            k[i] = sc.buckets[h[i].bucket_start + b];
        }
    }
    FuncPoint a[N];
    double xn[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        if (!table[i]) continue;
        const FuncPoint* pt = sc.points + h[i].start;
        a[i] = pt[k[i]];
        xn[i] = k[i] + 1 < h[i].n ? pt[k[i] + 1].x : HUGE_VAL;
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
        if (h[i].flags & FUNC_DEFAULT_ABSORPTION) { y[i] = HUGE_VAL; continue; }
        if (h[i].flags & FUNC_DEFAULT_REFRACTION) { y[i] = 1.0; continue; }
        if (h[i].flags & FUNC_DEFAULT_REFLECTION) { y[i] = 0.0; continue; }
        if (!table[i]) continue;
        // the bucket's segment, moved down or up until it holds x, as func_eval
        // does (x is sorted strictly, so at most one of the two moves happens)
        const FuncPoint* pt = sc.points + h[i].start;
        int32_t kk = k[i];
        if (kk > 0 && a[i].x > x) {
            do { kk--; } while (kk > 0 && pt[kk].x > x);
            a[i] = pt[kk];
        } else if (kk + 2 < h[i].n && xn[i] <= x) {
            do { kk++; } while (kk + 2 < h[i].n && pt[kk + 1].x <= x);
            a[i] = pt[kk];
        }
        y[i] = a[i].slope * x + a[i].y - a[i].slope * a[i].x;
    }
}

// ------------------------------------------------------------- transforms
struct D3 {
    double x, y, z;
};

// Vec3::operator*(v) = v.x*x + v.y*y + v.z*z (merlict/Vec3.cpp:31-33)
MLGPU_HD inline double dot(const D3& a, const D3& b) { return b.x * a.x + b.y * a.y + b.z * a.z; }
MLGPU_HD inline D3 normalized(const D3& a) {
    const double n = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return D3{a.x / n, a.y / n, a.z / n};
}
// HomTra3 position / orientation / *_inverse (merlict/HomTra3.cpp:154-216)
MLGPU_HD inline D3 orient(const LeafGeo& g, const D3& o) {
    const double* T = g.rot;
    return D3{o.x * T[0] + o.y * T[1] + o.z * T[2], o.x * T[3] + o.y * T[4] + o.z * T[5],
              o.x * T[6] + o.y * T[7] + o.z * T[8]};
}
MLGPU_HD inline D3 orient_inv(const LeafGeo& g, const D3& o) {
    const double* T = g.rot;
    return D3{o.x * T[0] + o.y * T[3] + o.z * T[6], o.x * T[1] + o.y * T[4] + o.z * T[7],
              o.x * T[2] + o.y * T[5] + o.z * T[8]};
}
MLGPU_HD inline D3 position(const LeafGeo& g, const D3& p) {
    const double* T = g.rot;
    return D3{p.x * T[0] + p.y * T[1] + p.z * T[2] + g.tx, p.x * T[3] + p.y * T[4] + p.z * T[5] + g.ty,
              p.x * T[6] + p.y * T[7] + p.z * T[8] + g.tz};
}
MLGPU_HD inline D3 position_inv(const LeafGeo& g, const D3& p) {
    const double* T = g.rot;
    return D3{p.x * T[0] + p.y * T[3] + p.z * T[6] - g.ti[0], p.x * T[1] + p.y * T[4] + p.z * T[7] - g.ti[1],
              p.x * T[2] + p.y * T[5] + p.z * T[8] - g.ti[2]};
}

// ----------------------------------------------------------------- leaves

// HexagonalPrismZ::is_inside (merlict/scenery/geometry/HexagonalPrismZ.cpp:14-29),
// unit vectors U = (0, 1, 0), V = (s, c, 0), W = (-s, c, 0)
MLGPU_HD inline bool hex_inside(const DeviceScene& sc, double inner, const D3& p) {
    const double u = p.x * 0.0 + p.y * 1.0 + p.z * 0.0;
    const double v = p.x * sc.hex_sin + p.y * sc.hex_cos + p.z * 0.0;
    const double w = p.x * -sc.hex_sin + p.y * sc.hex_cos + p.z * 0.0;
    return u < inner && u > -inner && v < inner && v > -inner && w < inner && w > -inner;
}

MLGPU_HD inline bool cap_bound_inside(const DeviceScene& sc, const LeafGeo& g, const D3& p) {
    if (g.kind == MLGPU_KIND_SPHERE_CAP_HEX) return hex_inside(sc, g.p[1], p);
    if (g.kind == MLGPU_KIND_SPHERE_CAP_CYL) return p.x * p.x + p.y * p.y < g.p[1];
    return fabs(p.x) < g.p[1] && fabs(p.y) < g.p[2];
}

MLGPU_HD inline D3 at(const D3& s, const D3& d, double t) { return D3{s.x + d.x * t, s.y + d.y * t, s.z + d.z * t}; }

// Vec3 != : squared distance > 1e-14 (the drop of self hits)
MLGPU_HD inline bool differs(const D3& a, const D3& b) {
    const D3 e{a.x - b.x, a.y - b.y, a.z - b.z};
    return dot(e, e) > kSelfHit2;
}

// calculate_intersection_with of one leaf for a world ray (s, d), in the
// leaf's object frame (RayAndFrame.cpp:109-116,151-158).  On a hit, t_out is
// the distance; the object-frame hit point is at(position_inv(s),
// orient_inv(d), t_out), bit for bit.
MLGPU_HD inline bool intersect_leaf(const DeviceScene& sc, const LeafGeo* __restrict__ gp, const D3& ws,
                                    const D3& wd, double& t_out) {
    const LeafGeo& g = *gp;
    const D3 s = position_inv(g, ws);
    const D3 d = orient_inv(g, wd);
    const int32_t kind = g.kind;
    double t;
    if (kind <= MLGPU_KIND_SPHERE_CAP_RECT) {
        // SphericalCapRayIntersectionEquation.cpp:6-40 and
        // SurfaceWithOuterPrismBound::add_causeal_intersection (…Bound.cpp:7-70)
        const double R = g.p[0];
        const double d2 = dot(d, d), ss = dot(s, s), sd = dot(s, d);
        const double P = 2.0 * (sd - R * d.z) / d2;
        const double Q = (ss - 2.0 * (R * s.z)) / d2;
        const double p2 = 0.5 * P;  // QuadraticEquation.cpp:12-30
        const double inner = p2 * p2 - Q;
        if (!(inner >= 0.0)) return false;
        const double sq = sqrt(inner);
        const double minus = -p2 - sq, plus = -p2 + sq;
        const D3 pi = at(s, d, plus), mi = at(s, d, minus);
        const bool p_in = cap_bound_inside(sc, g, pi), m_in = cap_bound_inside(sc, g, mi);
        const double rb = g.p[7];
        const bool p_c = plus > 0.0 && dot(pi, pi) <= rb * rb;
        const bool m_c = minus > 0.0 && dot(mi, mi) <= rb * rb;
        bool ok = false;
        t = minus;
        if (p_in && m_in) {
            if (p_c && m_c) { t = (minus >= plus) ? plus : minus; ok = true; }
            else if (p_c) { t = plus; ok = true; }
            else if (m_c) { t = minus; ok = true; }
        } else if (p_in && p_c) { t = plus; ok = true; }
        else if (m_in && m_c) { t = minus; ok = true; }
        if (!ok) return false;
        const D3 p = at(s, d, t);
        if (!differs(s, p)) return false;
    } else if (kind == MLGPU_KIND_CYLINDER) {
        // ZaxisCylinderRayIntersectionEquation.cpp:6-57, Cylinder.cpp:121-150
        const double dx2 = d.x * d.x, dy2 = d.y * d.y;
        const double sxdx = s.x * d.x, sydy = s.y * d.y;
        const double sx2 = s.x * s.x, sy2 = s.y * s.y;
        const double r2 = g.p[0] * g.p[0];
        const double dnom = dx2 + dy2;
        const double P = 2.0 * (sxdx + sydy) / dnom;
        const double Q = (sx2 + sy2 - r2) / dnom;
        const double p2 = 0.5 * P;
        const double inner = p2 * p2 - Q;
        if (!(inner >= 0.0)) return false;
        const double sq = sqrt(inner);
        const double plus = -p2 + sq, minus = -p2 - sq;
        bool causal = false;
        t = 0.0;
        if (minus < 0.0 && plus >= 0.0) { causal = true; t = plus; }
        if (plus >= 0.0 && minus >= 0.0) { causal = true; t = minus; }
        if (!causal) return false;
        const D3 p = at(s, d, t);
        if (!(fabs(p.z) <= g.p[1])) return false;
        if (!differs(s, p)) return false;
    } else if (kind == MLGPU_KIND_ELLIPTICAL_CAP_HEX) {
        // EllipticalCapRayIntersectionEquation.cpp:8-64 (the ellipsoid
        // x^2/A^2 + y^2/B^2 + (z - C)^2/C^2 = 1 through the origin; p0 = C,
        // p2..p4 = 1/A^2, 1/B^2, 1/C^2) and add_causeal_intersection with
        // the hexagonal bound, as for the spherical caps
        const double C = g.p[0], iAA = g.p[2], iBB = g.p[3], iCC = g.p[4];
        const double qa = (d.x * d.x) * iAA + (d.y * d.y) * iBB + (d.z * d.z) * iCC;
        const double qb = 2.0 * ((s.x * d.x) * iAA + (s.y * d.y) * iBB + d.z * (s.z - C) * iCC);
        const double qc = (s.x * s.x) * iAA + (s.y * s.y) * iBB + (s.z * s.z - 2.0 * s.z * C) * iCC;
        const double P = qb / qa, Q = qc / qa;
        const double p2 = 0.5 * P;
        const double inner = p2 * p2 - Q;
        if (!(inner >= 0.0)) return false;
        const double sq = sqrt(inner);
        const double minus = -p2 - sq, plus = -p2 + sq;
        const D3 pi = at(s, d, plus), mi = at(s, d, minus);
        const bool p_in = hex_inside(sc, g.p[1], pi), m_in = hex_inside(sc, g.p[1], mi);
        const double rb = g.p[7];
        const bool p_c = plus > 0.0 && dot(pi, pi) <= rb * rb;
        const bool m_c = minus > 0.0 && dot(mi, mi) <= rb * rb;
        bool ok = false;
        t = minus;
        if (p_in && m_in) {
            if (p_c && m_c) { t = (minus >= plus) ? plus : minus; ok = true; }
            else if (p_c) { t = plus; ok = true; }
            else if (m_c) { t = minus; ok = true; }
        } else if (p_in && p_c) { t = plus; ok = true; }
        else if (m_in && m_c) { t = minus; ok = true; }
        if (!ok) return false;
        const D3 p = at(s, d, t);
        if (!differs(s, p)) return false;
    } else if (kind == MLGPU_KIND_SPHERE) {
        // Sphere.cpp:62-105
        const double sd = dot(s, d), d2 = dot(d, d), ss = dot(s, s);
        const double r2 = g.p[0] * g.p[0];
        const double P = 2 * (sd / d2), Q = ss / d2 - r2;
        const double p2 = 0.5 * P;
        const double inner = p2 * p2 - Q;
        if (!(inner >= 0.0)) return false;
        const double sq = sqrt(inner);
        const double vP = -p2 + sq, vM = -p2 - sq;
        if (vP >= 0.0 && vM >= 0.0) t = vM;
        else if (!(vP < 0.0 && vM < 0.0)) t = vP;
        else return false;
        const D3 p = at(s, d, t);
        if (!differs(s, p)) return false;
    } else {
        // XyPlaneRayIntersectionEquation.cpp:6-41 and the planar primitives
        if (d.z == 0. && (d.x != 0. || d.y != 0.)) return false;
        t = -s.z / d.z;
        if (!(t > 0.0)) return false;
        const D3 p = at(s, d, t);
        bool in;
        switch (kind) {
        case MLGPU_KIND_PLANE: in = fabs(p.x) < g.p[1] && fabs(p.y) < g.p[2]; break;
        case MLGPU_KIND_DISC: in = p.x * p.x + p.y * p.y < g.p[1]; break;
        case MLGPU_KIND_HEX_PLANE: in = hex_inside(sc, g.p[1], p); break;
        case MLGPU_KIND_ANNULUS:
            in = (p.x * p.x + p.y * p.y < g.p[1]) && !(p.x * p.x + p.y * p.y < g.p[2]);
            break;
        case MLGPU_KIND_PLANE_DUAL_SPHERICAL: {
            // DualSphericalPrismZ::is_inside (DualSphericalPrismZ.cpp:21-27)
            const double ax = g.p[1] - p.x, ay = 0.0 - p.y;
            const double bx = -g.p[1] - p.x, by = 0.0 - p.y;
            in = (ax * ax + ay * ay <= g.p[2]) && (bx * bx + by * by <= g.p[2]);
            break;
        }
        case MLGPU_KIND_TRIANGLE: {
            // Triangle::is_inside_triangle (Triangle.cpp:92-102)
            const double Ax = g.p[0], Ay = g.p[1], Bx = g.p[2], By = g.p[3], Cx = g.p[4], Cy = g.p[5];
            const bool bA = ((p.x - Bx) * (Ay - By) - (Ax - Bx) * (p.y - By)) < 0.0;
            const bool bB = ((p.x - Cx) * (By - Cy) - (Bx - Cx) * (p.y - Cy)) < 0.0;
            const bool bC = ((p.x - Ax) * (Cy - Ay) - (Cx - Ax) * (p.y - Ay)) < 0.0;
            in = (bA == bB) && (bB == bC);
            break;
        }
        default: in = false;
        }
        if (!in) return false;
        if (!differs(s, p)) return false;
    }
    t_out = t;
    return true;
}

MLGPU_HD inline D3 leaf_normal(const LeafGeo& g, const D3& p) {
    const int32_t kind = g.kind;
    if (kind <= MLGPU_KIND_SPHERE_CAP_RECT) {
        // SphericalCapRayIntersectionEquation.cpp:52-59
        return normalized(D3{0.0 - p.x, 0.0 - p.y, g.p[0] - p.z});
    }
    if (kind == MLGPU_KIND_CYLINDER) return normalized(D3{p.x, p.y, 0.0});  // Cylinder.cpp:157-160
    if (kind == MLGPU_KIND_SPHERE) return normalized(p);                    // Sphere.cpp:80-81
    if (kind == MLGPU_KIND_ELLIPTICAL_CAP_HEX) {
        // EllipticalCapRayIntersectionEquation.cpp:72-99, as merlict writes
        // it: the factor takes sqrt(1 - x^2/A^2 - y^2/B^2) (not its inverse)
        // and both slopes use 1/A^2
        const double C = g.p[0], iAA = g.p[2], iBB = g.p[3];
        const double f = C * 0.5 * sqrt(1.0 - (p.x * p.x) * iAA - (p.y * p.y) * iBB);
        return normalized(D3{f * (-2.0 * p.x * iAA), f * (-2.0 * p.y * iAA), 1.0});
    }
    return D3{0.0, 0.0, 1.0};
}

// --------------------------------------------------------------- traversal
struct Closest {
    int32_t leaf;  // -1: no hit
    double t;
};

// Traversal stack of one lane: entry k lives at base[k * stride] (an LDS
// column per lane on the device, a local array on the host), room for
// kMaxTreeDepth entries.  An entry is a node index (>= 0) or a leaf deferred
// from node p, child k, kept as -1 - (kDeferredLeaf | (2 p + k)) so that its
// box can be tested again, against the best hit found meanwhile, before the
// FP64 solve.
constexpr int32_t kDeferredLeaf = 0x40000000;

struct LaneStack {
    int32_t* base;
    int stride;
};

// FP32 slab test of box [lo, hi] for the ray o + t d given as
// (1/d, o/d): entered at tmin within [0, limit]?
MLGPU_HD inline bool slab(const float* lo, const float* hi, const float idx, const float idy, const float idz,
                          const float oix, const float oiy, const float oiz, const float limit, float& tmin) {
    const float t1x = fmaf(lo[0], idx, -oix), t2x = fmaf(hi[0], idx, -oix);
    const float t1y = fmaf(lo[1], idy, -oiy), t2y = fmaf(hi[1], idy, -oiy);
    const float t1z = fmaf(lo[2], idz, -oiz), t2z = fmaf(hi[2], idz, -oiz);
    tmin = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fmaxf(fminf(t1z, t2z), 0.0f));
    const float tmax = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fminf(fmaxf(t1z, t2z), limit));
    return tmin <= tmax;
}

// 1 / x within 1 ulp (v_rcp_f32 on the GPU): the box tests only cull, and
// the boxes are inflated by far more than its rounding moves an entry point
MLGPU_HD inline float rcp_approx(const float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// FP32 test of the oriented box of a leaf for the ray o + t d: does the ray
// meet it within [0, limit]?
// Entry j of a leaf box's rotation (LeafBox::m).
MLGPU_HD inline float box_rotation(const LeafBox& b, const int j) {
    const uint32_t bits = (b.m[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
#if defined(__HIP_DEVICE_COMPILE__)
    return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(bits)));
#else
    const int e = static_cast<int>((bits >> 10) & 31u);
    const float mant = static_cast<float>(bits & 1023u);
    const float mag = e == 0 ? std::ldexp(mant, -24) : std::ldexp(1024.0f + mant, e - 25);
    return (bits & 0x8000u) ? -mag : mag;
#endif
}

MLGPU_HD inline bool leaf_box_hit(const LeafBox& b, const float ox, const float oy, const float oz, const float dx,
                                  const float dy, const float dz, const float limit) {
    const float px = ox - b.c[0], py = oy - b.c[1], pz = oz - b.c[2];
    float tmin = 0.0f, tmax = limit;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float r[3] = {box_rotation(b, 3 * k), box_rotation(b, 3 * k + 1), box_rotation(b, 3 * k + 2)};
        const float lo = fmaf(r[0], px, fmaf(r[1], py, r[2] * pz));
        const float ld = fmaf(r[0], dx, fmaf(r[1], dy, r[2] * dz));
        const float inv = rcp_approx(fabsf(ld) > 1e-30f ? ld : copysignf(1e-30f, ld));
        const float t1 = (-b.h[k] - lo) * inv, t2 = (b.h[k] - lo) * inv;
        tmin = fmaxf(tmin, fminf(t1, t2));
        tmax = fminf(tmax, fmaxf(t1, t2));
    }
    return tmin <= tmax;
}

// The float form of a ray for the box tests: origin o (the support, moved
// along the ray to the scope box when it starts far outside), direction,
// reciprocals and o / d.  t_adv is how far the origin was moved.
struct RayF {
    float ox, oy, oz;
    float dx, dy, dz;
    float idx, idy, idz;
    float oix, oiy, oiz;
    double t_adv;
    // byte selectors (perm_bytes) of a node's first child's faces the ray
    // enters (near) and leaves (far) on x, y, z: the lower face where the
    // direction component is positive, the upper one where it is negative
    uint32_t sel_near, sel_far;
};

MLGPU_HD inline void ray_setup(const DeviceScene& sc, const int32_t scope, const D3& s, const D3& d, RayF& r) {
    const Scope root = sc.scopes[scope];
    // direction reciprocals; a zero component becomes tiny so no 0 * inf occurs
    r.dx = static_cast<float>(d.x);
    r.dy = static_cast<float>(d.y);
    r.dz = static_cast<float>(d.z);
    r.idx = 1.0f / (fabsf(r.dx) > 1e-30f ? r.dx : copysignf(1e-30f, r.dx));
    r.idy = 1.0f / (fabsf(r.dy) > 1e-30f ? r.dy : copysignf(1e-30f, r.dy));
    r.idz = 1.0f / (fabsf(r.dz) > 1e-30f ? r.dz : copysignf(1e-30f, r.dz));
    // a support far outside the scope box is advanced (in FP64) to the box so
    // the float origin keeps the precision the box inflation assumes
    double t_adv = 0.0;
    {
        const double ax = (static_cast<double>(root.lo[0]) - s.x) * r.idx, bx = (static_cast<double>(root.hi[0]) - s.x) * r.idx;
        const double ay = (static_cast<double>(root.lo[1]) - s.y) * r.idy, by = (static_cast<double>(root.hi[1]) - s.y) * r.idy;
        const double az = (static_cast<double>(root.lo[2]) - s.z) * r.idz, bz = (static_cast<double>(root.hi[2]) - s.z) * r.idz;
        const double enter = fmax(fmax(fmin(ax, bx), fmin(ay, by)), fmin(az, bz));
        if (enter > 0.0) t_adv = enter * (1.0 - 1e-6) - sc.eps;
        if (t_adv < 0.0) t_adv = 0.0;
    }
    r.t_adv = t_adv;
    r.ox = static_cast<float>(s.x + d.x * t_adv);
    r.oy = static_cast<float>(s.y + d.y * t_adv);
    r.oz = static_cast<float>(s.z + d.z * t_adv);
    r.oix = r.ox * r.idx;
    r.oiy = r.oy * r.idy;
    r.oiz = r.oz * r.idz;
    // face byte 3 f + a of a child: lower (f = 0) or upper (f = 1) face on axis a
    const uint32_t sx = r.idx < 0.0f ? 1u : 0u, sy = r.idy < 0.0f ? 1u : 0u, sz = r.idz < 0.0f ? 1u : 0u;
    r.sel_near = (3u * sx) | (3u * sy + 1u) << 8 | (3u * sz + 2u) << 16;
    r.sel_far = (3u * (1u - sx)) | (3u * (1u - sy) + 1u) << 8 | (3u * (1u - sz) + 2u) << 16;
}

// Bytes of the 8-byte value (hi, lo) picked by sel: byte j of the result is
// byte (sel >> 8 j) & 7 of it (v_perm_b32 for selectors below 8).
MLGPU_HD inline uint32_t perm_bytes(const uint32_t hi, const uint32_t lo, const uint32_t sel) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(hi, lo, sel);
#else
    const uint64_t v = static_cast<uint64_t>(hi) << 32 | lo;
    uint32_t out = 0;
    for (int j = 0; j < 4; j++) out |= static_cast<uint32_t>((v >> (8 * ((sel >> (8 * j)) & 7u))) & 0xFFu) << (8 * j);
    return out;
#endif
}

// Slab test of a node's child boxes on its grid, in the ray's terms: a face
// of grid index k on axis a is crossed at t = k A_a + B_a, A_a = step_a / d_a
// (exact: a power of two times 1/d_a) and B_a = o_a / d_a - o_ray,a / d_a
// (one rounding more than a float face would take, far inside the boxes'
// inflation).  The near and far faces of each axis come from the ray's
// byte selectors, so no per-axis min / max is needed.
struct NodeRay {
    float ax, ay, az, bx, by, bz;
};

MLGPU_HD inline NodeRay node_ray(const PairNode& nd, const RayF& r) {
    NodeRay nr;
    const float ox = __builtin_bit_cast(float, nd.g[0] & 0xFFFFFF00u);
    const float oy = __builtin_bit_cast(float, nd.g[1] & 0xFFFFFF00u);
    const float oz = __builtin_bit_cast(float, nd.g[2] & 0xFFFFFF00u);
    nr.ax = __builtin_bit_cast(float, (nd.g[0] & 0xFFu) << 23) * r.idx;
    nr.ay = __builtin_bit_cast(float, (nd.g[1] & 0xFFu) << 23) * r.idy;
    nr.az = __builtin_bit_cast(float, (nd.g[2] & 0xFFu) << 23) * r.idz;
    nr.bx = fmaf(ox, r.idx, -r.oix);
    nr.by = fmaf(oy, r.idy, -r.oiy);
    nr.bz = fmaf(oz, r.idz, -r.oiz);
    return nr;
}

// near / far: the child's face bytes x, y, z in bytes 0-2 (perm_bytes)
MLGPU_HD inline bool node_child_hit(const NodeRay& nr, const uint32_t near, const uint32_t far, const float limit,
                                    float& tmin) {
    const float nx = fmaf(static_cast<float>(near & 0xFFu), nr.ax, nr.bx);
    const float ny = fmaf(static_cast<float>((near >> 8) & 0xFFu), nr.ay, nr.by);
    const float nz = fmaf(static_cast<float>((near >> 16) & 0xFFu), nr.az, nr.bz);
    const float fx = fmaf(static_cast<float>(far & 0xFFu), nr.ax, nr.bx);
    const float fy = fmaf(static_cast<float>((far >> 8) & 0xFFu), nr.ay, nr.by);
    const float fz = fmaf(static_cast<float>((far >> 16) & 0xFFu), nr.az, nr.bz);
    tmin = fmaxf(fmaxf(nx, ny), fmaxf(nz, 0.0f));
    const float tmax = fminf(fminf(fx, fy), fminf(fz, limit));
    return tmin <= tmax;
}

// Where a traversal stands: the next node or leaf code to visit (kNoChild:
// finished), the number of stack entries and the best hit so far.
struct TraceState {
    int32_t cur;
    int32_t depth;
    Closest best;
};

MLGPU_HD inline void trace_begin(const DeviceScene& sc, const int32_t scope, TraceState& ts) {
    ts.cur = sc.scopes[scope].entry;
    ts.depth = 0;
    ts.best.leaf = -1;
    ts.best.t = HUGE_VAL;
}

// Leaf `leaf` as a candidate: its oriented box is entered before the best
// hit so far and its FP64 solve gives the closest hit; an exact tie goes to
// the earlier candidate in merlict's order.
MLGPU_HD inline void try_leaf(const DeviceScene& sc, const RayF& r, const D3& s, const D3& d, const int32_t leaf,
                              Closest& best, float& best_f) {
    double t;
    if (leaf_box_hit(sc.boxes[leaf], r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, best_f) &&
        intersect_leaf(sc, sc.geo + leaf, s, d, t) &&
        (t < best.t || (t == best.t && sc.geo[leaf].rank < sc.geo[best.leaf].rank))) {
        best.leaf = leaf;
        best.t = t;
        best_f = static_cast<float>(t - r.t_adv) * (1.0f + 1e-6f) + sc.eps;
    }
}

#if defined(__HIPCC__)
// The top of the world scope's tree in LDS, copied there by the kernels that
// set DeviceScene::num_hot (one copy per block).
static __shared__ PairNode hot_nodes[kHotNodes];
#endif

// Node k; with Hot, from the LDS copy of the top of the tree that the
// calling kernel made (stage_hot).  Only the level-0 kernel gains from the
// copy: its traces stay in the top of the tree, while the deeper traces run
// slower with the extra branch in their node fetch.
template <bool Hot>
MLGPU_HD inline PairNode fetch_node(const DeviceScene& sc, const int32_t k) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (Hot && k < sc.num_hot) {
        const PairNode n = hot_nodes[k];
        // keeps this LDS read apart from the global read below: merged, the
        // two become one generic (flat) load through a selected address
        __builtin_amdgcn_sched_barrier(0);
        return n;
    }
#endif
    return sc.pairs[k];
}

// Runs the traversal on, for at most `budget` steps (node visits and leaf
// tests) when Budgeted; returns true when it has finished.  A traversal
// stopped early resumes from ts and the stack, and ends with the same
// closest hit.  The boxes only cull: every leaf whose box the ray enters
// before the best hit so far is solved in FP64, so the closest hit is
// merlict's (rays_first_intersection_with_frame, RayAndFrame.cpp:118-173);
// leaves at exactly the same distance (coplanar surfaces) are ordered by
// their rank in merlict's candidate order, so the winner does not depend on
// the traversal order.
// Nearer children are visited first so that the best hit shrinks early and
// prunes more boxes.
template <bool Budgeted, bool Hot = false>
MLGPU_HD inline bool trace_run(const DeviceScene& sc, const RayF& r, const D3& s, const D3& d, TraceState& ts,
                               const LaneStack stk, int32_t budget) {
    // best distance relative to the float origin, plus margin
    float best_f = ts.best.leaf >= 0 ? static_cast<float>(ts.best.t - r.t_adv) * (1.0f + 1e-6f) + sc.eps : HUGE_VALF;
    int32_t cur = ts.cur;
    int depth = ts.depth;
    if (cur == kNoChild) return true;
    for (;;) {
        if (Budgeted && budget-- <= 0) {
            ts.cur = cur;
            ts.depth = depth;
            return false;
        }
        if (cur >= 0) {
            const PairNode nd = fetch_node<Hot>(sc, cur);
            const NodeRay nr = node_ray(nd, r);
            float n0, n1;
            const bool h0 = node_child_hit(nr, perm_bytes(nd.q[1], nd.q[0], r.sel_near),
                                           perm_bytes(nd.q[1], nd.q[0], r.sel_far), best_f, n0) &&
                            nd.c0 != kNoChild;
            const bool h1 = node_child_hit(nr, perm_bytes(nd.q[2], nd.q[1], r.sel_near + 0x020202u),
                                           perm_bytes(nd.q[2], nd.q[1], r.sel_far + 0x020202u), best_f, n1) &&
                            nd.c1 != kNoChild;
            if (h0 && h1) {
                const bool near0 = n0 <= n1;
                const int32_t far = near0 ? nd.c1 : nd.c0;
                stk.base[depth * stk.stride] = far >= 0 ? far : -1 - (kDeferredLeaf | (2 * cur + (near0 ? 1 : 0)));
                depth++;
                cur = near0 ? nd.c0 : nd.c1;
                continue;
            }
            if (h0 || h1) {
                cur = h0 ? nd.c0 : nd.c1;
                continue;
            }
        } else {
            try_leaf(sc, r, s, d, -1 - cur, ts.best, best_f);
        }
        // next stack entry still in front of the best hit; a deferred leaf's
        // box is tested again first
        cur = kNoChild;
        while (depth > 0) {
            depth--;
            const int32_t e = stk.base[depth * stk.stride];
            if (e >= 0) {
                cur = e;
                break;
            }
            const int32_t p = (-1 - e) & (kDeferredLeaf - 1);
            const int32_t node = p >> 1;
            const PairNode nd = fetch_node<Hot>(sc, node);
            const bool second = (p & 1) != 0;
            const NodeRay nr = node_ray(nd, r);
            const uint32_t w_hi = second ? nd.q[2] : nd.q[1], w_lo = second ? nd.q[1] : nd.q[0];
            const uint32_t shift = second ? 0x020202u : 0u;
            float tn;
            if (node_child_hit(nr, perm_bytes(w_hi, w_lo, r.sel_near + shift), perm_bytes(w_hi, w_lo, r.sel_far + shift),
                               best_f, tn)) {
                cur = second ? nd.c1 : nd.c0;
                break;
            }
        }
        if (cur == kNoChild) break;
    }
    ts.cur = kNoChild;
    ts.depth = 0;
    return true;
}

// First intersection within scope `scope`.
template <bool Hot = false>
MLGPU_HD inline void first_hit(const DeviceScene& sc, const int32_t scope, const D3& s, const D3& d, Closest& best,
                               const LaneStack stk) {
    RayF r;
    ray_setup(sc, scope, s, d, r);
    TraceState ts;
    trace_begin(sc, scope, ts);
    trace_run<false, Hot>(sc, r, s, d, ts, stk, 0);
    best = ts.best;
}

// ------------------------------------------------------------- propagator
// State of one photon between bounces (merlict's Photon reduced to what the
// remaining propagation and the final record need).
struct PhotonState {
    D3 s, d;                 // ray support and unit direction, world frame
    double tof;              // running Photon::time_of_flight
    double dist;             // running accumulated distance
    int32_t root;            // scope of env.root_frame
    uint32_t n_hist;         // history length including production
    uint8_t err;
    uint8_t last_interaction;
    int32_t last_frame, last_sensor;
    // the result's x, y, cx, cy, rounded to float when the bounce computes
    // them (the record stores floats)
    float out_x, out_y, out_cx, out_cy;
};

MLGPU_HD inline void photon_init(PhotonState& st, const mlgpu_photons& in, const int64_t i, const int32_t world_scope) {
    st.s = D3{in.support_x[i], in.support_y[i], in.support_z[i]};
    st.d = normalized(D3{in.direction_x[i], in.direction_y[i], in.direction_z[i]});  // Ray.cpp:16-20
    st.tof = 0.0;
    st.dist = 0.0;
    st.root = world_scope;
    st.n_hist = 1;  // production
    st.err = in.wavelength[i] > 0.0 ? MLGPU_PHOTON_OK : MLGPU_PHOTON_BAD_WAVELENGTH;  // Photon.cpp:20-28
    st.last_interaction = MLGPU_PRODUCTION;
    st.last_frame = -1;
    st.last_sensor = -1;
    st.out_x = static_cast<float>(st.s.x);
    st.out_y = static_cast<float>(st.s.y);
    st.out_cx = 0.0f;
    st.out_cy = 0.0f;
}

// The history entry RayForPropagation::push_back_intersection_and_interaction
// appends (RayForPropagation.cpp): the Intersection's constructor arguments
// (Intersection.cpp:14-27) and the interaction.
MLGPU_HD inline void record_bounce(mlgpu_bounce* const hist, const uint32_t k, const D3& position, const D3& normal,
                                   const double distance, const int32_t frame, const uint8_t interaction,
                                   const bool from_outside) {
    mlgpu_bounce& b = hist[k];
    b.position[0] = position.x;
    b.position[1] = position.y;
    b.position[2] = position.z;
    b.normal[0] = normal.x;
    b.normal[1] = normal.y;
    b.normal[2] = normal.z;
    b.distance = distance;
    b.frame = frame;
    b.interaction = interaction;
    b.from_outside = from_outside ? 1 : 0;
    b.reserved[0] = 0;
    b.reserved[1] = 0;
}

// One pass of Propagator::work_on_first_causal_intersection and what follows
// (merlict/PhotonAndFrame.cpp:23-124) for the closest hit (leaf, t) of the
// current ray (leaf < 0: none).  Returns true when the photon propagates on.
// hist (optional): the photon's history entries after the production; the
// bounce is written to hist[n_hist - 1] when it is pushed.
MLGPU_HD inline bool photon_bounce(const DeviceScene& sc, PhotonState& st, const int32_t leaf, const double t,
                                   const double lambda, Rng& rng, const uint32_t max_interactions,
                                   mlgpu_bounce* const hist = nullptr) {
    if (leaf < 0) {
        // no hit: absorbed in void at 1e4 m, n = 1 on that leg
        double nv;
        if (!func_eval(sc, sc.default_refraction, lambda, nv)) { st.err = MLGPU_PHOTON_OUT_OF_TABLE; return false; }
        st.tof += nv * kVoidDistance / kC;
        st.dist += kVoidDistance;
        const D3 e = at(st.s, st.d, kVoidDistance);
        // the void intersection of CausalIntersection::calculate_closest_intersection
        // (RayAndFrame.cpp:160-167): world point at 1e4 m, normal = direction
        if (hist) record_bounce(hist, st.n_hist - 1, e, st.d, kVoidDistance, -1, MLGPU_ABSORPTION_IN_VOID, false);
        st.n_hist++;
        st.last_interaction = MLGPU_ABSORPTION_IN_VOID;
        st.last_frame = -1;
        st.last_sensor = -1;
        st.out_x = static_cast<float>(e.x);
        st.out_y = static_cast<float>(e.y);
        const D3 inc = normalized(D3{e.x - st.s.x, e.y - st.s.y, e.z - st.s.z});
        st.out_cx = static_cast<float>(-1.0 * inc.x);
        st.out_cy = static_cast<float>(-1.0 * inc.y);
        return false;
    }
    const LeafGeo& g = sc.geo[leaf];
    const LeafPhys ph = sc.phys[leaf];
    // the object-frame ray and hit, as the leaf test computed them
    const D3 ls = position_inv(g, st.s);
    const D3 ld = orient_inv(g, st.d);
    const D3 hp = at(ls, ld, t);
    const D3 n = leaf_normal(g, hp);
    // Intersection::ray_runs_from_outside_to_inside (Intersection.cpp:133-139)
    const bool outside_in = dot(n, ld) < 0.0;
    const D3 hit_world = position(g, hp);

    uint8_t interaction;
    bool stop = true;
    D3 nd = st.d;
    int32_t new_root = st.root;
    // the functions a bounce can need, looked up in two pairs (absorption and
    // the index it comes from; reflection and the index it would enter); a
    // lookup outside its table is an error only where merlict evaluates it
    double fv[2];
    bool fok[2];
    {
        const int32_t fns[2] = {ph.surf[outside_in ? MLGPU_OUTER_ABSORPTION : MLGPU_INNER_ABSORPTION],
                                ph.surf[outside_in ? MLGPU_OUTER_REFRACTION : MLGPU_INNER_REFRACTION]};
        func_eval_n(sc, fns, lambda, fv, fok);
    }
    // absorbed_in_medium_before_reaching_surface (PhotonAndFrame.cpp:32-38)
    if (!fok[0]) {
        st.err = MLGPU_PHOTON_OUT_OF_TABLE;
        return false;
    }
    const double ell = fv[0];
    // the refractive index on the side the photon comes from: Fresnel's n1
    // and the index of the leg's time of flight (Photon.cpp:44-56)
    if (!fok[1]) {
        st.err = MLGPU_PHOTON_OUT_OF_TABLE;
        return false;
    }
    const double n_from = fv[1];
    // an infinite absorption length (merlict's default) survives exactly:
    // exp(-t / inf) = exp(-0) = 1
    const double survival = ell == HUGE_VAL ? 1.0 : exp(-t / ell);
    if (rng.uniform() > survival) {
        interaction = MLGPU_ABSORPTION_IN_VOID;
    } else {
        // interact_with_object (PhotonAndFrame.cpp:40-48)
        double gv[2];
        bool gok[2];
        {
            const int32_t fns[2] = {ph.surf[outside_in ? MLGPU_OUTER_REFLECTION : MLGPU_INNER_REFLECTION],
                                    ph.surf[outside_in ? MLGPU_INNER_REFRACTION : MLGPU_OUTER_REFRACTION]};
            func_eval_n(sc, fns, lambda, gv, gok);
        }
        if (!gok[0]) {
            st.err = MLGPU_PHOTON_OUT_OF_TABLE;
            return false;
        }
        const double refl = gv[0];
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
                // fresnel_refraction_and_reflection (PhotonAndFrame.cpp:75-87), Fresnel.cpp:7-38
                if (!gok[1]) {
                    st.err = MLGPU_PHOTON_OUT_OF_TABLE;
                    return false;
                }
                const double n_to = gv[1];
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
                    interaction = MLGPU_FRESNEL_REFLECTION_ON_SURFACE;
                    mirror = true;
                } else {
                    // pass_the_boundary_layer / propagate_on_after_boundary_layer (:89-118)
                    interaction = outside_in ? MLGPU_REFRACTION_TO_INSIDE : MLGPU_REFRACTION_TO_OUTSIDE;
                    const bool to_default = (outside_in ? ph.surf[MLGPU_INNER_REFRACTION]
                                                        : ph.surf[MLGPU_OUTER_REFRACTION]) == sc.default_refraction;
                    new_root = (ph.allowed_scope >= 0 && !to_default) ? ph.allowed_scope : ph.root_scope;
                    const double k = eta * cosI - cosT;
                    const D3 refr{ld.x * eta + nrm.x * k, ld.y * eta + nrm.y * k, ld.z * eta + nrm.z * k};
                    nd = normalized(orient(g, refr));
                    stop = false;
                }
            }
        }
        if (mirror) {
            // Intersection::reflection_direction_in_root_frame with the mirror
            // matrix I - 2 n n^T (Intersection.cpp:74-79, Vec3.cpp:59-113)
            const double x = n.x, y = n.y, z = n.z;
            const D3 m{(1. - 2. * x * x) * ld.x + -2. * x * y * ld.y + -2. * x * z * ld.z,
                       -2. * x * y * ld.x + (1. - 2. * y * y) * ld.y + -2. * y * z * ld.z,
                       -2. * x * z * ld.x + -2. * y * z * ld.y + (1. - 2. * z * z) * ld.z};
            nd = normalized(orient(g, m));
            stop = false;
        }
    }
    // push_back_intersection_and_interaction; Photon::time_of_flight (Photon.cpp:44-56)
    st.tof += n_from * t / kC;
    st.dist += t;
    if (hist) record_bounce(hist, st.n_hist - 1, hp, n, t, ph.frame, interaction, outside_in);
    st.n_hist++;
    st.last_interaction = interaction;
    st.last_frame = ph.frame;
    st.last_sensor = ph.sensor;
    // Propagator::propagate (PhotonAndFrame.cpp:13-21)
    const bool go = !stop && st.n_hist <= max_interactions;
    if (!go) {
        // the final record, from the photon's last bounce (a bounce it goes
        // on from is overwritten by the next one)
        st.out_x = static_cast<float>(hp.x);
        st.out_y = static_cast<float>(hp.y);
        // final_intersection_incident_direction_wrt_frame (RayForPropagation.cpp:114-136)
        const D3 inc = normalized(D3{hit_world.x - st.s.x, hit_world.y - st.s.y, hit_world.z - st.s.z});
        const D3 loc = orient_inv(g, inc);
        st.out_cx = static_cast<float>(-1.0 * loc.x);
        st.out_cy = static_cast<float>(-1.0 * loc.y);
    }
    if (stop) return false;
    st.s = hit_world;
    st.d = nd;
    st.root = new_root;
    return go;
}

MLGPU_HD inline void photon_finish(const PhotonState& st, const mlgpu_results& out, const int64_t i) {
    if (st.err != MLGPU_PHOTON_OK) {
        // merlict aborts the batch with an exception here; the record of an
        // erroneous photon is its error code with a neutral final state
        out.final_frame[i] = -1;
        out.sensor_id[i] = -1;
        out.final_interaction[i] = MLGPU_PRODUCTION;
        out.num_interactions[i] = 0;
        out.error[i] = st.err;
        out.time_of_flight[i] = 0.0;
        out.x[i] = 0.0f;
        out.y[i] = 0.0f;
        out.cx[i] = 0.0f;
        out.cy[i] = 0.0f;
        if (out.accumulated_distance) out.accumulated_distance[i] = 0.0;
        return;
    }
    out.final_frame[i] = st.last_frame;
    out.sensor_id[i] = st.last_sensor;
    out.final_interaction[i] = st.last_interaction;
    out.num_interactions[i] = static_cast<uint8_t>(st.n_hist);
    out.error[i] = st.err;
    out.time_of_flight[i] = st.tof;
    out.x[i] = st.out_x;
    out.y[i] = st.out_y;
    out.cx[i] = st.out_cx;
    out.cy[i] = st.out_cy;
    if (out.accumulated_distance) out.accumulated_distance[i] = st.dist;
}

// Propagator (merlict/PhotonAndFrame.cpp:6-124) for photon i of the batch,
// all bounces in one lane, on the Philox stream of photon id `id`.
MLGPU_HD inline void propagate_photon(const DeviceScene& sc, const int32_t world_scope, const mlgpu_photons& in,
                                      const int64_t i, const uint32_t max_interactions, const uint64_t seed,
                                      const uint64_t id, const mlgpu_results& out, const LaneStack stk) {
    PhotonState st;
    photon_init(st, in, i, world_scope);
    const double lambda = in.wavelength[i];
    Rng rng;
    rng.init(seed, id, 0u);
    bool go = st.err == MLGPU_PHOTON_OK && st.n_hist <= max_interactions;
    while (go) {
        Closest c;
        first_hit(sc, st.root, st.s, st.d, c, stk);
        go = photon_bounce(sc, st, c.leaf, c.t, lambda, rng, max_interactions);
    }
    photon_finish(st, out, i);
}

// propagate_photon that also writes the photon's history: its bounces to
// bounces[0, num_bounces), the count, and the last ray (support, direction;
// RayForPropagation keeps the ray of the last leg) to final_ray[6] if given.
MLGPU_HD inline void propagate_photon_history(const DeviceScene& sc, const int32_t world_scope,
                                              const mlgpu_photons& in, const int64_t i,
                                              const uint32_t max_interactions, const uint64_t seed,
                                              const uint64_t id, const mlgpu_results& out, const LaneStack stk,
                                              mlgpu_bounce* const bounces, int32_t* const num_bounces,
                                              double* const final_ray) {
    PhotonState st;
    photon_init(st, in, i, world_scope);
    const double lambda = in.wavelength[i];
    Rng rng;
    rng.init(seed, id, 0u);
    bool go = st.err == MLGPU_PHOTON_OK && st.n_hist <= max_interactions;
    while (go) {
        Closest c;
        first_hit(sc, st.root, st.s, st.d, c, stk);
        go = photon_bounce(sc, st, c.leaf, c.t, lambda, rng, max_interactions, bounces);
    }
    photon_finish(st, out, i);
    *num_bounces = static_cast<int32_t>(st.n_hist) - 1;
    if (final_ray) {
        final_ray[0] = st.s.x;
        final_ray[1] = st.s.y;
        final_ray[2] = st.s.z;
        final_ray[3] = st.d.x;
        final_ray[4] = st.d.y;
        final_ray[5] = st.d.z;
    }
}

}  // namespace mlgpu
