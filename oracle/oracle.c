/*
 * oracle.c -- FP64 CPU restatement of merlict's photon propagation.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Every function names the
 * reference file:line it restates.  Arithmetic is written in the reference's
 * operation order and compiled without FMA contraction, so on the same libm
 * it reproduces the reference's doubles.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static __thread char g_err[512];
static void set_err(const char* msg) { snprintf(g_err, sizeof(g_err), "%s", msg); }
const char* oracle_last_error(void) { return g_err; }

/* ------------------------------------------------------------ Vec3 math */
typedef struct { double x, y, z; } v3;

static v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
/* Vec3::operator*(v) = v.x*x + v.y*y + v.z*z  (merlict/Vec3.cpp:31-33) */
static double dot(v3 a, v3 b) { return b.x * a.x + b.y * a.y + b.z * a.z; }
static v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 mul(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
static v3 dvd(v3 a, double s) { return V(a.x / s, a.y / s, a.z / s); }
static double norm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static v3 normalized(v3 a) { return dvd(a, norm(a)); }
/* Vec3 != : squared distance > 1e-14 (merlict/Vec3.cpp:121-127) */
static int vec_neq(v3 a, v3 b) { v3 d = sub(a, b); return dot(d, d) > 1e-14; }
/* Vec3::norm_is_less_equal_than (Vec3.cpp:146-149) */
static int norm_le(v3 a, double r) { return dot(a, a) <= r * r; }
/* Vec3::octant (Vec3.cpp:151-166) */
static int octant(v3 a) { return 4 * (a.x >= 0.) + 2 * (a.y >= 0.) + 1 * (a.z >= 0.); }

/* ---------------------------------------------------------------- HomTra3 */
typedef struct { double T[3][4]; } homtra;

static homtra ht_unit(void) {
    homtra h = {{{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}}};
    return h;
}

/* HomTra3::operator* (merlict/HomTra3.cpp:245-268) */
static homtra ht_mul(const homtra* A, const homtra* G) {
    homtra r;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            r.T[i][j] = A->T[i][0] * G->T[0][j] + A->T[i][1] * G->T[1][j] + A->T[i][2] * G->T[2][j];
        r.T[i][3] = A->T[i][0] * G->T[0][3] + A->T[i][1] * G->T[1][3] + A->T[i][2] * G->T[2][3] + A->T[i][3];
    }
    return r;
}

/* Rot3 semantics (merlict/Rot3.cpp:10-33) + HomTra3 rotation components
 * (HomTra3.cpp:67-112) + set_transformation = Translation*Rotation (:43-54) */
static homtra ht_from_pos_rot(v3 pos, int rot_mode, const double* rot) {
    homtra R = ht_unit();
    int xyz = (rot_mode == MLGPU_ROT_XYZ);
    double rx = rot[0], ry = rot[1], rz = rot[2];
    v3 axis = V(rot[0], rot[1], rot[2]);
    double angle = rot[3];
    if (!xyz && angle == 0.0) { xyz = 1; rx = 0.0; ry = 0.0; rz = 0.0; }
    if (xyz) {
        const double cosRx = cos(rx), cosRy = cos(ry), cosRz = cos(rz);
        const double sinRx = sin(rx), sinRy = sin(ry), sinRz = sin(rz);
        R.T[0][0] = cosRy * cosRz;
        R.T[0][1] = cosRx * sinRz + sinRx * sinRy * cosRz;
        R.T[0][2] = sinRx * sinRz - cosRx * sinRy * cosRz;
        R.T[1][0] = -cosRy * sinRz;
        R.T[1][1] = cosRx * cosRz - sinRx * sinRy * sinRz;
        R.T[1][2] = sinRx * cosRz + cosRx * sinRy * sinRz;
        R.T[2][0] = sinRy;
        R.T[2][1] = -sinRx * cosRy;
        R.T[2][2] = cosRx * cosRy;
    } else {
        if (rot_mode != MLGPU_ROT_AXIS_STORED)
            axis = dvd(axis, norm(axis)); /* Rot3::set normalises          */
        axis = dvd(axis, norm(axis));     /* set_rotation_..._rot_axis again */
        const double x = axis.x, y = axis.y, z = axis.z;
        const double sinR = sin(angle), cosR = cos(angle);
        R.T[0][0] = cosR + x * x * (1. - cosR);
        R.T[0][1] = x * y * (1. - cosR) - z * sinR;
        R.T[0][2] = x * z * (1. - cosR) + y * sinR;
        R.T[1][0] = y * x * (1. - cosR) + z * sinR;
        R.T[1][1] = cosR + y * y * (1. - cosR);
        R.T[1][2] = y * z * (1. - cosR) - x * sinR;
        R.T[2][0] = z * x * (1. - cosR) - y * sinR;
        R.T[2][1] = z * y * (1. - cosR) + x * sinR;
        R.T[2][2] = cosR + z * z * (1. - cosR);
    }
    homtra Tr = ht_unit();
    Tr.T[0][3] = pos.x; Tr.T[1][3] = pos.y; Tr.T[2][3] = pos.z;
    return ht_mul(&Tr, &R);
}

/* HomTra3::position / orientation / *_inverse (HomTra3.cpp:154-216) */
static v3 ht_orient(const homtra* h, v3 o) {
    const double (*T)[4] = h->T;
    return V(o.x * T[0][0] + o.y * T[0][1] + o.z * T[0][2],
             o.x * T[1][0] + o.y * T[1][1] + o.z * T[1][2],
             o.x * T[2][0] + o.y * T[2][1] + o.z * T[2][2]);
}
static v3 ht_orient_inv(const homtra* h, v3 o) {
    const double (*T)[4] = h->T;
    return V(o.x * T[0][0] + o.y * T[1][0] + o.z * T[2][0],
             o.x * T[0][1] + o.y * T[1][1] + o.z * T[2][1],
             o.x * T[0][2] + o.y * T[1][2] + o.z * T[2][2]);
}
static v3 ht_pos(const homtra* h, v3 p) {
    const double (*T)[4] = h->T;
    return V(p.x * T[0][0] + p.y * T[0][1] + p.z * T[0][2] + T[0][3],
             p.x * T[1][0] + p.y * T[1][1] + p.z * T[1][2] + T[1][3],
             p.x * T[2][0] + p.y * T[2][1] + p.z * T[2][2] + T[2][3]);
}
static v3 ht_pos_inv(const homtra* h, v3 p) {
    const double (*T)[4] = h->T;
    return V(p.x * T[0][0] + p.y * T[1][0] + p.z * T[2][0] - (T[0][0] * T[0][3] + T[1][0] * T[1][3] + T[2][0] * T[2][3]),
             p.x * T[0][1] + p.y * T[1][1] + p.z * T[2][1] - (T[0][1] * T[0][3] + T[1][1] * T[1][3] + T[2][1] * T[2][3]),
             p.x * T[0][2] + p.y * T[1][2] + p.z * T[2][2] - (T[0][2] * T[0][3] + T[1][2] * T[1][3] + T[2][2] * T[2][3]));
}

/* ------------------------------------------------------------------ Func1 */
typedef struct { double x, y, slope; } fpoint;
typedef struct { fpoint* p; int n; double lower, upper; } func1;

static int cmp_fpoint(const void* a, const void* b) {
    double xa = ((const fpoint*)a)->x, xb = ((const fpoint*)b)->x;
    return (xa < xb) ? -1 : (xa > xb) ? 1 : 0;
}

/* Func1::init: slopes from table order, then sort (Func1.cpp:22-50) */
static int func1_init(func1* f, const double* x, const double* y, int n) {
    if (n < 1) { set_err("Func1: empty table"); return 1; }
    f->p = (fpoint*)malloc(sizeof(fpoint) * n);
    f->n = n;
    for (int i = 0; i < n; i++) {
        f->p[i].x = x[i];
        f->p[i].y = y[i];
        f->p[i].slope = (i != n - 1) ? (y[i] - y[i + 1]) / (x[i] - x[i + 1]) : 0.0;
    }
    qsort(f->p, n, sizeof(fpoint), cmp_fpoint);
    for (int i = 0; i + 1 < n; i++)
        if (f->p[i].x == f->p[i + 1].x) { set_err("Func1: duplicate argument"); return 1; }
    f->lower = f->p[0].x;
    f->upper = f->p[n - 1].x;
    return 0;
}

/* the work counts of the calling worker thread (NULL: not counted) */
static __thread oracle_work* g_work = NULL;
static __thread oracle_work g_last_work;

/* Func1::evaluate (Func1.cpp:66-71,118-130), Limits::assert_contains */
static int func1_eval(const func1* f, double x, double* out) {
    if (g_work) {
        int c = 0;
        while ((1 << c) < f->n) c++;
        g_work->func_lookups++;
        g_work->func_compares += c;
    }
    if (x < f->lower || x >= f->upper) return 1;
    /* std::upper_bound: first point with P.x > x */
    int lo = 0, hi = f->n;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (f->p[mid].x > x) hi = mid; else lo = mid + 1;
    }
    if (lo == 0 || (lo == f->n && x > f->p[f->n - 1].x)) return 1;
    const fpoint p0 = f->p[lo - 1];
    *out = p0.slope * x + p0.y - p0.slope * p0.x;
    return 0;
}

/* ------------------------------------------------------------------ scene */
typedef struct {
    int mother;
    int* children;
    int n_children, cap_children;
    v3 pos;
    int rot_mode;
    double rot[4];
    homtra t2m, t2w;
    double radius;
    int kind;
    double p[MLGPU_NUM_PARAMS];
    int surf[MLGPU_NUM_SURFACE_SLOTS];  /* resolved function index */
    int allowed;
    int sensor;
    int root;
    /* primitive constants */
    double hex_inner;       /* HexagonalPrismZ inner radius          */
    double cyl_r2;          /* CylinderPrismZ radius_square          */
    double cyl_r2_inner;
    double half_x, half_y;  /* RectangularPrismZ                     */
    double ds_dist, ds_r;   /* DualSphericalPrismZ                   */
    double ell_C, ell_iAA, ell_iBB, ell_iCC;  /* EllipticalCapWithHexagonalBound */
} oframe;

struct oracle_scene {
    oframe* f;
    int n, cap;
    int n_desc;
    func1* funcs;
    int n_funcs;        /* desc functions; defaults follow */
    int def_refl, def_refr, def_abs;
};

static const double UNIT_V_X_INIT = 0.0; /* placeholder for static init */
static v3 UNIT_U, UNIT_V, UNIT_W;
static int g_units_ready = 0;
static void init_hex_units(void) {
    if (g_units_ready) return;
    /* HexagonalPrismZ.h:10-16 */
    const double s = sin(2. / 3. * M_PI), c = cos(2. / 3. * M_PI);
    UNIT_U = V(0., 1., 0.);
    UNIT_V = add(mul(V(1., 0., 0.), +s), mul(V(0., 1., 0.), c));
    UNIT_W = add(mul(V(1., 0., 0.), -s), mul(V(0., 1., 0.), c));
    (void)UNIT_V_X_INIT;
    g_units_ready = 1;
}

static int push_child(oframe* m, int c) {
    if (m->n_children == m->cap_children) {
        m->cap_children = m->cap_children ? 2 * m->cap_children : 4;
        m->children = (int*)realloc(m->children, sizeof(int) * m->cap_children);
    }
    m->children[m->n_children++] = c;
    return 0;
}

static int new_frame(oracle_scene* s) {
    if (s->n == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 64;
        s->f = (oframe*)realloc(s->f, sizeof(oframe) * s->cap);
    }
    memset(&s->f[s->n], 0, sizeof(oframe));
    s->f[s->n].allowed = -1;
    s->f[s->n].sensor = -1;
    s->f[s->n].mother = -1;
    return s->n++;
}

/* bounding radius a primitive sets on construction
 * (post_initialize_radius_of_enclosing_sphere of each primitive). */
static void primitive_init(oframe* f) {
    const double* p = f->p;
    switch (f->kind) {
    case MLGPU_KIND_SPHERE_CAP_HEX: {     /* SphereCapWithHexagonalBound.cpp:8-43 */
        double R = p[0], r = p[1];
        double r_member = r > R ? R : r;
        f->hex_inner = r * cos(M_PI / 6.0);  /* hexBounds gets the parameter */
        double z = sqrt(R * R - r_member * r_member);
        f->radius = hypot(r_member, R - z);
        break;
    }
    case MLGPU_KIND_SPHERE_CAP_CYL: {     /* SphereCapWithCylinderBound.cpp:8-38 */
        double R = p[0], r = p[1];
        double r_member = r > R ? R : r;
        f->cyl_r2 = r * r;
        double z = sqrt(R * R - r_member * r_member);
        f->radius = hypot(r_member, R - z);
        break;
    }
    case MLGPU_KIND_SPHERE_CAP_RECT:      /* SphereCapWithRectangularBound.cpp:8-30 */
        f->half_x = 0.5 * p[1];
        f->half_y = 0.5 * p[2];
        f->radius = hypot(f->half_x, f->half_y);
        break;
    case MLGPU_KIND_PLANE:                /* Plane.cpp:13-23 */
        f->half_x = 0.5 * p[0];
        f->half_y = 0.5 * p[1];
        f->radius = hypot(f->half_x, f->half_y);
        break;
    case MLGPU_KIND_DISC:                 /* Disc.cpp:8-15, CylinderPrismZ.cpp */
        f->cyl_r2 = p[0] * p[0];
        f->radius = sqrt(f->cyl_r2);
        break;
    case MLGPU_KIND_HEX_PLANE:            /* HexPlane.cpp:9-16 */
        f->hex_inner = p[0] * cos(M_PI / 6.0);
        f->radius = p[0];
        break;
    case MLGPU_KIND_ANNULUS:              /* Annulus.cpp:8-19 */
        f->cyl_r2 = p[0] * p[0];
        f->cyl_r2_inner = p[1] * p[1];
        f->radius = sqrt(f->cyl_r2);
        break;
    case MLGPU_KIND_PLANE_DUAL_SPHERICAL: { /* DualSphericalPrismZ.cpp:7-19 */
        double h = p[0], w = p[1];
        f->ds_dist = (w * w - h * h) / (2.0 * h);
        f->ds_r = f->ds_dist / 2.0 + h / 2.0;
        f->radius = w * 0.5;
        break;
    }
    case MLGPU_KIND_CYLINDER:             /* Cylinder.cpp:104-107 */
        f->radius = hypot(0.5 * p[1], p[0]);
        break;
    case MLGPU_KIND_TRIANGLE: {           /* Triangle.cpp:72-80 */
        double a = sqrt(p[0] * p[0] + p[1] * p[1]);
        double b = sqrt(p[2] * p[2] + p[3] * p[3]);
        double c = sqrt(p[4] * p[4] + p[5] * p[5]);
        double m = a;
        if (b > m) m = b;
        if (c > m) m = c;
        f->radius = m;
        break;
    }
    case MLGPU_KIND_SPHERE:               /* Sphere.cpp:13-21 */
        f->radius = p[0];
        break;
    case MLGPU_KIND_ELLIPTICAL_CAP_HEX: { /* EllipticalCapWithHexagonalBound.cpp:10-52 */
        /* p: focal length, x/y curvature ratio, hex rotation (unused), outer hex radius */
        const double mean_curvature_radius = p[0] * 0.5;
        const double X = mean_curvature_radius * p[1];
        const double Y = mean_curvature_radius / p[1];
        const double Z = 0.5 * (X + Y);
        double outer = p[3];
        f->ell_C = Z;
        f->ell_iAA = 1.0 / (X * X);
        f->ell_iBB = 1.0 / (Y * Y);
        f->ell_iCC = 1.0 / (Z * Z);
        if (outer > Z) outer = Z;             /* restricted before hexBounds takes it */
        f->hex_inner = outer * cos(M_PI / 6.0);
        f->radius = hypot(outer, Z - sqrt(Z * Z - outer * outer));
        break;
    }
    default:
        f->radius = 0.0;
    }
}

/* bound::position_mean_in_mother (Frames.cpp:92-108) */
static v3 position_mean(const oracle_scene* s, const int* idx, int n) {
    v3 sum = V(0., 0., 0.);
    for (int i = 0; i < n; i++) sum = add(sum, s->f[idx[i]].pos);
    return dvd(sum, (double)n);
}

/* bound::bounding_sphere_radius (Frames.cpp:137-206) */
static double bounding_radius(const oracle_scene* s, const int* idx, int n, v3 center) {
    double radius = 0.0;
    for (int i = 0; i < n; i++) {
        const oframe* c = &s->f[idx[i]];
        double need = norm(sub(center, c->pos)) + c->radius;
        if (need > radius) radius = need;
    }
    return radius;
}

/* bound::bounding_sphere_center_alternative (Frames.cpp:57-90) */
static v3 center_alternative(const oracle_scene* s, const int* idx, int n) {
    if (n == 1) return s->f[idx[0]].pos;
    int start = -1, end = -1;
    double maximum = 0.0;
    for (int i = 0; i < n; i++) {
        for (int j = i; j < n; j++) {
            if (i != j) {
                const oframe* a = &s->f[idx[i]];
                const oframe* b = &s->f[idx[j]];
                double dist = norm(sub(a->pos, b->pos)) + a->radius + b->radius;
                if (dist > maximum) { maximum = dist; start = idx[i]; end = idx[j]; }
            }
        }
    }
    /* both VOID_FRAME (pos 0, radius 0) when all distances are 0 */
    v3 sp = start >= 0 ? s->f[start].pos : V(0., 0., 0.);
    v3 ep = end >= 0 ? s->f[end].pos : V(0., 0., 0.);
    double sr = start >= 0 ? s->f[start].radius : 0.0;
    v3 dir = sub(ep, sp);
    dir = dvd(dir, norm(dir));
    return add(sp, mul(dir, 0.5 * maximum - sr));
}

/* bound::bounding_sphere_center (Frames.cpp:39-55) */
static v3 bounding_center(const oracle_scene* s, const int* idx, int n) {
    v3 center = position_mean(s, idx, n);
    if (n < 16) {
        v3 alt = center_alternative(s, idx, n);
        double m_r = bounding_radius(s, idx, n, center);
        double a_r = bounding_radius(s, idx, n, alt);
        if (m_r > a_r) center = alt;
    }
    return center;
}

/* bound::position_spread_in_mother (Frames.cpp:110-124) */
static int too_close(const oracle_scene* s, const int* idx, int n) {
    if (n < 2) return 0;
    v3 mean = position_mean(s, idx, n);
    v3 u = V(0., 0., 0.);
    for (int i = 0; i < n; i++) {
        v3 r = sub(s->f[idx[i]].pos, mean);
        u = add(u, V(r.x * r.x, r.y * r.y, r.z * r.z));
    }
    u = dvd(u, (double)n);
    return sqrt(norm(u)) < 1e-6;  /* FRAME_MIN_STRUCTURE_SIZE */
}

/* Frame::cluster_children (Frame.cpp:190-255) */
static void cluster_children(oracle_scene* s, int fi) {
    if (s->f[fi].n_children > 16) {
        int n = s->f[fi].n_children;
        int* old = (int*)malloc(sizeof(int) * n);
        memcpy(old, s->f[fi].children, sizeof(int) * n);
        int* oct[8];
        int noct[8] = {0};
        for (int k = 0; k < 8; k++) oct[k] = (int*)malloc(sizeof(int) * n);
        for (int i = 0; i < n; i++) {
            int c = old[i];
            int k = octant(s->f[c].pos);
            oct[k][noct[k]++] = c;
        }
        s->f[fi].n_children = 0;
        for (int k = 0; k < 8; k++) {
            if (too_close(s, oct[k], noct[k])) {
                for (int i = 0; i < noct[k]; i++) {
                    s->f[oct[k][i]].mother = fi;
                    push_child(&s->f[fi], oct[k][i]);
                }
            } else if (noct[k] == 1) {
                s->f[oct[k][0]].mother = fi;
                push_child(&s->f[fi], oct[k][0]);
            } else if (noct[k] > 1) {
                v3 center = bounding_center(s, oct[k], noct[k]);
                int o = new_frame(s);   /* may move s->f */
                oframe* of = &s->f[o];
                of->mother = fi;
                of->kind = MLGPU_KIND_FRAME;
                of->pos = center;
                of->rot_mode = MLGPU_ROT_XYZ;
                of->t2m = ht_from_pos_rot(center, MLGPU_ROT_XYZ, of->rot);
                for (int q = 0; q < MLGPU_NUM_SURFACE_SLOTS; q++) of->surf[q] = -1;
                push_child(&s->f[fi], o);
                for (int i = 0; i < noct[k]; i++) {
                    oframe* c = &s->f[oct[k][i]];
                    c->pos = sub(c->pos, center);
                    c->t2m = ht_from_pos_rot(c->pos, c->rot_mode, c->rot);
                    c->mother = o;
                    push_child(&s->f[o], oct[k][i]);
                }
            }
        }
        for (int k = 0; k < 8; k++) free(oct[k]);
        free(old);
    }
    int nc = s->f[fi].n_children;
    for (int i = 0; i < nc; i++) cluster_children(s, s->f[fi].children[i]);
}

static void init_root_and_t2w(oracle_scene* s, int fi, int root) {
    oframe* f = &s->f[fi];
    f->root = root;
    if (f->mother >= 0) f->t2w = ht_mul(&s->f[f->mother].t2w, &f->t2m);
    else f->t2w = f->t2m;
    for (int i = 0; i < f->n_children; i++) init_root_and_t2w(s, f->children[i], root);
}

/* Frame::update_bounding_sphere (Frame.cpp:141-147) */
static void update_radius(oracle_scene* s, int fi) {
    oframe* f = &s->f[fi];
    for (int i = 0; i < f->n_children; i++) update_radius(s, f->children[i]);
    f = &s->f[fi];
    if (f->n_children > 0)
        f->radius = bounding_radius(s, f->children, f->n_children, V(0., 0., 0.));
}

static int resolve_func(const oracle_scene* s, int slot, int id) {
    if (id >= 0) return id;
    if (slot == MLGPU_OUTER_REFLECTION || slot == MLGPU_INNER_REFLECTION) return s->def_refl;
    if (slot == MLGPU_OUTER_REFRACTION || slot == MLGPU_INNER_REFRACTION) return s->def_refr;
    return s->def_abs;
}

oracle_scene* oracle_scene_create(const mlgpu_scene_desc* d) {
    init_hex_units();
    if (!d || d->n_frames < 1) { set_err("empty scene"); return NULL; }
    oracle_scene* s = (oracle_scene*)calloc(1, sizeof(oracle_scene));
    s->n_desc = d->n_frames;
    s->n_funcs = d->n_functions;
    s->funcs = (func1*)calloc(d->n_functions + 3, sizeof(func1));
    for (int i = 0; i < d->n_functions; i++) {
        int a = d->func_start[i], b = d->func_start[i + 1];
        if (func1_init(&s->funcs[i], d->func_x + a, d->func_y + a, b - a)) {
            oracle_scene_destroy(s);
            return NULL;
        }
    }
    /* SurfaceEntity.cpp:12-33 */
    const double dx[2] = {200e-9, 1200e-9};
    const double refl[2] = {0, 0}, refr[2] = {1, 1}, absr[2] = {INFINITY, INFINITY};
    s->def_refl = d->n_functions;
    s->def_refr = d->n_functions + 1;
    s->def_abs = d->n_functions + 2;
    func1_init(&s->funcs[s->def_refl], dx, refl, 2);
    func1_init(&s->funcs[s->def_refr], dx, refr, 2);
    func1_init(&s->funcs[s->def_abs], dx, absr, 2);

    for (int i = 0; i < d->n_frames; i++) {
        int fi = new_frame(s);
        oframe* f = &s->f[fi];
        f->mother = d->mother[i];
        if (i == 0 && f->mother != -1) { set_err("frame 0 must be the root"); oracle_scene_destroy(s); return NULL; }
        if (i > 0 && (f->mother < 0 || f->mother >= i)) { set_err("mother must precede child"); oracle_scene_destroy(s); return NULL; }
        f->pos = V(d->pos[3 * i], d->pos[3 * i + 1], d->pos[3 * i + 2]);
        f->rot_mode = d->rot_mode[i];
        memcpy(f->rot, d->rot + 4 * i, sizeof(double) * 4);
        f->t2m = ht_from_pos_rot(f->pos, f->rot_mode, f->rot);
        f->kind = d->kind[i];
        memcpy(f->p, d->params + MLGPU_NUM_PARAMS * i, sizeof(double) * MLGPU_NUM_PARAMS);
        for (int q = 0; q < MLGPU_NUM_SURFACE_SLOTS; q++) {
            int id = d->surface[MLGPU_NUM_SURFACE_SLOTS * i + q];
            if (id >= d->n_functions) { set_err("function index out of range"); oracle_scene_destroy(s); return NULL; }
            f->surf[q] = resolve_func(s, q, id);
        }
        f->allowed = d->allowed_frame ? d->allowed_frame[i] : -1;
        f->sensor = d->sensor_id ? d->sensor_id[i] : -1;
        primitive_init(f);
        if (f->mother >= 0) push_child(&s->f[f->mother], fi);
    }
    /* Frame::init_tree_based_on_mother_child_relations (Frame.cpp:125-130) */
    cluster_children(s, 0);
    init_root_and_t2w(s, 0, 0);
    update_radius(s, 0);
    return s;
}

void oracle_scene_destroy(oracle_scene* s) {
    if (!s) return;
    for (int i = 0; i < s->n; i++) free(s->f[i].children);
    free(s->f);
    if (s->funcs)
        for (int i = 0; i < s->n_funcs + 3; i++) free(s->funcs[i].p);
    free(s->funcs);
    free(s);
}

int oracle_num_frames(const oracle_scene* s) { return s->n; }

static void dump_rec(const oracle_scene* s, int fi, int32_t* frame, int32_t* nch,
                     double* t2w, double* radius, int cap, int* k) {
    if (*k < cap) {
        frame[*k] = fi;
        nch[*k] = s->f[fi].n_children;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 4; c++) t2w[12 * (*k) + 4 * r + c] = s->f[fi].t2w.T[r][c];
        radius[*k] = s->f[fi].radius;
    }
    (*k)++;
    for (int i = 0; i < s->f[fi].n_children; i++)
        dump_rec(s, s->f[fi].children[i], frame, nch, t2w, radius, cap, k);
}

int oracle_tree_dump(const oracle_scene* s, int32_t* frame, int32_t* n_children,
                     double* frame2world, double* radius, int capacity) {
    int k = 0;
    dump_rec(s, 0, frame, n_children, frame2world, radius, capacity, &k);
    return k;
}

int oracle_func_evaluate(const oracle_scene* s, int32_t f, double x, double* y) {
    int idx = f;
    if (f == -1) idx = s->def_refl;
    else if (f == -2) idx = s->def_refr;
    else if (f == -3) idx = s->def_abs;
    if (idx < 0 || idx >= s->n_funcs + 3) return 2;
    return func1_eval(&s->funcs[idx], x, y);
}

/* ---------------------------------------------------------- intersections */
typedef struct {
    int frame;           /* -1: void */
    v3 position;         /* object frame */
    v3 normal;           /* object frame */
    double distance;
    int outside_in;
} isect;

static void isect_make(isect* is, int frame, v3 pos, v3 normal, double dist, v3 incident) {
    is->frame = frame;
    is->position = pos;
    is->normal = normal;
    is->distance = dist;
    /* Intersection::ray_runs_from_outside_to_inside (Intersection.cpp:133-139) */
    is->outside_in = dot(normal, incident) < 0.0;
}

static int hex_inside(double inner, v3 v) {
    /* HexagonalPrismZ::is_inside (HexagonalPrismZ.cpp:14-29) */
    const double u = dot(UNIT_U, v), vv = dot(UNIT_V, v), w = dot(UNIT_W, v);
    return u < inner && u > -inner && vv < inner && vv > -inner && w < inner && w > -inner;
}

static int bound_inside(const oframe* f, v3 v) {
    switch (f->kind) {
    case MLGPU_KIND_SPHERE_CAP_HEX:
    case MLGPU_KIND_ELLIPTICAL_CAP_HEX: return hex_inside(f->hex_inner, v);
    case MLGPU_KIND_SPHERE_CAP_CYL: return v.x * v.x + v.y * v.y < f->cyl_r2;
    case MLGPU_KIND_SPHERE_CAP_RECT: return fabs(v.x) < f->half_x && fabs(v.y) < f->half_y;
    default: return 0;
    }
}

/* SphericalCapRayIntersectionEquation (…Equation.cpp:6-40) +
 * SurfaceWithOuterPrismBound::add_causeal_intersection (…Bound.cpp:7-70) */
static int intersect_cap(const oframe* f, int fi, v3 s, v3 d, isect* out) {
    const double R = f->p[0];
    const double dd = dot(s, s), sd = dot(s, d), d2 = dot(d, d);
    const double Rdz = R * d.z, Rsz = R * s.z;
    const double P = 2.0 * (sd - Rdz) / d2;
    const double Q = (dd - 2.0 * Rsz) / d2;
    /* QuadraticEquation.cpp:12-30 */
    const double p_over_2 = 0.5 * P;
    const double inner = p_over_2 * p_over_2 - Q;
    if (!(inner >= 0.0)) return 0;
    const double sq = sqrt(inner);
    const double minus = -p_over_2 - sq, plus = -p_over_2 + sq;
    const v3 pi = add(s, mul(d, plus)), mi = add(s, mul(d, minus));
    const int p_in = bound_inside(f, pi), m_in = bound_inside(f, mi);
    const int p_c = plus > 0.0 && norm_le(pi, f->radius);
    const int m_c = minus > 0.0 && norm_le(mi, f->radius);
    double t = minus;
    int ok = 0;
    if (p_in && m_in) {
        if (p_c && m_c) { t = (minus >= plus) ? plus : minus; ok = 1; }
        else if (p_c) { t = plus; ok = 1; }
        else if (m_c) { t = minus; ok = 1; }
    } else if (p_in && p_c) { t = plus; ok = 1; }
    else if (m_in && m_c) { t = minus; ok = 1; }
    if (!ok) return 0;
    const v3 p = add(s, mul(d, t));
    if (!vec_neq(s, p)) return 0;
    v3 n = sub(V(0.0, 0.0, R), p);
    n = dvd(n, norm(n));
    isect_make(out, fi, p, n, t, d);
    return 1;
}

/* the causal solution of SurfaceWithOuterPrismBound::add_causeal_intersection
 * (SurfaceWithOuterPrismBound.cpp:7-70) among the two roots of a surface
 * equation: 1 and *t set when one is inside the bound and the enclosing sphere */
static int causal_root(const oframe* f, v3 s, v3 d, double minus, double plus, double* t_out) {
    const v3 pi = add(s, mul(d, plus)), mi = add(s, mul(d, minus));
    const int p_in = bound_inside(f, pi), m_in = bound_inside(f, mi);
    const int p_c = plus > 0.0 && norm_le(pi, f->radius);
    const int m_c = minus > 0.0 && norm_le(mi, f->radius);
    if (p_in && m_in) {
        if (p_c && m_c) { *t_out = (minus >= plus) ? plus : minus; return 1; }
        if (p_c) { *t_out = plus; return 1; }
        if (m_c) { *t_out = minus; return 1; }
        return 0;
    }
    if (p_in && p_c) { *t_out = plus; return 1; }
    if (m_in && m_c) { *t_out = minus; return 1; }
    return 0;
}

/* EllipticalCapRayIntersectionEquation.cpp:8-99 (the ellipsoid
 * x^2/A^2 + y^2/B^2 + (z-C)^2/C^2 = 1; its normal as merlict computes it:
 * factor C/2 sqrt(1 - x^2/A^2 - y^2/B^2), both slopes with 1/A^2) +
 * add_causeal_intersection with the hexagonal bound */
static int intersect_elliptical_cap(const oframe* f, int fi, v3 s, v3 d, isect* out) {
    const double C = f->ell_C, iAA = f->ell_iAA, iBB = f->ell_iBB, iCC = f->ell_iCC;
    const double a = (d.x * d.x) * iAA + (d.y * d.y) * iBB + (d.z * d.z) * iCC;
    const double b = 2.0 * ((s.x * d.x) * iAA + (s.y * d.y) * iBB + d.z * (s.z - C) * iCC);
    const double c = (s.x * s.x) * iAA + (s.y * s.y) * iBB + (s.z * s.z - 2.0 * s.z * C) * iCC;
    const double p_over_2 = 0.5 * (b / a);
    const double inner = p_over_2 * p_over_2 - c / a;
    double t;
    if (!(inner >= 0.0)) return 0;
    {
        const double sq = sqrt(inner);
        if (!causal_root(f, s, d, -p_over_2 - sq, -p_over_2 + sq, &t)) return 0;
    }
    const v3 p = add(s, mul(d, t));
    if (!vec_neq(s, p)) return 0;
    {
        const double factor = C * 0.5 * sqrt(1.0 - (p.x * p.x) * iAA - (p.y * p.y) * iBB);
        const v3 n = normalized(V(factor * (-2.0 * p.x * iAA), factor * (-2.0 * p.y * iAA), 1.0));
        isect_make(out, fi, p, n, t, d);
    }
    return 1;
}

/* XyPlaneRayIntersectionEquation.cpp:6-41 + each planar primitive */
static int intersect_planar(const oframe* f, int fi, v3 s, v3 d, isect* out) {
    const int parallel = (d.z == 0. && (d.x != 0. || d.y != 0.));
    if (parallel) return 0;
    const double v = -s.z / d.z;
    if (!(v > 0.0)) return 0;
    const v3 p = add(s, mul(d, v));
    int inside = 0;
    switch (f->kind) {
    case MLGPU_KIND_PLANE: inside = fabs(p.x) < f->half_x && fabs(p.y) < f->half_y; break;
    case MLGPU_KIND_DISC: inside = p.x * p.x + p.y * p.y < f->cyl_r2; break;
    case MLGPU_KIND_HEX_PLANE: inside = hex_inside(f->hex_inner, p); break;
    case MLGPU_KIND_ANNULUS:
        inside = (p.x * p.x + p.y * p.y < f->cyl_r2) && !(p.x * p.x + p.y * p.y < f->cyl_r2_inner);
        break;
    case MLGPU_KIND_PLANE_DUAL_SPHERICAL: {
        /* DualSphericalPrismZ::is_inside (DualSphericalPrismZ.cpp:21-27) */
        const double c1x = +f->ds_dist / 2.0, c2x = -f->ds_dist / 2.0;
        const double ax = c1x - p.x, ay = 0.0 - p.y;
        const double bx = c2x - p.x, by = 0.0 - p.y;
        inside = (ax * ax + ay * ay <= f->ds_r * f->ds_r) && (bx * bx + by * by <= f->ds_r * f->ds_r);
        break;
    }
    case MLGPU_KIND_TRIANGLE: {
        /* Triangle::is_inside_triangle (Triangle.cpp:92-102) */
        const double Ax = f->p[0], Ay = f->p[1], Bx = f->p[2], By = f->p[3], Cx = f->p[4], Cy = f->p[5];
        const int bA = ((p.x - Bx) * (Ay - By) - (Ax - Bx) * (p.y - By)) < 0.0;
        const int bB = ((p.x - Cx) * (By - Cy) - (Bx - Cx) * (p.y - Cy)) < 0.0;
        const int bC = ((p.x - Ax) * (Cy - Ay) - (Cx - Ax) * (p.y - Ay)) < 0.0;
        inside = (bA == bB) && (bB == bC);
        break;
    }
    default: inside = 0;
    }
    if (!inside) return 0;
    if (!vec_neq(s, p)) return 0;
    isect_make(out, fi, p, V(0.0, 0.0, 1.0), v, d);
    return 1;
}

/* ZaxisCylinderRayIntersectionEquation.cpp:6-57 + Cylinder.cpp:121-150 */
static int intersect_cylinder(const oframe* f, int fi, v3 s, v3 d, isect* out) {
    const double radius = f->p[0], length = f->p[1];
    const double dx2 = d.x * d.x, dy2 = d.y * d.y;
    const double sxdx = s.x * d.x, sydy = s.y * d.y;
    const double sx2 = s.x * s.x, sy2 = s.y * s.y;
    const double r2 = radius * radius;
    const double dnom = dx2 + dy2;
    const double p = 2.0 * (sxdx + sydy) / dnom;
    const double q = (sx2 + sy2 - r2) / dnom;
    const double p_over_2 = 0.5 * p;
    const double inner = p_over_2 * p_over_2 - q;
    if (!(inner >= 0.0)) return 0;
    const double sq = sqrt(inner);
    const double plus = -p_over_2 + sq, minus = -p_over_2 - sq;
    int causal = 0;
    double v = 0.0;
    if (minus < 0.0 && plus >= 0.0) { causal = 1; v = plus; }
    if (plus >= 0.0 && minus >= 0.0) { causal = 1; v = minus; }
    if (!causal) return 0;
    const v3 pt = add(s, mul(d, v));
    if (!(fabs(pt.z) <= length / 2.0)) return 0;
    if (!vec_neq(s, pt)) return 0;
    v3 n = V(pt.x, pt.y, 0.0);
    n = dvd(n, norm(n));
    isect_make(out, fi, pt, n, v, d);
    return 1;
}

/* Sphere.cpp:62-105 */
static int intersect_sphere(const oframe* f, int fi, v3 s, v3 d, isect* out) {
    const double sd = dot(s, d), d2 = dot(d, d), ss = dot(s, s);
    const double r2 = f->p[0] * f->p[0];
    const double P = 2 * (sd / d2), Q = ss / d2 - r2;
    const double p_over_2 = 0.5 * P;
    const double inner = p_over_2 * p_over_2 - Q;
    if (!(inner >= 0.0)) return 0;
    const double sq = sqrt(inner);
    const double vP = -p_over_2 + sq, vM = -p_over_2 - sq;
    double t;
    if (vP >= 0.0 && vM >= 0.0) t = vM;
    else if (!(vP < 0.0 && vM < 0.0)) t = vP;
    else return 0;
    const v3 p = add(s, mul(d, t));
    const v3 n = dvd(p, norm(p));
    if (!vec_neq(s, p)) return 0;
    isect_make(out, fi, p, n, t, d);
    return 1;
}

static int intersect_leaf(const oframe* f, int fi, v3 s, v3 d, isect* out) {
    if (g_work) {
        switch (f->kind) {
        case MLGPU_KIND_SPHERE_CAP_HEX:
        case MLGPU_KIND_SPHERE_CAP_CYL:
        case MLGPU_KIND_SPHERE_CAP_RECT:
        case MLGPU_KIND_ELLIPTICAL_CAP_HEX: g_work->leaf_cap++; break;
        case MLGPU_KIND_CYLINDER: g_work->leaf_cylinder++; break;
        case MLGPU_KIND_SPHERE: g_work->leaf_sphere++; break;
        default: g_work->leaf_planar++;
        }
    }
    switch (f->kind) {
    case MLGPU_KIND_SPHERE_CAP_HEX:
    case MLGPU_KIND_SPHERE_CAP_CYL:
    case MLGPU_KIND_SPHERE_CAP_RECT: return intersect_cap(f, fi, s, d, out);
    case MLGPU_KIND_PLANE:
    case MLGPU_KIND_DISC:
    case MLGPU_KIND_HEX_PLANE:
    case MLGPU_KIND_ANNULUS:
    case MLGPU_KIND_PLANE_DUAL_SPHERICAL:
    case MLGPU_KIND_TRIANGLE: return intersect_planar(f, fi, s, d, out);
    case MLGPU_KIND_CYLINDER: return intersect_cylinder(f, fi, s, d, out);
    case MLGPU_KIND_SPHERE: return intersect_sphere(f, fi, s, d, out);
    case MLGPU_KIND_ELLIPTICAL_CAP_HEX: return intersect_elliptical_cap(f, fi, s, d, out);
    default: return 0;
    }
}

/* ray_intersects_frames_bounding_sphere (RayAndFrame.cpp:16-107) */
static int sphere_hit(const oframe* f, v3 s, v3 d) {
    const v3 c = V(f->t2w.T[0][3], f->t2w.T[1][3], f->t2w.T[2][3]);
    const double alpha = dot(d, c) - dot(s, d);
    const v3 q = add(s, mul(d, alpha));
    const v3 sc = sub(c, q);
    const double dist2 = dot(sc, sc);
    if (dist2 > f->radius * f->radius) return 0;
    if (alpha < 0.0) return norm_le(sub(c, s), f->radius);
    return 1;
}

typedef struct {
    const oracle_scene* s;
    v3 sup, dir;
    isect best;
    int have;
    int64_t n_sphere, n_leaf;
} query;

/* CausalIntersection (RayAndFrame.cpp:126-173): DFS in child order, then the
 * first minimum distance among candidates. */
static void dfs(query* q, int fi) {
    const oframe* f = &q->s->f[fi];
    q->n_sphere++;
    if (g_work) g_work->sphere_tests++;
    if (!sphere_hit(f, q->sup, q->dir)) return;
    if (f->n_children > 0) {
        for (int i = 0; i < f->n_children; i++) dfs(q, f->children[i]);
        return;
    }
    q->n_leaf++;
    const v3 ls = ht_pos_inv(&f->t2w, q->sup);
    const v3 ld = ht_orient_inv(&f->t2w, q->dir);
    isect is;
    if (intersect_leaf(f, fi, ls, ld, &is)) {
        if (!q->have || is.distance < q->best.distance) { q->best = is; q->have = 1; }
    }
}

static isect first_hit(const oracle_scene* s, int root, v3 sup, v3 dir, int64_t* ns, int64_t* nl) {
    if (g_work) g_work->queries++;
    query q;
    q.s = s; q.sup = sup; q.dir = dir; q.have = 0; q.n_sphere = 0; q.n_leaf = 0;
    dfs(&q, root);
    if (ns) *ns += q.n_sphere;
    if (nl) *nl += q.n_leaf;
    if (q.have) return q.best;
    isect v;
    /* void intersection at 1e4 m (RayAndFrame.cpp:161-167) */
    isect_make(&v, -1, add(sup, mul(dir, 1e4)), dir, 1e4, dir);
    return v;
}

int oracle_first_hit(const oracle_scene* s, int32_t root_frame, const double support[3],
                     const double direction[3], int32_t* frame, double* distance,
                     double position[3], int64_t* sphere_tests, int64_t* leaf_tests) {
    if (root_frame < 0 || root_frame >= s->n) { set_err("bad root frame"); return 1; }
    v3 sup = V(support[0], support[1], support[2]);
    v3 dir = normalized(V(direction[0], direction[1], direction[2]));
    isect is = first_hit(s, root_frame, sup, dir, sphere_tests, leaf_tests);
    *frame = is.frame;
    *distance = is.distance;
    v3 p = is.frame >= 0 ? ht_pos(&s->f[is.frame].t2w, is.position) : is.position;
    position[0] = p.x; position[1] = p.y; position[2] = p.z;
    return 0;
}

/* ------------------------------------------------------------------ Philox */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void oracle_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = counter[0], c1 = counter[1], c2 = counter[2], c3 = counter[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += PHILOX_W0; k1 += PHILOX_W1; }
        const uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        const uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double oracle_uniform(uint64_t seed, uint32_t stream, uint64_t photon_id, uint32_t draw) {
    const uint32_t ctr[4] = {draw >> 2, stream, (uint32_t)photon_id, (uint32_t)(photon_id >> 32)};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    oracle_philox4x32_10(ctr, key, o);
    return (double)(o[draw & 3] >> 8) * (1.0 / 16777216.0);
}

typedef struct {
    uint32_t ctr[4];
    uint32_t key[2];
    uint32_t buf[4];
    uint32_t next;  /* draw index */
} philox_gen;

static void gen_init(philox_gen* g, uint64_t seed, uint64_t id) {
    g->ctr[0] = 0; g->ctr[1] = 0;
    g->ctr[2] = (uint32_t)id; g->ctr[3] = (uint32_t)(id >> 32);
    g->key[0] = (uint32_t)seed; g->key[1] = (uint32_t)(seed >> 32);
    g->next = 0;
}

static double gen_uniform(philox_gen* g) {
    if ((g->next & 3) == 0) {
        g->ctr[0] = g->next >> 2;
        oracle_philox4x32_10(g->ctr, g->key, g->buf);
    }
    const uint32_t w = g->buf[g->next & 3];
    g->next++;
    return (double)(w >> 8) * (1.0 / 16777216.0);
}

/* ----------------------------------------------------------- photon sources
 * The photon sources of the library restated: Philox4x32-10 blocks
 * (block, 1, id_lo, id_hi); block 0 direction, block 1 wavelength, block
 * 2 + k disc try k; 53-bit uniforms; log and sin/cos by series so that only
 * correctly rounded IEEE operations occur. */
static void src_block(uint64_t seed, uint64_t id, uint32_t block, uint32_t w[4]) {
    const uint32_t ctr[4] = {block, 1u, (uint32_t)id, (uint32_t)(id >> 32)};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    oracle_philox4x32_10(ctr, key, w);
}

static double src_u53(uint32_t a, uint32_t b) {
    /* (a >> 5) 2^-27 + (b >> 6) 2^-53 */
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0;
}

double oracle_log_series(double x) {
    /* x = m 2^e, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh((m-1)/(m+1)) */
    int e;
    double m = frexp(x, &e);  /* m in [0.5, 1) */
    if (m < 0.70710678118654752440) { m = m * 2.0; e = e - 1; }
    const double z = (m - 1.0) / (m + 1.0);
    const double z2 = z * z;
    double acc = 1.0 / 25.0;
    for (int k = 11; k >= 0; k--) acc = acc * z2 + 1.0 / (double)(2 * k + 1);
    return (double)e * 6.93147180559945286227e-01 + ((double)e * 2.31904681384629955842e-17 + 2.0 * z * acc);
}

/* sin, cos of t in [0, pi/2]: nested Taylor factors down to t^25 */
static void src_sincos(double t, double* s, double* c) {
    const double t2 = t * t;
    double fs = 1.0, fc = 1.0;
    for (int k = 12; k >= 1; k--) {
        fs = 1.0 - t2 / ((double)(2 * k) * (double)(2 * k + 1)) * fs;
        fc = 1.0 - t2 / ((double)(2 * k - 1) * (double)(2 * k)) * fc;
    }
    *s = t * fs;
    *c = fc;
}

void oracle_sincos_turn(double u, double* s, double* c) {
    const double u4 = u * 4.0;
    const int q = (int)u4;
    double a, b;
    src_sincos((u4 - (double)q) * 1.57079632679489661923, &a, &b);
    switch (q) {
    case 0: *s = a; *c = b; break;
    case 1: *s = b; *c = -a; break;
    case 2: *s = -a; *c = -b; break;
    default: *s = -b; *c = a; break;
    }
}

/* photon id of the source before the pose: support (px, py), direction
   about +z, wavelength */
static void src_local(const mlgpu_source* src, uint64_t seed, uint64_t id, double* px_out, double* py_out,
                      double d[3], double* lambda_out) {
    {
        uint32_t w[4];
        double dx = 0.0, dy = 0.0, dz = 1.0, lambda, px = 0.0, py = 0.0;
        src_block(seed, id, 0u, w);
        const double u0 = src_u53(w[0], w[1]), u1 = src_u53(w[2], w[3]);
        if (src->direction == MLGPU_DIRECTION_TILT) {
            double st, ct, sp, cp;
            src_sincos(src->angle * u0, &st, &ct);
            oracle_sincos_turn(u1, &sp, &cp);
            dx = st * cp; dy = st * sp; dz = ct;
        } else if (src->direction == MLGPU_DIRECTION_GAUSSIAN) {
            const double rad = sqrt(-2.0 * oracle_log_series(1.0 - u0));
            double sp, cp;
            oracle_sincos_turn(u1, &sp, &cp);
            const double tx = src->angle * (rad * cp), ty = src->angle * (rad * sp);
            const double len = sqrt(tx * tx + ty * ty + 1.0);
            dx = tx / len; dy = ty / len; dz = 1.0 / len;
        } else if (src->direction == MLGPU_DIRECTION_CAP) {
            double sa, ca, sp, cp;
            src_sincos(src->angle, &sa, &ca);
            const double cz = 1.0 - u0 * (1.0 - ca);
            const double sz = sqrt(1.0 - cz * cz);
            oracle_sincos_turn(u1, &sp, &cp);
            dx = sz * cp; dy = sz * sp; dz = cz;
        } else if (src->direction == MLGPU_DIRECTION_RING) {
            const double rad = sqrt(-2.0 * oracle_log_series(1.0 - u0));
            double sb, cb, st, ct, sp, cp;
            oracle_sincos_turn(u1, &sb, &cb);
            double zen = src->angle + src->angle_width * (rad * cb);
            if (zen < 0.0) zen = -zen;
            if (zen > 1.57079632679489661923) zen = 1.57079632679489661923;
            src_sincos(zen, &st, &ct);
            src_block(seed, id, 1u, w);
            oracle_sincos_turn(src_u53(w[2], w[3]), &sp, &cp);
            dx = st * cp; dy = st * sp; dz = ct;
        }
        src_block(seed, id, 1u, w);
        if (src->lambda_min == src->lambda_max) {
            lambda = src->lambda_min;
        } else if (src->spectrum == MLGPU_SPECTRUM_UNIFORM) {
            lambda = src->lambda_min + src_u53(w[0], w[1]) * (src->lambda_max - src->lambda_min);
        } else {
            const double inv_lo = 1.0 / src->lambda_min, inv_hi = 1.0 / src->lambda_max;
            lambda = 1.0 / (inv_lo - src_u53(w[0], w[1]) * (inv_lo - inv_hi));
        }
        if (src->support == MLGPU_SUPPORT_DISC) {
            /* Generator::get_point_on_xy_disc_within_radius (Generator.cpp:22-34),
             * at most 64 tries */
            const double r = src->radius;
            for (uint32_t k = 0; k < 64; k++) {
                src_block(seed, id, 2u + k, w);
                const double x = r * (src_u53(w[0], w[1]) * 2.0 - 1.0);
                const double y = r * (src_u53(w[2], w[3]) * 2.0 - 1.0);
                if (x * x + y * y <= r * r) { px = x; py = y; break; }
            }
        }
        *px_out = px; *py_out = py;
        d[0] = dx; d[1] = dy; d[2] = dz;
        *lambda_out = lambda;
    }
}

void oracle_generate_photons(const mlgpu_source* src, int64_t n, uint64_t seed,
                             uint64_t photon_id_offset, double* const out[7]) {
    for (int64_t i = 0; i < n; i++) {
        double px, py, d[3], lambda;
        src_local(src, seed, photon_id_offset + (uint64_t)i, &px, &py, d, &lambda);
        const double dx = d[0], dy = d[1], dz = d[2];
        const double g = src->start_distance;
        const double sx = px - dx * g, sy = py - dy * g, sz = 0.0 - dz * g;
        const double* R = src->rot;
        out[0][i] = sx * R[0] + sy * R[1] + sz * R[2] + src->pos[0];
        out[1][i] = sx * R[3] + sy * R[4] + sz * R[5] + src->pos[1];
        out[2][i] = sx * R[6] + sy * R[7] + sz * R[8] + src->pos[2];
        out[3][i] = dx * R[0] + dy * R[1] + dz * R[2];
        out[4][i] = dx * R[3] + dy * R[4] + dz * R[5];
        out[5][i] = dx * R[6] + dy * R[7] + dz * R[8];
        out[6][i] = lambda;
    }
}

/* Arrival-time jitter: Box-Muller normal from Philox block (0, 2, id_lo, id_hi) */
double oracle_arrival_jitter(uint64_t seed, uint64_t photon_id) {
    const uint32_t ctr[4] = {0u, 2u, (uint32_t)photon_id, (uint32_t)(photon_id >> 32)};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    double s, c;
    oracle_philox4x32_10(ctr, key, w);
    const double rad = sqrt(-2.0 * oracle_log_series(1.0 - src_u53(w[0], w[1])));
    oracle_sincos_turn(src_u53(w[2], w[3]), &s, &c);
    return rad * c;
}

void oracle_histogram_arrivals(const int32_t* sensor_id, const uint8_t* error,
                               const double* time_of_flight, int64_t n,
                               const mlgpu_histogram* h, uint32_t* counts) {
    for (int64_t i = 0; i < n; i++) {
        const int32_t sensor = sensor_id[i];
        if (error[i] != 0 || sensor < 0 || sensor >= h->num_sensors) continue;
        double t = time_of_flight[i];
        if (h->jitter_sigma > 0.0)
            t = t + h->jitter_sigma * oracle_arrival_jitter(h->seed, h->photon_id_offset + (uint64_t)i);
        const double bin = (t - h->t_start) / h->bin_width;
        if (bin >= 0.0 && bin < (double)h->num_bins)
            counts[(int64_t)sensor * h->num_bins + (int64_t)bin] += 1u;
    }
}

/* -------------------------------------------------------------- propagator */
#define C_VACUUM 299792458.0  /* PhysicalConstants.h:7 */

typedef struct {
    int frame;
    v3 pos_obj;      /* object frame position */
    v3 pos_world;
    double distance;
    int outside_in;
} hist_entry;

static int surf_func(const oracle_scene* s, int fi, int slot) {
    if (fi < 0) return resolve_func(s, slot, -1);  /* VOID_SURFACE_ENTITY */
    return s->f[fi].surf[slot];
}

/* Intersection property lookups (Intersection.cpp:83-121) */
static int eval_coming_from_refraction(const oracle_scene* s, const isect* is, double lam, double* n) {
    int f = surf_func(s, is->frame, is->outside_in ? MLGPU_OUTER_REFRACTION : MLGPU_INNER_REFRACTION);
    return func1_eval(&s->funcs[f], lam, n);
}

static void propagate_one(const oracle_scene* s, int world, v3 sup, v3 dir, double lam,
                          uint32_t max_int, uint64_t seed, uint64_t id,
                          int64_t k, oracle_results* out, int64_t* ns, int64_t* nl) {
    philox_gen g;
    gen_init(&g, seed, id);
    int root = world;
    int n_hist = 1;
    int last_interaction = MLGPU_PRODUCTION;
    isect last_is;
    int have_last = 0;
    v3 prev_world = sup, last_world = sup;
    double tof = 0.0, acc = 0.0;
    int err = MLGPU_PHOTON_OK;
    dir = normalized(dir);                              /* Ray.cpp:16-20 */
    if (!(lam > 0.0)) err = MLGPU_PHOTON_BAD_WAVELENGTH; /* Photon.cpp:20-28 */

    /* Propagator::propagate loop (PhotonAndFrame.cpp:13-21) */
    while (!err && (uint32_t)n_hist <= max_int) {
        if (g_work) g_work->bounces++;
        isect is = first_hit(s, root, sup, dir, ns, nl);
        int interaction;
        int stop = 1;
        v3 new_sup = sup, new_dir = dir;
        int new_root = root;
        if (is.frame < 0) {
            interaction = MLGPU_ABSORPTION_IN_VOID;
        } else {
            /* absorbed_in_medium_before_reaching_surface (:32-38) */
            const oframe* f = &s->f[is.frame];
            double ell;
            int fa = f->surf[is.outside_in ? MLGPU_OUTER_ABSORPTION : MLGPU_INNER_ABSORPTION];
            if (func1_eval(&s->funcs[fa], lam, &ell)) { err = MLGPU_PHOTON_OUT_OF_TABLE; break; }
            const double survival = exp(-is.distance / ell);
            if (gen_uniform(&g) > survival) {
                interaction = MLGPU_ABSORPTION_IN_VOID;
            } else {
                /* interact_with_object (:40-48) */
                double refl;
                int fr = f->surf[is.outside_in ? MLGPU_OUTER_REFLECTION : MLGPU_INNER_REFLECTION];
                if (func1_eval(&s->funcs[fr], lam, &refl)) { err = MLGPU_PHOTON_OUT_OF_TABLE; break; }
                const v3 hit_world = ht_pos(&f->t2w, is.position);
                if (refl >= gen_uniform(&g)) {
                    /* reflect_on_surface_and_propagate_on (:56-66) */
                    v3 d_obj = ht_orient_inv(&f->t2w, dir);
                    const double x = is.normal.x, y = is.normal.y, z = is.normal.z;
                    /* Vec3::mirror (Vec3.cpp:59-113) */
                    v3 m = V((1. - 2. * x * x) * d_obj.x + -2. * x * y * d_obj.y + -2. * x * z * d_obj.z,
                             -2. * x * y * d_obj.x + (1. - 2. * y * y) * d_obj.y + -2. * y * z * d_obj.z,
                             -2. * x * z * d_obj.x + -2. * y * z * d_obj.y + (1. - 2. * z * z) * d_obj.z);
                    new_sup = hit_world;
                    new_dir = normalized(ht_orient(&f->t2w, m));
                    interaction = MLGPU_REFLECTION_ON_SURFACE;
                    stop = 0;
                } else {
                    const int transparent =
                        f->surf[MLGPU_OUTER_REFRACTION] != s->def_refr ||
                        f->surf[MLGPU_INNER_REFRACTION] != s->def_refr;
                    if (!transparent) {
                        interaction = MLGPU_ABSORPTION_ON_SURFACE;
                    } else {
                        /* fresnel_refraction_and_reflection (:75-87) */
                        double n_from, n_to;
                        int ff = f->surf[is.outside_in ? MLGPU_OUTER_REFRACTION : MLGPU_INNER_REFRACTION];
                        int ft = f->surf[is.outside_in ? MLGPU_INNER_REFRACTION : MLGPU_OUTER_REFRACTION];
                        if (func1_eval(&s->funcs[ff], lam, &n_from) ||
                            func1_eval(&s->funcs[ft], lam, &n_to)) { err = MLGPU_PHOTON_OUT_OF_TABLE; break; }
                        const v3 inc = ht_orient_inv(&f->t2w, dir);
                        const v3 nrm = is.outside_in ? is.normal : mul(is.normal, -1.0);
                        /* Fresnel.cpp:7-38 */
                        if (g_work) g_work->fresnel++;
                        const double cosI = dot(nrm, inc) * (-1.0);
                        const double eta = n_from / n_to;
                        const double sinT2 = eta * eta * (1.0 - cosI * cosI);
                        const double cosT = sqrt(1.0 - sinT2);
                        double R;
                        if (sinT2 > 1.0) {
                            R = 1.0;
                        } else {
                            const double rOrth = (n_from * cosI - n_to * cosT) / (n_from * cosI + n_to * cosT);
                            const double rPar = (n_to * cosI - n_from * cosT) / (n_to * cosI + n_from * cosT);
                            R = (rOrth * rOrth + rPar * rPar) / 2.0;
                        }
                        if (R > gen_uniform(&g)) {
                            v3 d_obj = ht_orient_inv(&f->t2w, dir);
                            const double x = is.normal.x, y = is.normal.y, z = is.normal.z;
                            v3 m = V((1. - 2. * x * x) * d_obj.x + -2. * x * y * d_obj.y + -2. * x * z * d_obj.z,
                                     -2. * x * y * d_obj.x + (1. - 2. * y * y) * d_obj.y + -2. * y * z * d_obj.z,
                                     -2. * x * z * d_obj.x + -2. * y * z * d_obj.y + (1. - 2. * z * z) * d_obj.z);
                            new_sup = hit_world;
                            new_dir = normalized(ht_orient(&f->t2w, m));
                            interaction = MLGPU_FRESNEL_REFLECTION_ON_SURFACE;
                            stop = 0;
                        } else {
                            /* pass_the_boundary_layer (:89-118) */
                            interaction = is.outside_in ? MLGPU_REFRACTION_TO_INSIDE : MLGPU_REFRACTION_TO_OUTSIDE;
                            const int going_to_default = is.outside_in
                                ? f->surf[MLGPU_INNER_REFRACTION] == s->def_refr
                                : f->surf[MLGPU_OUTER_REFRACTION] == s->def_refr;
                            if (f->allowed >= 0 && !going_to_default) new_root = f->allowed;
                            else new_root = f->root;
                            const v3 refr = add(mul(inc, eta), mul(nrm, eta * cosI - cosT));
                            new_sup = hit_world;
                            new_dir = normalized(ht_orient(&f->t2w, refr));
                            stop = 0;
                        }
                    }
                }
            }
        }
        /* push_back_intersection_and_interaction */
        n_hist++;
        last_interaction = interaction;
        last_is = is;
        have_last = 1;
        prev_world = last_world;
        last_world = is.frame >= 0 ? ht_pos(&s->f[is.frame].t2w, is.position) : is.position;
        acc += is.distance;
        {
            double n;
            if (eval_coming_from_refraction(s, &is, lam, &n)) { err = MLGPU_PHOTON_OUT_OF_TABLE; break; }
            tof += n * is.distance / C_VACUUM;
        }
        sup = new_sup;
        dir = new_dir;
        root = new_root;
        if (stop) break;
    }

    out->error[k] = (uint8_t)err;
    if (err) {
        /* merlict aborts the batch with an exception; the record of an
         * erroneous photon is its error code with a neutral final state */
        out->num_interactions[k] = 0;
        out->final_interaction[k] = MLGPU_PRODUCTION;
        out->time_of_flight[k] = 0.0;
        if (out->accumulated_distance) out->accumulated_distance[k] = 0.0;
        out->final_frame[k] = -1;
        out->sensor_id[k] = -1;
        out->x[k] = 0.0; out->y[k] = 0.0; out->cx[k] = 0.0; out->cy[k] = 0.0;
        return;
    }
    out->num_interactions[k] = (uint8_t)n_hist;
    out->final_interaction[k] = (uint8_t)last_interaction;
    out->time_of_flight[k] = tof;
    if (out->accumulated_distance) out->accumulated_distance[k] = acc;
    if (!have_last) {
        out->final_frame[k] = -1;
        out->sensor_id[k] = -1;
        out->x[k] = sup.x; out->y[k] = sup.y;
        out->cx[k] = 0.0; out->cy[k] = 0.0;
        return;
    }
    out->final_frame[k] = last_is.frame;
    out->sensor_id[k] = last_is.frame >= 0 ? s->f[last_is.frame].sensor : -1;
    out->x[k] = last_is.position.x;
    out->y[k] = last_is.position.y;
    /* final_intersection_incident_direction_wrt_frame (RayForPropagation.cpp:114-136) */
    v3 inc = normalized(sub(last_world, prev_world));
    v3 loc = last_is.frame >= 0 ? ht_orient_inv(&s->f[last_is.frame].t2w, inc) : inc;
    out->cx[k] = -1.0 * loc.x;
    out->cy[k] = -1.0 * loc.y;
}

typedef struct {
    const oracle_scene* s;
    int world;
    const mlgpu_photons* ph;
    int64_t begin, end;
    uint32_t max_int;
    uint64_t seed, id0;
    oracle_results* out;
    int64_t ns, nl;
    oracle_work wk;
} work;

static void* worker(void* arg) {
    work* w = (work*)arg;
    g_work = &w->wk;
    for (int64_t i = w->begin; i < w->end; i++) {
        v3 sup = V(w->ph->support_x[i], w->ph->support_y[i], w->ph->support_z[i]);
        v3 dir = V(w->ph->direction_x[i], w->ph->direction_y[i], w->ph->direction_z[i]);
        propagate_one(w->s, w->world, sup, dir, w->ph->wavelength[i], w->max_int, w->seed,
                      w->id0 + (uint64_t)i, i, w->out, &w->ns, &w->nl);
    }
    g_work = NULL;
    return NULL;
}

void oracle_last_work(oracle_work* out) { *out = g_last_work; }

int oracle_propagate(const oracle_scene* s, int32_t world, const mlgpu_photons* ph, int64_t n,
                     uint32_t max_int, uint64_t seed, uint64_t id0, oracle_results* out,
                     int num_threads, int64_t* sphere_tests, int64_t* leaf_tests) {
    if (!s || !ph || !out || n < 0) { set_err("bad argument"); return 1; }
    if (world < 0 || world >= s->n) { set_err("bad world frame"); return 1; }
    if (num_threads < 1) num_threads = 1;
    if (num_threads > 256) num_threads = 256;
    work* w = (work*)calloc(num_threads, sizeof(work));
    pthread_t* th = (pthread_t*)calloc(num_threads, sizeof(pthread_t));
    const int64_t chunk = (n + num_threads - 1) / num_threads;
    for (int t = 0; t < num_threads; t++) {
        w[t].s = s; w[t].world = world; w[t].ph = ph;
        w[t].begin = t * chunk < n ? t * chunk : n;
        w[t].end = (t + 1) * chunk < n ? (t + 1) * chunk : n;
        w[t].max_int = max_int; w[t].seed = seed; w[t].id0 = id0; w[t].out = out;
    }
    if (num_threads == 1) {
        worker(&w[0]);
    } else {
        for (int t = 0; t < num_threads; t++) pthread_create(&th[t], NULL, worker, &w[t]);
        for (int t = 0; t < num_threads; t++) pthread_join(th[t], NULL);
    }
    int64_t ns = 0, nl = 0;
    memset(&g_last_work, 0, sizeof(g_last_work));
    for (int t = 0; t < num_threads; t++) {
        ns += w[t].ns; nl += w[t].nl;
        const oracle_work* k = &w[t].wk;
        g_last_work.queries += k->queries; g_last_work.sphere_tests += k->sphere_tests;
        g_last_work.leaf_cap += k->leaf_cap; g_last_work.leaf_planar += k->leaf_planar;
        g_last_work.leaf_cylinder += k->leaf_cylinder; g_last_work.leaf_sphere += k->leaf_sphere;
        g_last_work.func_lookups += k->func_lookups; g_last_work.func_compares += k->func_compares;
        g_last_work.fresnel += k->fresnel; g_last_work.bounces += k->bounces;
    }
    if (sphere_tests) *sphere_tests = ns;
    if (leaf_tests) *leaf_tests = nl;
    free(w);
    free(th);
    return 0;
}

/* ------------------------------------------------------ CORSIKA bunches
 * merlict_corsika/PhotonFactory.cpp: weight check :138-148 and collapse
 * :8-24 (floor(w) photons, one more when a uniform <= fmod(w, 1)),
 * production point :30-70,126-128 (the ground point x, y [cm -> m] moved
 * against the direction by c t + 1e3 m), direction :72-124 (cx, cy,
 * -sqrt(1 - cx^2 - cy^2)), wavelength :134-136; Ray and Photon normalise
 * their directions (Ray.cpp:16-20, Vec3.cpp:20-26). */
int64_t oracle_corsika_photons(const float* bunches, int64_t n, uint64_t seed, uint64_t id_offset,
                               double* const out[7], int32_t* truth, int64_t capacity) {
    int64_t k = 0;
    for (int64_t b = 0; b < n; b++) {
        const float* c = bunches + 8 * b;
        const double weight = c[6];
        if (weight < 0.0 || weight > 16.0) return -1;
        uint32_t num = (uint32_t)floor(weight);
        const double rest = fmod(weight, 1.0);
        const uint64_t id = id_offset + (uint64_t)b;
        const uint32_t ctr[4] = {0u, 5u, (uint32_t)id, (uint32_t)(id >> 32)};
        const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        uint32_t w[4];
        oracle_philox4x32_10(ctr, key, w);
        if ((double)(w[0] >> 8) * (1.0 / 16777216.0) <= rest) num += 1;
        const double cx = c[2], cy = c[3];
        const double d[3] = {cx, cy, -sqrt(1.0 - cx * cx - cy * cy)};
        double up[3] = {d[0] * -1.0, d[1] * -1.0, d[2] * -1.0};
        const double un = sqrt(up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
        up[0] = up[0] / un; up[1] = up[1] / un; up[2] = up[2] / un;
        const double gx = c[0] * 1e-2, gy = c[1] * 1e-2;
        const double s = c[4] * 1e-9 * 299792458.0 + 1e3;
        const double sx = gx + up[0] * s, sy = gy + up[1] * s, sz = 0.0 + up[2] * s;
        const double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double v[7] = {sx, sy, sz, d[0] / dn, d[1] / dn, d[2] / dn, fabs(c[7] * 1e-9)};
        for (uint32_t j = 0; j < num; j++) {
            if (k >= capacity) return -2;
            for (int q = 0; q < 7; q++) out[q][k] = v[q];
            if (truth) truth[k] = (int32_t)b;
            k++;
        }
    }
    return k;
}

/* ------------------------------------------------------- signal chain */
typedef struct { int32_t sensor; double t; int64_t idx; } sig_key;

static int sig_cmp(const void* a, const void* b) {
    const sig_key* x = (const sig_key*)a;
    const sig_key* y = (const sig_key*)b;
    if (x->sensor != y->sensor) return x->sensor < y->sensor ? -1 : 1;
    if (x->t != y->t) return x->t < y->t ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

int64_t oracle_sensor_pipelines(const int32_t* sensor_id, const uint8_t* error, const double* time_of_flight,
                                const double* wavelength, const int32_t* truth_id, int64_t n, int32_t num_sensors,
                                int64_t* offsets, double* arrival_time, double* out_wavelength,
                                int32_t* out_truth) {
    sig_key* k = (sig_key*)malloc(sizeof(sig_key) * (size_t)(n > 0 ? n : 1));
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++)
        if (sensor_id[i] >= 0 && sensor_id[i] < num_sensors && error[i] == 0) {
            k[m].sensor = sensor_id[i]; k[m].t = time_of_flight[i]; k[m].idx = i; m++;
        }
    qsort(k, (size_t)m, sizeof(sig_key), sig_cmp);
    int64_t j = 0;
    for (int32_t s = 0; s <= num_sensors; s++) {
        while (j < m && k[j].sensor < s) j++;
        offsets[s] = j;
    }
    for (int64_t q = 0; q < m; q++) {
        arrival_time[q] = k[q].t;
        out_wavelength[q] = wavelength[k[q].idx];
        out_truth[q] = truth_id ? truth_id[k[q].idx] : (int32_t)k[q].idx;
    }
    free(k);
    return m;
}

/* sequential 24-bit uniforms of Philox stream `stream` of (seed, id) */
typedef struct { uint32_t key[2], ctr[4], buf[4], next; } sig_gen;
static void sig_init(sig_gen* g, uint64_t seed, uint64_t id, uint32_t stream) {
    g->key[0] = (uint32_t)seed; g->key[1] = (uint32_t)(seed >> 32);
    g->ctr[0] = 0u; g->ctr[1] = stream; g->ctr[2] = (uint32_t)id; g->ctr[3] = (uint32_t)(id >> 32);
    g->next = 0u;
}
static double sig_uniform(sig_gen* g) {
    if ((g->next & 3u) == 0u) { g->ctr[0] = g->next >> 2; oracle_philox4x32_10(g->ctr, g->key, g->buf); }
    const uint32_t w = g->buf[g->next & 3u];
    g->next++;
    return (double)(w >> 8) * (1.0 / 16777216.0);
}

/* Func1 of (x, y) with the slopes of the unsorted table (Func1.cpp:22-71) */
typedef struct { double x, y, slope; } sig_pt;
static int sig_pt_cmp(const void* a, const void* b) {
    const double x = ((const sig_pt*)a)->x, y = ((const sig_pt*)b)->x;
    return x < y ? -1 : (x > y);
}
static int sig_eval(const sig_pt* p, int32_t n, double x, double* y) {
    if (!(x >= p[0].x) || !(x < p[n - 1].x)) return 0;
    int32_t lo = 0, hi = n;
    while (lo < hi) { const int32_t mid = (lo + hi) / 2; if (p[mid].x > x) hi = mid; else lo = mid + 1; }
    *y = p[lo - 1].slope * x + p[lo - 1].y - p[lo - 1].slope * p[lo - 1].x;
    return 1;
}

int64_t oracle_photoelectric(const int64_t* offsets, const double* arrival_time, const double* wavelength,
                             const int32_t* truth_id, int32_t num_sensors, const double* qe_x, const double* qe_y,
                             int32_t qe_n, double dark_rate, double crosstalk, double exposure_time, uint64_t seed,
                             int64_t* out_offsets, double* out_time, int32_t* out_truth, int64_t capacity) {
    sig_pt* p = (sig_pt*)malloc(sizeof(sig_pt) * (size_t)qe_n);
    for (int32_t i = 0; i < qe_n; i++) {
        p[i].x = qe_x[i]; p[i].y = qe_y[i];
        p[i].slope = (i != qe_n - 1) ? (qe_y[i] - qe_y[i + 1]) / (qe_x[i] - qe_x[i + 1]) : 0.0;
    }
    qsort(p, (size_t)qe_n, sizeof(sig_pt), sig_pt_cmp);
    int64_t k = 0;
    int64_t rc = 0;
    /* add_pulse (PhotoElectricConverter.cpp:53-66): the pulse, then a
       crosstalk pulse while crosstalk >= u, all drawn from g */
#define SIG_PUSH(T, TR) do { if (k >= capacity) { rc = -2; goto done; } out_time[k] = (T); out_truth[k] = (TR); k++; \
        while (crosstalk >= sig_uniform(&g)) { if (k >= capacity) { rc = -2; goto done; } \
            out_time[k] = (T); out_truth[k] = -202; k++; } } while (0)
    for (int32_t s = 0; s < num_sensors; s++) {
        out_offsets[s] = k;
        /* get_pulse_pipeline_for_photon_pipeline (:31-51), photon j of the
           sorted pipelines drawing from stream 6 of (seed, j) */
        for (int64_t j = offsets[s]; j < offsets[s + 1]; j++) {
            sig_gen g;
            sig_init(&g, seed, (uint64_t)j, 6u);
            double q;
            if (!sig_eval(p, qe_n, wavelength[j], &q)) { rc = -1; goto done; }
            if (q >= sig_uniform(&g)) SIG_PUSH(arrival_time[j], truth_id[j]);
        }
        /* add_accidental_pulse (:68-89), sensor s drawing from stream 8 of
           (seed, s) */
        sig_gen g;
        sig_init(&g, seed, (uint64_t)s, 8u);
        double sum = -log(sig_uniform(&g)) / dark_rate;
        while (sum < exposure_time) {
            const double next = -log(sig_uniform(&g)) / dark_rate;
            SIG_PUSH(sum, -201);
            sum += next;
        }
    }
    out_offsets[num_sensors] = k;
    rc = k;
done:
#undef SIG_PUSH
    free(p);
    return rc;
}

void oracle_extract_pulses(const int64_t* offsets, const double* arrival_time, int32_t num_sensors,
                           double duration, double std_dev, uint64_t seed, int16_t* slice) {
    /* extract_pulses (pulse_extraction.cpp:9-30), pulse j drawing its normal
       from stream 7 of (seed, j) */
    for (int64_t j = 0; j < offsets[num_sensors]; j++) {
        sig_gen g;
        sig_init(&g, seed, (uint64_t)j, 7u);
        const double u1 = 1.0 - sig_uniform(&g), u2 = sig_uniform(&g);
        const double t = arrival_time[j] + (0.0 + std_dev * sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2));
        const int32_t sl = (int32_t)round(t / duration);
        slice[j] = (sl >= 0 && sl < 100) ? (int16_t)sl : (int16_t)-1;
    }
}

/* ------------------------------------------------ night-sky background */

/* Func1 from (x, y) pairs in the given order: slopes first, then sorted;
   0 on a duplicate x (Func1.cpp:22-50) */
static int nsb_func(const double* x, const double* y, int32_t n, sig_pt* p) {
    for (int32_t i = 0; i < n; i++) {
        p[i].x = x[i]; p[i].y = y[i];
        p[i].slope = (i != n - 1) ? (y[i] - y[i + 1]) / (x[i] - x[i + 1]) : 0.0;
    }
    qsort(p, (size_t)n, sizeof(sig_pt), sig_pt_cmp);
    for (int32_t i = 0; i + 1 < n; i++) if (p[i].x == p[i + 1].x) return 0;
    return 1;
}

#define NSB_STEPS 4096
/* SamplesFromDistribution (SamplesFromDistribution.cpp:8-17) over
   function::integral and function::inverse (tools.cpp:37-76); cdf gets
   NSB_STEPS + 1 points; returns the total integral, or -1 on failure */
static double nsb_cdf(const double* fx, const double* fy, int32_t n, sig_pt* cdf) {
    sig_pt* f = (sig_pt*)malloc(sizeof(sig_pt) * (size_t)n);
    sig_pt* F = (sig_pt*)malloc(sizeof(sig_pt) * (NSB_STEPS + 1));
    double* X = (double*)malloc(sizeof(double) * (NSB_STEPS + 1));
    double* Y = (double*)malloc(sizeof(double) * (NSB_STEPS + 1));
    double total = -1.0, v;
    if (!nsb_func(fx, fy, n, f)) goto out;
    {
        const double lo = f[0].x, hi = f[n - 1].x, step = (hi - lo) / (double)NSB_STEPS;
        X[0] = lo; Y[0] = 0.0;
        for (int i = 1; i < NSB_STEPS; i++) {
            X[i] = lo + (double)i * step;
            if (!sig_eval(f, n, X[i], &v)) goto out;
            Y[i] = v * step + Y[i - 1];
        }
        if (!sig_eval(f, n, nextafter(hi, lo), &v)) goto out;
        X[NSB_STEPS] = hi; Y[NSB_STEPS] = v * step + Y[NSB_STEPS - 1];
        if (!nsb_func(X, Y, NSB_STEPS + 1, F)) goto out;
        double mx = F[0].y;
        for (int i = 0; i <= NSB_STEPS; i++) if (F[i].y > mx) mx = F[i].y;
        /* the inverse on the grid of F */
        const double Flo = F[0].x, Fhi = F[NSB_STEPS].x, Fstep = (Fhi - Flo) / (double)NSB_STEPS;
        if (!sig_eval(F, NSB_STEPS + 1, Flo, &v)) goto out;
        X[0] = v; Y[0] = Flo;
        for (int i = 1; i < NSB_STEPS; i++) {
            const double xi = Flo + (double)i * Fstep;
            if (!sig_eval(F, NSB_STEPS + 1, xi, &v)) goto out;
            X[i] = v; Y[i] = xi;
        }
        if (!sig_eval(F, NSB_STEPS + 1, nextafter(Fhi, Flo), &v)) goto out;
        X[NSB_STEPS] = v; Y[NSB_STEPS] = Fhi;
        if (!nsb_func(X, Y, NSB_STEPS + 1, cdf)) goto out;
        total = mx;
    }
out:
    free(f); free(F); free(X); free(Y);
    return total;
}

double oracle_nsb_light_rate(const double* fx, const double* fy, int32_t n, double max_fov_radius,
                             double max_aperture_radius) {
    sig_pt* cdf = (sig_pt*)malloc(sizeof(sig_pt) * (NSB_STEPS + 1));
    const double total = nsb_cdf(fx, fy, n, cdf);
    free(cdf);
    if (total < 0.0) return -1.0;
    /* Light.cpp:15-27: fov 1.2 x, aperture 1.05 x (NightSkyBackground.h) */
    const double fov = 1.20 * max_fov_radius, ap = 1.05 * max_aperture_radius;
    const double solid = 2.0 * M_PI * (1.0 - cos(fov));
    const double area = M_PI * ap * ap;
    return total * area * solid;
}

typedef struct { double t; int64_t order; } nsb_key;
static int nsb_key_cmp(const void* a, const void* b) {
    const nsb_key* x = (const nsb_key*)a;
    const nsb_key* y = (const nsb_key*)b;
    if (x->t != y->t) return x->t < y->t ? -1 : 1;
    return x->order < y->order ? -1 : (x->order > y->order);
}

/* inject_nsb_into_photon_pipeline (Injector.cpp:11-107) with sensor s
   drawing from Philox stream 9 of (seed, s); returns the merged count, -1
   for a bad table or draw, -2 when capacity is too small */
int64_t oracle_inject_nsb(const int64_t* offsets, const double* arrival_time, const double* wavelength,
                          const int32_t* truth_id, int32_t num_sensors, const float* eff, const double* flux_x,
                          const double* flux_y, int32_t flux_n, double rate, double exposure_time, uint64_t seed,
                          int64_t* out_offsets, double* out_time, double* out_wavelength, int32_t* out_truth,
                          int64_t capacity, double* start_out) {
    *start_out = 0.0;
    if (num_sensors == 0) return 0;
    sig_pt* cdf = (sig_pt*)malloc(sizeof(sig_pt) * (NSB_STEPS + 1));
    const double total = nsb_cdf(flux_x, flux_y, flux_n, cdf);
    if (total < 0.0) { free(cdf); return -1; }
    const int64_t m = offsets[num_sensors];
    /* the mode of the arrival times (Injector.cpp:23-66) */
    double lo = 1.7976931348623157e308, hi = 2.2250738585072014e-308;
    for (int32_t s = 0; s < num_sensors; s++)
        if (offsets[s + 1] > offsets[s]) {
            if (arrival_time[offsets[s]] < lo) lo = arrival_time[offsets[s]];
            if (arrival_time[offsets[s + 1] - 1] > hi) hi = arrival_time[offsets[s + 1] - 1];
        }
    double mode = 0.0;
    if (m > 0) {
        const unsigned count = (unsigned)(2u + sqrt((double)(unsigned)m));
        double* edge = (double*)malloc(sizeof(double) * count);
        unsigned* bins = (unsigned*)calloc(count, sizeof(unsigned));
        const double step = (hi - lo) / (double)(count - 1u);
        for (unsigned i = 0; i < count; i++) edge[i] = lo + step * (double)i;
        for (int64_t j = 0; j < m; j++) {
            /* first edge > t */
            unsigned a = 0, b = count;
            while (a < b) { const unsigned mid = (a + b) / 2; if (edge[mid] > arrival_time[j]) b = mid; else a = mid + 1; }
            if (a == 0 || a == count) continue;
            bins[a - 1]++;
        }
        unsigned best = 0, best_c = 0;
        for (unsigned i = 0; i + 1 < count; i++) if (bins[i] > best_c) { best_c = bins[i]; best = i; }
        mode = 0.5 * (edge[best + 1] + edge[best]);
        free(edge); free(bins);
    }
    const double start = mode - 0.5 * exposure_time;
    *start_out = start;
    int64_t k = 0, rc = 0;
    nsb_key* keys = NULL;
    double* bt = NULL;
    double* bl = NULL;
    size_t cap_keys = 0, cap_b = 0;
    for (int32_t s = 0; s < num_sensors; s++) {
        const double r = rate * (double)eff[s] / (double)num_sensors;
        if (!(r >= 0.0) || r * exposure_time > 1e7) { rc = -1; goto done; }
        sig_gen g;
        sig_init(&g, seed, (uint64_t)s, 9u);
        int64_t nb = 0;
        double sum = -log(sig_uniform(&g)) / r;
        while (sum < exposure_time) {
            const double next = -log(sig_uniform(&g)) / r;
            if ((size_t)nb >= cap_b) {
                cap_b = cap_b ? 2 * cap_b : 64;
                bt = (double*)realloc(bt, sizeof(double) * cap_b);
                bl = (double*)realloc(bl, sizeof(double) * cap_b);
            }
            bt[nb++] = start + sum;
            sum += next;
        }
        for (int64_t q = 0; q < nb; q++)
            if (!sig_eval(cdf, NSB_STEPS + 1, sig_uniform(&g) * total, &bl[q])) { rc = -1; goto done; }
        const int64_t na = offsets[s + 1] - offsets[s];
        if ((size_t)(na + nb) > cap_keys) {
            cap_keys = (size_t)(na + nb) * 2;
            keys = (nsb_key*)realloc(keys, sizeof(nsb_key) * cap_keys);
        }
        for (int64_t a = 0; a < na; a++) { keys[a].t = arrival_time[offsets[s] + a]; keys[a].order = a; }
        for (int64_t q = 0; q < nb; q++) { keys[na + q].t = bt[q]; keys[na + q].order = na + q; }
        /* stable by time: a pipeline photon before a background photon of
           the same time */
        qsort(keys, (size_t)(na + nb), sizeof(nsb_key), nsb_key_cmp);
        out_offsets[s] = k;
        for (int64_t i = 0; i < na + nb; i++) {
            if (k >= capacity) { rc = -2; goto done; }
            const int64_t o = keys[i].order;
            out_time[k] = keys[i].t - start;
            if (o < na) { out_wavelength[k] = wavelength[offsets[s] + o]; out_truth[k] = truth_id[offsets[s] + o]; }
            else { out_wavelength[k] = bl[o - na]; out_truth[k] = -100; }
            k++;
        }
    }
    out_offsets[num_sensors] = k;
    rc = k;
done:
    free(cdf); free(keys); free(bt); free(bl);
    return rc;
}

/* ---------------------------------------------- light-field calibration */

/* one_photon's result (Calibrator.cpp:48-102) from the photon's draw and its
   propagation result */
void oracle_calibration_photons(const mlgpu_source* src, uint64_t seed, uint64_t photon_id_offset,
                                const int32_t* sensor_id, const uint8_t* error, const double* time_of_flight,
                                int64_t n, int32_t num_lixel, mlgpu_calibration_photon* out) {
    for (int64_t i = 0; i < n; i++) {
        mlgpu_calibration_photon c;
        memset(&c, 0, sizeof(c));
        if (sensor_id[i] >= 0 && sensor_id[i] < num_lixel && error[i] == 0) {
            double px, py, d[3], lambda;
            src_local(src, seed, photon_id_offset + (uint64_t)i, &px, &py, d, &lambda);
            const double* R = src->rot;
            c.reached_sensor = 1u;
            c.lixel_id = (uint32_t)sensor_id[i];
            c.x_pos_on_principal_aperture = (float)(px * R[0] + py * R[1] + 0.0 * R[2] + src->pos[0]);
            c.y_pos_on_principal_aperture = (float)(px * R[3] + py * R[4] + 0.0 * R[5] + src->pos[1]);
            c.x_tilt_vs_optical_axis = (float)(-(d[0] * R[0] + d[1] * R[1] + d[2] * R[2]));
            c.y_tilt_vs_optical_axis = (float)(-(d[0] * R[3] + d[1] * R[4] + d[2] * R[5]));
            c.time_of_flight = (float)(time_of_flight[i] - src->start_distance / 299792458.0);
        }
        out[i] = c;
    }
}

/* LixelStatisticsFiller over n records in order (LixelStatisticsFiller.cpp:
   8-62, OnlineStatistics.cpp:15-45); -1 for a reached lixel id out of range */
int oracle_lixel_statistics(const mlgpu_calibration_photon* rec, int64_t n, int64_t num_lixel, int64_t num_blocks,
                            int64_t photons_per_block, mlgpu_lixel_statistic* out) {
    const double emitted = (double)num_blocks * (double)photons_per_block / (double)num_lixel;
    unsigned* count = (unsigned*)calloc((size_t)num_lixel, sizeof(unsigned));
    double* acc = (double*)calloc((size_t)num_lixel * 15, sizeof(double)); /* (sum, mean, m2) x 5 */
    int rc = 0;
    for (int64_t i = 0; i < n; i++) {
        if (!rec[i].reached_sensor) continue;
        if (rec[i].lixel_id >= (uint64_t)num_lixel) { rc = -1; break; }
        const uint32_t l = rec[i].lixel_id;
        count[l]++;
        const double ns = (double)count[l];
        const double v[5] = {rec[i].x_tilt_vs_optical_axis, rec[i].y_tilt_vs_optical_axis,
                             rec[i].x_pos_on_principal_aperture, rec[i].y_pos_on_principal_aperture,
                             rec[i].time_of_flight};
        for (int q = 0; q < 5; q++) {
            double* a = acc + (size_t)l * 15 + 3 * q;
            const double delta = v[q] - a[1];
            a[1] += delta / ns;
            a[2] += delta * (v[q] - a[1]);
            a[0] += v[q];
        }
    }
    for (int64_t l = 0; rc == 0 && l < num_lixel; l++) {
        const double ns = (double)count[l];
        float m[5], sd[5];
        for (int q = 0; q < 5; q++) {
            const double* a = acc + (size_t)l * 15 + 3 * q;
            m[q] = (float)(a[0] / ns);
            sd[q] = (float)sqrt(ns < 2.0 ? NAN : a[2] / ns);
        }
        mlgpu_lixel_statistic r;
        r.efficiency = (float)((double)count[l] / emitted);
        r.efficiency_std = (float)((double)sqrtf((float)count[l]) / emitted);
        r.cx_mean = m[0]; r.cx_std = sd[0];
        r.cy_mean = m[1]; r.cy_std = sd[1];
        r.x_mean = m[2]; r.x_std = sd[2];
        r.y_mean = m[3]; r.y_std = sd[3];
        r.time_delay_mean = m[4]; r.time_delay_std = sd[4];
        out[l] = r;
    }
    free(count); free(acc);
    return rc;
}

/* -------------------------------------------------------------- rendering */

static v3 mirror_v(v3 n, v3 v) {
    /* Vec3::mirror (Vec3.cpp:59-113) */
    const double x = n.x, y = n.y, z = n.z;
    return V((1. - 2. * x * x) * v.x + -2. * x * y * v.y + -2. * x * z * v.z,
             -2. * x * y * v.x + (1. - 2. * y * y) * v.y + -2. * y * z * v.z,
             -2. * x * z * v.x + -2. * y * z * v.y + (1. - 2. * z * z) * v.z);
}

/* PinHoleCamera (PinHoleCamera.cpp:10-77) and Tracer (Tracer.cpp:21-145)
   per pixel, pixel p drawing from Philox stream 10 of (seed, p); the ray
   ends black after max_interaction_depth reflections and refractions.
   image: rows x cols x 3; returns the number of pixels whose ray met a
   function that does not cover 533 nm (NaN colour), -1 on bad arguments. */
int64_t oracle_render_pinhole(const oracle_scene* s, int32_t world, const mlgpu_pinhole_camera* cam,
                              const mlgpu_render_config* cfg, const float* colors, uint64_t seed, float* image) {
    if (world < 0 || world >= s->n || cam->num_cols < 1 || cam->num_rows < 1) return -1;
    const double lam = 533e-9;
    const double rot[4] = {cam->rotation[0], cam->rotation[1], cam->rotation[2], 0.0};
    const homtra T = ht_from_pos_rot(V(cam->position[0], cam->position[1], cam->position[2]), MLGPU_ROT_XYZ, rot);
    const v3 axis = normalized(ht_orient(&T, V(0.0, 0.0, 1.0)));
    const v3 scol = ht_orient(&T, V(0.0, 1.0, 0.0));
    const v3 srow = ht_orient(&T, V(1.0, 0.0, 0.0));
    const double dist = ((double)(unsigned)cam->num_cols / 2.0) / tan(cam->field_of_view / 2.0);
    const v3 pp = mul(axis, dist);
    const v3 pos = V(cam->position[0], cam->position[1], cam->position[2]);
    const v3 light = V(cfg->incoming_direction[0], cfg->incoming_direction[1], cfg->incoming_direction[2]);
    int64_t errors = 0;
    const int64_t npix = (int64_t)cam->num_rows * cam->num_cols;
    for (int64_t p = 0; p < npix; p++) {
        const int row = (int)(p / cam->num_cols), col = (int)(p % cam->num_cols);
        v3 dir = normalized(add(add(pp, mul(srow, (double)(row - cam->num_rows / 2))),
                                mul(scol, (double)(col - cam->num_cols / 2))));
        v3 sup = pos;
        int root = world;
        sig_gen g;
        sig_init(&g, seed, (uint64_t)p, 10u);
        float rgb[3] = {0.0f, 0.0f, 0.0f};
        int bad = 0, bounces = 0;
        for (;;) {
            const isect is = first_hit(s, root, sup, dir, NULL, NULL);
            if (is.frame < 0) {
                for (int k = 0; k < 3; k++) rgb[k] = cfg->sky_color[k];
                break;
            }
            const oframe* f = &s->f[is.frame];
            const v3 inc = ht_orient_inv(&f->t2w, dir);
            const v3 hit_world = ht_pos(&f->t2w, is.position);
            double refl;
            if (func1_eval(&s->funcs[f->surf[is.outside_in ? MLGPU_OUTER_REFLECTION : MLGPU_INNER_REFLECTION]], lam,
                           &refl)) { bad = 1; break; }
            int reflect = 0;
            v3 refr = V(0.0, 0.0, 0.0);
            int new_root = root;
            if (refl >= sig_uniform(&g)) {
                reflect = 1;
            } else if (f->surf[MLGPU_OUTER_REFRACTION] == s->def_refr && f->surf[MLGPU_INNER_REFRACTION] == s->def_refr) {
                const float* fc = colors ? colors + 6 * is.frame + (is.outside_in ? 0 : 3) : NULL;
                for (int k = 0; k < 3; k++) rgb[k] = fc ? fc[k] : 25.0f;
                if (cfg->global_illumination) {
                    const v3 spec = ht_orient(&f->t2w, mirror_v(is.normal, ht_orient_inv(&f->t2w, light)));
                    double dk = dot(spec, dir);
                    if (dk < 0.0) dk = 0.0;
                    if (dk > 0.750) dk = 0.750;
                    const isect sh = first_hit(s, root, hit_world, normalized(light), NULL, NULL);
                    const v3 nr = ht_orient(&f->t2w, is.normal);
                    double pr = dot(nr, light);
                    if (dot(nr, dir) > 0.0) pr = pr * -1.0;
                    if (sh.frame >= 0 || pr < 0.0) dk = dk * 0.25;
                    const double m = 0.750 - dk;
                    for (int k = 0; k < 3; k++) rgb[k] = (float)((1. - m) * rgb[k] + m * 0.0f);
                }
                break;
            } else {
                double n_from, n_to;
                int ff = f->surf[is.outside_in ? MLGPU_OUTER_REFRACTION : MLGPU_INNER_REFRACTION];
                int ft = f->surf[is.outside_in ? MLGPU_INNER_REFRACTION : MLGPU_OUTER_REFRACTION];
                if (func1_eval(&s->funcs[ff], lam, &n_from) || func1_eval(&s->funcs[ft], lam, &n_to)) { bad = 1; break; }
                const v3 nrm = is.outside_in ? is.normal : mul(is.normal, -1.0);
                const double cosI = dot(nrm, inc) * (-1.0);
                const double eta = n_from / n_to;
                const double sinT2 = eta * eta * (1.0 - cosI * cosI);
                const double cosT = sqrt(1.0 - sinT2);
                double R;
                if (sinT2 > 1.0) {
                    R = 1.0;
                } else {
                    const double ro = (n_from * cosI - n_to * cosT) / (n_from * cosI + n_to * cosT);
                    const double rp = (n_to * cosI - n_from * cosT) / (n_to * cosI + n_from * cosT);
                    R = (ro * ro + rp * rp) / 2.0;
                }
                if (R > sig_uniform(&g)) {
                    reflect = 1;
                } else {
                    const int to_default = is.outside_in ? f->surf[MLGPU_INNER_REFRACTION] == s->def_refr
                                                         : f->surf[MLGPU_OUTER_REFRACTION] == s->def_refr;
                    new_root = (f->allowed >= 0 && !to_default) ? f->allowed : f->root;
                    refr = add(mul(inc, eta), mul(nrm, eta * cosI - cosT));
                }
            }
            if (bounces == cfg->max_interaction_depth) {
                rgb[0] = rgb[1] = rgb[2] = 0.0f;
                break;
            }
            bounces++;
            dir = normalized(ht_orient(&f->t2w, reflect ? mirror_v(is.normal, inc) : refr));
            sup = hit_world;
            root = new_root;
        }
        if (bad) {
            rgb[0] = rgb[1] = rgb[2] = NAN;
            errors++;
        }
        for (int k = 0; k < 3; k++) image[3 * p + k] = rgb[k];
    }
    return errors;
}

/* ------------------------------------------------- FACT photon stream */
static int cmp_double(const void* a, const void* b) {
    const double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : (x > y);
}

/* electric_pipelines_to_photon_stream (merlict_fact/apps/fact_propagation.cpp:
   34-94): the pulses' slices relative to the event's median arrival time,
   pixel by pixel, each pixel ended by 255; returns the length, -1 when
   capacity is too small */
int64_t oracle_fact_photon_stream(const int64_t* offsets, const double* arrival_time, int32_t num_pixels,
                                  double trigger_time, uint8_t* raw, int64_t capacity) {
    const int64_t n = offsets[num_pixels];
    double median = NAN;
    if (n > 0) {
        double* t = (double*)malloc(sizeof(double) * (size_t)n);
        memcpy(t, arrival_time, sizeof(double) * (size_t)n);
        qsort(t, (size_t)n, sizeof(double), cmp_double);
        median = t[(unsigned)((double)n * 0.5)];
        free(t);
    }
    const double dur = (double)0.5e-9f;
    int64_t k = 0;
    for (int32_t s = 0; s < num_pixels; s++) {
        for (int64_t j = offsets[s]; j < offsets[s + 1]; j++) {
            const double x = round((arrival_time[j] - median + trigger_time) / dur);
            if (x >= 30.0 && x < 130.0) {
                if (k >= capacity) return -1;
                raw[k++] = (uint8_t)(int)x;
            }
        }
        if (k >= capacity) return -1;
        raw[k++] = 255;
    }
    return k;
}
