// Scene builders: the sceneries the propagator is measured on, written as
// flat frame lists.  Geometry helpers restate merlict's factories in the
// same operation order so that poses agree bit for bit with the reference:
//   HexGridAnnulus        merlict/scenery/geometry/HexGridAnnulus.cpp:23-45
//   HexGridFlower         merlict/scenery/geometry/HexGridFlower.cpp:8-47
//   lens_maker            merlict/scenery/geometry/lens_maker.cpp:15-46
//   BiConvexLens          merlict/scenery/primitive/BiConvexLens.cpp:7-36
//   BiConvexLensHexBound  merlict/scenery/primitive/BiConvexLensHexBound.cpp:7-64
//   segmented reflector   merlict/scenery/segmented_imaging_reflector/Geometry.cpp:14-146
//   light-field sensor    merlict_portal_plenoscope/light_field_sensor/{Geometry,Factory}.cpp
#include <algorithm>
#include <array>
#include <cmath>
#include <string>
#include <vector>

#include "fact_tables.inc"
#include "scene.hpp"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace mlgpu {
namespace {

Vec3d v3(double x, double y, double z) { return Vec3d{x, y, z}; }
Vec3d operator+(Vec3d a, Vec3d b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
Vec3d operator-(Vec3d a, Vec3d b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
Vec3d operator*(Vec3d a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
Vec3d operator/(Vec3d a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
double dot(Vec3d a, Vec3d b) { return b.x * a.x + b.y * a.y + b.z * a.z; }
double norm(Vec3d a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
Vec3d cross(Vec3d a, Vec3d v) {
    return v3(a.y * v.z - a.z * v.y, a.z * v.x - a.x * v.z, a.x * v.y - a.y * v.x);
}
const Vec3d UNIT_X{1., 0., 0.};
const Vec3d UNIT_Y{0., 1., 0.};
const Vec3d UNIT_Z{0., 0., 1.};

// Vec3::angle_in_between (merlict/Vec3.cpp:115-119)
double angle_in_between(Vec3d a, Vec3d b) {
    const Vec3d an = a / norm(a);
    const Vec3d bn = b / norm(b);
    return std::acos(dot(an, bn));
}

std::vector<Vec3d> hex_grid_annulus(double outer_radius, double inner_radius, double spacing) {
    const Vec3d b = UNIT_Y * spacing;
    const Vec3d a = (UNIT_Y * 0.5 + UNIT_X * std::sqrt(3.0) / 2.0) * spacing;
    const int sample_radius = static_cast<int>(2.0 * std::floor(outer_radius / spacing));
    std::vector<Vec3d> grid;
    for (int i = -sample_radius; i <= sample_radius; i++) {
        for (int j = -sample_radius; j <= sample_radius; j++) {
            const Vec3d cell = a * i + b * j;
            const double n = norm(cell);
            if (n <= outer_radius && n >= inner_radius) grid.push_back(cell);
        }
    }
    return grid;
}

// HexagonalPrismZ with outer radius r (HexagonalPrismZ.cpp:8-30)
struct HexPrism {
    double inner;
    Vec3d u, v, w;
    explicit HexPrism(double outer) {
        inner = outer * std::cos(M_PI / 6.0);
        u = UNIT_Y;
        v = UNIT_X * +std::sin(2. / 3. * M_PI) + UNIT_Y * std::cos(2. / 3. * M_PI);
        w = UNIT_X * -std::sin(2. / 3. * M_PI) + UNIT_Y * std::cos(2. / 3. * M_PI);
    }
    bool inside(Vec3d p) const {
        const double pu = dot(u, p), pv = dot(v, p), pw = dot(w, p);
        return pu < inner && pu > -inner && pv < inner && pv > -inner && pw < inner && pw > -inner;
    }
};

struct Flower {
    double spacing;
    std::vector<Vec3d> grid;
};

Flower hex_grid_flower(double outer_hex_grid_radius, unsigned int count_on_diameter) {
    if (count_on_diameter % 2 == 0) count_on_diameter++;
    const unsigned int rows = count_on_diameter;
    const double inner_grid = outer_hex_grid_radius * (std::sqrt(3.0) / 2.0);
    const double inner_grid_diameter = 2.0 * inner_grid;
    const double outer_facet = inner_grid_diameter / (0.5 * (3.0 * static_cast<double>(rows) + 1.0));
    const double inner_facet = outer_facet * (std::sqrt(3.0) / 2.0);
    Flower f;
    f.spacing = 2.0 * inner_facet;
    const HexPrism prism(outer_hex_grid_radius);
    const Vec3d b = UNIT_X * f.spacing;
    const Vec3d a = (UNIT_X * 0.5 + UNIT_Y * std::sqrt(3.0) / 2.0) * f.spacing;
    const int sample_radius = static_cast<int>(2.0 * std::floor(outer_hex_grid_radius / f.spacing));
    for (int i = -sample_radius; i <= sample_radius; i++) {
        for (int j = -sample_radius; j <= sample_radius; j++) {
            const Vec3d cell = a * i + b * j;
            if (prism.inside(cell)) f.grid.push_back(cell);
        }
    }
    return f;
}

double lens_maker_curvature_radius(double f, double r, double n) {
    double R = f;
    double deviation = 0.0;
    unsigned int iteration = 0;
    do {
        if (++iteration > 1000) throw std::runtime_error("lens_maker: no convergence");
        R = R - deviation * 0.1;
        const double t = 2.0 * R - 2.0 * std::sqrt(R * R - r * r);
        const double f_inv = (n - 1.0) * (2.0 / R - ((n - 1.0) * t) / (n * R * R));
        const double f_expected = 1.0 / f_inv;
        deviation = f_expected - f;
    } while (std::fabs(deviation) > 1e-6);
    return R;
}

double cap_height(double R, double r) { return R - std::sqrt(R * R - r * r); }

// function::mean over num_samples points (merlict/function/tools.cpp:29-35)
double func_mean_constant_table(double lower, double upper, double value, uint64_t num) {
    double arg = lower;
    const double increment = (upper - lower) / static_cast<double>(num);
    double sum = 0.0;
    for (uint64_t i = 0; i < num; i++) {
        // a two-node constant table: slope 0, so evaluate() returns value
        const double y = 0.0 * arg + value - 0.0 * lower;
        sum = sum + y;
        arg = arg + increment;
    }
    return sum / static_cast<double>(num);
}

struct ReflectorConfig {
    double focal_length;
    double dc_over_parabolic;
    double max_outer_aperture_radius;
    double min_inner_aperture_radius;
    int outer_aperture_shape_hex;
    double facet_inner_hex_radius;
    double gap_between_facets;
};

struct Facet {
    Vec3d pos;
    Rot rot;
};

std::vector<Facet> reflector_facets(const ReflectorConfig& c) {
    const double spacing = c.facet_inner_hex_radius * 2.0 + c.gap_between_facets;
    const double max_r = c.max_outer_aperture_radius - spacing / 2.0;
    const double min_r = c.min_inner_aperture_radius + spacing / 2.0;
    std::vector<Vec3d> pos;
    if (c.outer_aperture_shape_hex == 1) {
        const Vec3d B = UNIT_Y * spacing;
        const Vec3d A = (UNIT_Y * 0.5 + UNIT_X * std::sqrt(3.0) / 2.0) * spacing;
        const Vec3d U = UNIT_X;
        const Vec3d V = UNIT_Y * +std::sin(2. / 3. * M_PI) + UNIT_X * std::cos(2. / 3. * M_PI);
        const Vec3d W = UNIT_Y * -std::sin(2. / 3. * M_PI) + UNIT_X * std::cos(2. / 3. * M_PI);
        const double R = (std::sqrt(3.0) / 2.0) * max_r;
        const int N = static_cast<int>(2.0 * std::ceil(c.max_outer_aperture_radius / spacing));
        for (int a = -N; a <= N; a++) {
            for (int b = -N; b <= N; b++) {
                const Vec3d fc = A * a + B * b;
                const double u = dot(U, fc), v = dot(V, fc), w = dot(W, fc);
                const bool in_hex = u < R && u > -R && v < R && v > -R && w < R && w > -R;
                const bool out_disc = std::hypot(fc.x, fc.y) > min_r;
                if (in_hex && out_disc) pos.push_back(fc);
            }
        }
    } else {
        pos = hex_grid_annulus(max_r, min_r, spacing);
    }
    const double f = c.focal_length;
    for (auto& p : pos) {
        const double dist = std::sqrt(p.x * p.x + p.y * p.y + 0. * 0.);
        const double z_dc = f - std::sqrt(f * f - dist * dist);
        const double z_pa = 1.0 / (4.0 * f) * dist * dist;
        p.z = c.dc_over_parabolic * z_dc + (1.0 - c.dc_over_parabolic) * z_pa;
    }
    // make_average_facet_distance_to_focal_point_match_focal_length (:78-91)
    const Vec3d focal = v3(0.0, 0.0, f);
    unsigned int iteration = 0;
    double deviation = 0.0;
    do {
        if (++iteration > 100) throw std::runtime_error("reflector: no convergence");
        double sum = 0.0;
        for (const auto& p : pos) sum = sum + norm(p - focal);
        const double avg = sum / static_cast<double>(pos.size());
        deviation = f - avg;
        const double step = -deviation * 0.5;
        for (auto& p : pos) p.z = p.z + step;
    } while (std::fabs(deviation) > f * 1e-6);
    std::vector<Facet> facets;
    for (const auto& p : pos) {
        // get_rotation_for_facet_position (:115-123)
        const Vec3d fp = focal - p;
        const Vec3d axis = cross(fp, UNIT_Z);
        const double angle = angle_in_between(UNIT_Z, fp);
        facets.push_back(Facet{p, Rot::axis(axis, -(0.5 * angle))});
    }
    return facets;
}

void add_reflector(mlgpu_desc* d, int32_t mother, const ReflectorConfig& c, int32_t reflection_fn) {
    const int32_t refl = d->add_frame(mother, v3(0, 0, 0), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
    Surface s;
    s.f[MLGPU_OUTER_REFLECTION] = reflection_fn;
    const double R = c.focal_length * 2.0;
    const double outer_hex = c.facet_inner_hex_radius * 2.0 / std::sqrt(3.0);
    for (const auto& fa : reflector_facets(c))
        d->add_frame(refl, fa.pos, fa.rot, MLGPU_KIND_SPHERE_CAP_HEX, {R, outer_hex}, s);
}

std::vector<double> wavelengths(double lo, double hi, int n) {
    std::vector<double> x;
    for (int i = 0; i < n; i++) x.push_back(lo + (hi - lo) * i / (n - 1));
    return x;
}

// PMMA refractive index, Sellmeier fit n^2 - 1 = 1.1819 l^2 / (l^2 - 0.011313),
// l in micrometres (Sultanova, Kasarova, Nikolov, Acta Phys. Pol. A 116 (2009) 585).
double pmma_index(double lambda_m) {
    const double l2 = (lambda_m * 1e6) * (lambda_m * 1e6);
    return std::sqrt(1.0 + 1.1819 * l2 / (l2 - 0.011313));
}

// Synthetic PMMA absorption length: 1 m at 280 nm to 10 m at 700 nm,
// log-linear in between, clamped outside.
double pmma_absorption_length(double lambda_m) {
    const double l = std::min(std::max(lambda_m, 280e-9), 700e-9);
    return std::pow(10.0, (l - 280e-9) / 420e-9);
}

int32_t add_constant_function(mlgpu_desc* d, double value) {
    return d->add_function({200e-9, 1200e-9}, {value, value});
}

// Triangle::set_normal_and_3_vertecies (merlict/scenery/primitive/Triangle.cpp:18-48)
void add_triangle(mlgpu_desc* d, int32_t mother, Vec3d normal, Vec3d a, Vec3d b, Vec3d c,
                  const Surface& s, int32_t allowed) {
    const Vec3d pos = (a + b + c) / 3.0;
    a = a - pos; b = b - pos; c = c - pos;
    Rot rot = Rot::xyz(0, 0, 0);
    const Vec3d dn = normal - UNIT_Z;
    if (dot(dn, dn) > 1e-14) {
        const Vec3d axis = cross(UNIT_Z, normal);
        const double angle = angle_in_between(UNIT_Z, normal);
        rot = Rot::axis(axis, angle);
        // inverse rotation of the vertices: R^T for the axis-angle matrix
        Vec3d ax = axis / norm(axis);
        ax = ax / norm(ax);
        const double x = ax.x, y = ax.y, z = ax.z;
        const double sn = std::sin(angle), cs = std::cos(angle);
        double T[3][3];
        T[0][0] = cs + x * x * (1. - cs);
        T[0][1] = x * y * (1. - cs) - z * sn;
        T[0][2] = x * z * (1. - cs) + y * sn;
        T[1][0] = y * x * (1. - cs) + z * sn;
        T[1][1] = cs + y * y * (1. - cs);
        T[1][2] = y * z * (1. - cs) - x * sn;
        T[2][0] = z * x * (1. - cs) - y * sn;
        T[2][1] = z * y * (1. - cs) + x * sn;
        T[2][2] = cs + z * z * (1. - cs);
        auto inv = [&](Vec3d v) {
            return v3(v.x * T[0][0] + v.y * T[1][0] + v.z * T[2][0],
                      v.x * T[0][1] + v.y * T[1][1] + v.z * T[2][1],
                      v.x * T[0][2] + v.y * T[1][2] + v.z * T[2][2]);
        };
        a = inv(a); b = inv(b); c = inv(c);
    }
    d->add_frame(mother, pos, rot, MLGPU_KIND_TRIANGLE, {a.x, a.y, b.x, b.y, c.x, c.y}, s, allowed);
}

}  // namespace
}  // namespace mlgpu

using namespace mlgpu;

extern "C" {

int mlgpu_build_lens_scene(double focal_length, double aperture_radius, double refractive_index,
                           double cell_pitch, mlgpu_desc** out) {
    if (!out || !(focal_length > 0) || !(aperture_radius > 0) || !(refractive_index > 1.0) || !(cell_pitch > 0)) {
        set_last_error("mlgpu_build_lens_scene: bad argument");
        return MLGPU_ERR_ARGUMENT;
    }
    try {
        auto* d = new mlgpu_desc();
        const int32_t root = d->add_frame(-1, v3(0, 0, 0), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
        const int32_t n_fn = add_constant_function(d, refractive_index);
        const double R = lens_maker_curvature_radius(focal_length, aperture_radius, refractive_index);
        const double h = cap_height(R, aperture_radius);
        // the lens: a SurfaceEntity whose inner_refraction is the glass; its
        // caps adopt the surface inside out and are restricted to the lens.
        Surface lens;
        lens.f[MLGPU_INNER_REFRACTION] = n_fn;
        const int32_t lf = d->add_frame(root, v3(0, 0, 0), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME, {}, lens);
        const Surface cap = lens.inside_out();
        d->add_frame(lf, v3(0.0, 0.0, -h), Rot::xyz(0, 0, 0), MLGPU_KIND_SPHERE_CAP_CYL, {R, aperture_radius}, cap, lf);
        d->add_frame(lf, v3(0.0, 0.0, h), Rot::xyz(M_PI, 0.0, 0.0), MLGPU_KIND_SPHERE_CAP_CYL, {R, aperture_radius}, cap, lf);
        // sensor: 19 hexagonal cells (two rings) at the focal plane
        const int32_t plane = d->add_frame(root, v3(0, 0, -focal_length), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
        const std::vector<Vec3d> cells = hex_grid_annulus(2.0 * cell_pitch * 1.01, 0.0, cell_pitch);
        int32_t sid = 0;
        for (const auto& c : cells)
            d->add_frame(plane, c, Rot::xyz(0, 0, 0), MLGPU_KIND_HEX_PLANE, {cell_pitch / std::sqrt(3.0)},
                         Surface(), MLGPU_NO_FRAME, sid++);
        validate_desc(d->view());
        *out = d;
        return MLGPU_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return MLGPU_ERR_SCENE;
    }
}

int mlgpu_build_fact_scene(int camera, mlgpu_desc** out) {
    if (!out || camera < 0 || camera > 2) { set_last_error("null argument or bad camera"); return MLGPU_ERR_ARGUMENT; }
    const bool with_light_guides = camera == 1;
    try {
        auto* d = new mlgpu_desc();
        const int32_t root = d->add_frame(-1, v3(0, 0, 0), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
        // the FACT mirror reflectivity (fact_propagation.cpp:145-259)
        std::vector<double> rx, ry;
        for (int k = 0; k < kFactMirrorReflectivityPoints; k++) {
            rx.push_back(kFactMirrorReflectivity[k][0]);
            ry.push_back(kFactMirrorReflectivity[k][1]);
        }
        const int32_t mirror_fn = d->add_function(rx, ry);
        ReflectorConfig rc;
        rc.focal_length = 4.889;
        rc.dc_over_parabolic = 0.5;
        rc.max_outer_aperture_radius = 2.1;
        rc.min_inner_aperture_radius = 0.25;
        rc.outer_aperture_shape_hex = 0;
        rc.facet_inner_hex_radius = 0.3;
        rc.gap_between_facets = 0.01;
        add_reflector(d, root, rc, mirror_fn);

        const int32_t cam = d->add_frame(root, v3(0, 0, 4.889), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
        d->add_frame(cam, v3(0, 0, 0.001), Rot::xyz(0, 0, 0), MLGPU_KIND_DISC, {0.33});

        // the 1440 pixels of the FACT camera in chid order
        // (fact_propagation.cpp:316-341): the app's float millimetres widened
        // to double and divided by 1e3, hexagons of outer radius 4.95 mm
        // turned by 30 deg; sensor id = chid
        std::vector<Vec3d> centre(kFactNumPixels);
        for (int k = 0; k < kFactNumPixels; k++)
            centre[k] = v3(static_cast<double>(kFactPixelXmm[k]) / 1e3, static_cast<double>(kFactPixelYmm[k]) / 1e3,
                           0.0);
        const double pixel_r = 0.00495;
        const double guide_h = 0.015;
        if (camera == 2) {
            d->add_frame(cam, v3(0, 0, -guide_h), Rot::xyz(0, 0, 0), MLGPU_KIND_DISC, {0.33}, Surface(),
                         MLGPU_NO_FRAME, 0);
            validate_desc(d->view());
            *out = d;
            return MLGPU_OK;
        }
        // light guides: entry hexagons (flat sides facing +-x, +-60 deg,
        // +-120 deg, like the pixels) as wide as the tightest pair of
        // neighbouring pixels allows, so that no two guides overlap
        double min_gap = 1.0;
        for (int a = 0; a < kFactNumPixels; a++)
            for (int b = a + 1; b < kFactNumPixels; b++) {
                const double dx = centre[b].x - centre[a].x, dy = centre[b].y - centre[a].y;
                if (dx * dx + dy * dy > 4.0 * pixel_r * pixel_r * 4.0) continue;
                double sep = 0.0;  // separation along the flat normals
                for (int e = 0; e < 3; e++) {
                    const double ang = e * M_PI / 3.0;
                    sep = std::max(sep, std::fabs(dx * std::cos(ang) + dy * std::sin(ang)));
                }
                min_gap = std::min(min_gap, sep);
            }
        const double entry_r = 0.5 * min_gap / std::cos(M_PI / 6.0) * 0.995;
        int32_t n_index = -1, n_abs = -1;
        if (with_light_guides) {
            const std::vector<double> lx = wavelengths(200e-9, 1200e-9, 51);
            std::vector<double> ny, ay;
            for (double l : lx) { ny.push_back(pmma_index(l)); ay.push_back(pmma_absorption_length(l)); }
            n_index = d->add_function(lx, ny);
            n_abs = d->add_function(lx, ay);
        }
        for (int32_t k = 0; k < kFactNumPixels; k++) {
            const Vec3d& p = centre[k];
            if (!with_light_guides) {
                d->add_frame(cam, p, Rot::xyz(0, 0, M_PI / 6.0), MLGPU_KIND_HEX_PLANE, {pixel_r},
                             Surface(), MLGPU_NO_FRAME, k);
                continue;
            }
            // light guide: a hexagonal PMMA frustum from an entry hexagon at
            // z = -h down to the pixel at z = 0, frame rotated like the pixel.
            const int32_t g = d->add_frame(cam, p, Rot::xyz(0, 0, M_PI / 6.0), MLGPU_KIND_FRAME);
            Surface pix;
            pix.f[MLGPU_INNER_ABSORPTION] = n_abs;  // the medium below the pixel is PMMA
            d->add_frame(g, v3(0, 0, 0), Rot::xyz(0, 0, 0), MLGPU_KIND_HEX_PLANE, {pixel_r}, pix,
                         MLGPU_NO_FRAME, k);
            Surface pmma;  // outer side (the normal side) is the guide's PMMA
            pmma.f[MLGPU_OUTER_REFRACTION] = n_index;
            pmma.f[MLGPU_OUTER_ABSORPTION] = n_abs;
            d->add_frame(g, v3(0, 0, -guide_h), Rot::xyz(0, 0, 0), MLGPU_KIND_HEX_PLANE, {entry_r}, pmma, g);
            for (int e = 0; e < 6; e++) {
                const double a0 = e * M_PI / 3.0, a1 = (e + 1) * M_PI / 3.0;
                const Vec3d E0 = v3(entry_r * std::cos(a0), entry_r * std::sin(a0), -guide_h);
                const Vec3d E1 = v3(entry_r * std::cos(a1), entry_r * std::sin(a1), -guide_h);
                const Vec3d X0 = v3(pixel_r * std::cos(a0), pixel_r * std::sin(a0), 0.0);
                const Vec3d X1 = v3(pixel_r * std::cos(a1), pixel_r * std::sin(a1), 0.0);
                // inward normal of the wall through E0, E1, X1
                Vec3d nrm = cross(E1 - E0, X1 - E0);
                nrm = nrm / norm(nrm);
                const Vec3d mid = (E0 + E1 + X0 + X1) / 4.0;
                if (dot(nrm, mid) > 0.0) nrm = nrm * -1.0;
                add_triangle(d, g, nrm, E0, E1, X1, pmma, g);
                add_triangle(d, g, nrm, E0, X1, X0, pmma, g);
            }
        }
        validate_desc(d->view());
        *out = d;
        return MLGPU_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return MLGPU_ERR_SCENE;
    }
}

int mlgpu_build_plenoscope_scene(double focal_length, double max_outer_aperture_radius,
                                 double min_inner_aperture_radius, double facet_inner_hex_radius,
                                 double gap_between_facets, double max_FoV_diameter_deg,
                                 double hex_pixel_FoV_flat2flat_deg, int num_paxel_on_pixel_diagonal,
                                 double housing_overhead, double lens_refraction,
                                 double mirror_reflection, mlgpu_desc** out) {
    if (!out || !(focal_length > 0) || !(max_outer_aperture_radius > 0) || num_paxel_on_pixel_diagonal < 1) {
        set_last_error("mlgpu_build_plenoscope_scene: bad argument");
        return MLGPU_ERR_ARGUMENT;
    }
    try {
        auto* d = new mlgpu_desc();
        const int32_t root = d->add_frame(-1, v3(0, 0, 0), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
        const int32_t mirror_fn = add_constant_function(d, mirror_reflection);
        const int32_t lens_fn = add_constant_function(d, lens_refraction);
        const int32_t portal = d->add_frame(root, v3(0, 0, 0), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
        ReflectorConfig rc;
        rc.focal_length = focal_length;
        rc.dc_over_parabolic = 0.0;
        rc.max_outer_aperture_radius = max_outer_aperture_radius;
        rc.min_inner_aperture_radius = min_inner_aperture_radius;
        rc.outer_aperture_shape_hex = 0;
        rc.facet_inner_hex_radius = facet_inner_hex_radius;
        rc.gap_between_facets = gap_between_facets;
        add_reflector(d, portal, rc, mirror_fn);

        // light_field_sensor::Geometry (Geometry.cpp:14-187)
        const double deg = M_PI / 180.0;
        const double f_exp = focal_length;
        const double ap_exp = max_outer_aperture_radius;
        const double max_fov_diameter = (M_PI * max_FoV_diameter_deg) / 180.0;
        const double pixel_fov = (M_PI * hex_pixel_FoV_flat2flat_deg) / 180.0;
        (void)deg;
        const double max_fov_radius = max_fov_diameter / 2.0;
        const double max_sensor_r = f_exp * std::tan(max_fov_radius);
        const double lens_inner_ap = f_exp * std::tan(pixel_fov / 2.0);
        const double lens_outer_ap = lens_inner_ap * 2.0 / std::sqrt(3.0);
        const double pixel_spacing = lens_inner_ap * 2.0;
        const Flower paxel = hex_grid_flower(lens_outer_ap, static_cast<unsigned int>(num_paxel_on_pixel_diagonal));
        const double lixel_spacing = paxel.spacing;
        const double lixel_inner_r = 0.5 * lixel_spacing;
        const double lixel_outer_r = lixel_inner_r * 2.0 / std::sqrt(3.0);
        const double ap_image_r = lens_inner_ap - 0.5 * lixel_inner_r;
        const double naive_f_over_d = f_exp / (2.0 * ap_exp);
        const double lens_f = naive_f_over_d * 2.0 * ap_image_r;
        const double paxel_dist = lens_f;
        const std::vector<Vec3d> pixels = hex_grid_annulus(max_sensor_r - lens_outer_ap, 0.0, pixel_spacing);
        const double rr = 1.0 + paxel_dist / f_exp;
        const double mean_n = func_mean_constant_table(200e-9, 1200e-9, lens_refraction, 137);
        const double lens_R = 1.20 * lens_maker_curvature_radius(lens_f, lens_outer_ap, mean_n);
        const double housing_r = max_sensor_r * housing_overhead;
        const double bin_h = paxel_dist * 0.25;

        const int32_t lfs = d->add_frame(portal, v3(0, 0, focal_length), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
        // front: lens array + face plate
        const int32_t front = d->add_frame(lfs, v3(0, 0, 0), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
        const int32_t lens_array = d->add_frame(front, v3(0, 0, 0), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
        Surface lens_surface;
        lens_surface.f[MLGPU_INNER_REFRACTION] = lens_fn;
        const Surface cap = lens_surface.inside_out();
        const double h = cap_height(lens_R, lens_outer_ap);
        const double in_ap = lens_outer_ap * std::sqrt(3.0) / 2.0;
        const double wall_h = 2.0 * (cap_height(lens_R, lens_outer_ap) +
                                     std::sqrt(lens_R * lens_R - in_ap * in_ap) - lens_R);
        for (const auto& pp : pixels) {
            const int32_t lf = d->add_frame(lens_array, pp, Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME, {}, lens_surface);
            d->add_frame(lf, v3(0.0, 0.0, -h), Rot::xyz(0, 0, 0), MLGPU_KIND_SPHERE_CAP_HEX, {lens_R, lens_outer_ap}, cap, lf);
            d->add_frame(lf, v3(0.0, 0.0, h), Rot::xyz(M_PI, 0.0, 0.0), MLGPU_KIND_SPHERE_CAP_HEX, {lens_R, lens_outer_ap}, cap, lf);
            for (unsigned int i = 0; i < 6; i++) {
                const double phi = static_cast<double>(i) * 1.0 / 3.0 * M_PI;
                d->add_frame(lf, v3(in_ap * std::sin(phi), in_ap * std::cos(phi), 0.0),
                             Rot::xyz(M_PI * 0.5, M_PI * 0.5, phi), MLGPU_KIND_PLANE_DUAL_SPHERICAL,
                             {wall_h, lens_outer_ap});
            }
        }
        const int32_t face_plate = d->add_frame(front, v3(0, 0, 0), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
        for (const auto& fp : hex_grid_annulus(max_sensor_r + pixel_spacing, max_sensor_r - lens_outer_ap, pixel_spacing))
            d->add_frame(face_plate, fp, Rot::xyz(0, 0, 0), MLGPU_KIND_HEX_PLANE, {lens_outer_ap});
        d->add_frame(face_plate, v3(0, 0, 0), Rot::xyz(0, 0, 0), MLGPU_KIND_ANNULUS, {housing_r, max_sensor_r});
        // housing
        const double housing_h = .67 * housing_r;
        const int32_t housing = d->add_frame(lfs, v3(0, 0, 0), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
        d->add_frame(housing, v3(0.0, 0.0, housing_h), Rot::xyz(0, 0, 0), MLGPU_KIND_DISC, {housing_r});
        {
            // Cylinder::set_cylinder (Cylinder.cpp:16-45): centre between start and end
            const Vec3d start = v3(0.0, 0.0, 0.0), end = v3(0.0, 0.0, housing_h);
            const Vec3d axis = end - start;
            const double length = norm(end - start);
            d->add_frame(housing, start + axis / 2.0, Rot::xyz(0.0, 0.0, 0.0), MLGPU_KIND_CYLINDER, {housing_r, length});
        }
        // bins
        Surface bin_surface;
        bin_surface.f[MLGPU_OUTER_REFLECTION] = mirror_fn;
        const int32_t bin_array = d->add_frame(lfs, v3(0.0, 0.0, paxel_dist), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
        for (const auto& pp : pixels) {
            const int32_t bin = d->add_frame(bin_array, pp * rr, Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
            for (unsigned int i = 0; i < 6; i++) {
                const double phi = static_cast<double>(i) * 1.0 / 3.0 * M_PI;
                d->add_frame(bin, v3(lens_inner_ap * std::sin(phi), lens_inner_ap * std::cos(phi), -0.5 * bin_h),
                             Rot::xyz(-M_PI * 0.5, 0.0, phi), MLGPU_KIND_PLANE, {lens_outer_ap, bin_h}, bin_surface);
            }
        }
        // lixels
        const int32_t lixel_array = d->add_frame(lfs, v3(0.0, 0.0, paxel_dist), Rot::xyz(0, 0, 0), MLGPU_KIND_FRAME);
        int32_t sid = 0;
        for (const auto& pp : pixels) {
            const Vec3d center = pp * rr;
            for (const auto& q : paxel.grid)
                d->add_frame(lixel_array, center + q, Rot::xyz(0.0, 0.0, 1.0 / 6.0 * M_PI), MLGPU_KIND_HEX_PLANE,
                             {lixel_outer_r}, Surface(), MLGPU_NO_FRAME, sid++);
        }
        validate_desc(d->view());
        *out = d;
        return MLGPU_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return MLGPU_ERR_SCENE;
    }
}

}  // extern "C"
