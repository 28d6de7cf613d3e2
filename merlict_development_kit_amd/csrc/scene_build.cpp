// Host-side flattening of a scene description into the device layout of
// device_scene.hpp.  Replaces merlict's post-initialisation
// (Frame::init_tree_based_on_mother_child_relations, merlict/Frame.cpp:125-147)
// for the GPU: frame2world poses, leaf geometry, surface tables and a
// bounding-sphere hierarchy in depth-first order with skip indices.
#include "scene_build.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <numeric>
#include <string>
#include <stdexcept>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace mlgpu {
namespace {

struct M34 {
    double T[3][4];
};

M34 unit() {
    M34 m = {{{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}}};
    return m;
}

// composition A*G of 3x4 homogeneous transforms (merlict/HomTra3.cpp:245-268)
M34 compose(const M34& A, const M34& G) {
    M34 r;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            r.T[i][j] = A.T[i][0] * G.T[0][j] + A.T[i][1] * G.T[1][j] + A.T[i][2] * G.T[2][j];
        r.T[i][3] = A.T[i][0] * G.T[0][3] + A.T[i][1] * G.T[1][3] + A.T[i][2] * G.T[2][3] + A.T[i][3];
    }
    return r;
}

// pose in mother: translation * rotation, the rotation from Tait-Bryan
// angles or axis-angle (merlict/HomTra3.cpp:43-112, merlict/Rot3.cpp:21-33)
M34 pose(const double* pos, int mode, const double* rot) {
    M34 R = unit();
    bool xyz = mode == MLGPU_ROT_XYZ;
    double rx = rot[0], ry = rot[1], rz = rot[2];
    if (!xyz && rot[3] == 0.0) { xyz = true; rx = ry = rz = 0.0; }
    if (xyz) {
        const double cx = std::cos(rx), cy = std::cos(ry), cz = std::cos(rz);
        const double sx = std::sin(rx), sy = std::sin(ry), sz = std::sin(rz);
        R.T[0][0] = cy * cz;
        R.T[0][1] = cx * sz + sx * sy * cz;
        R.T[0][2] = sx * sz - cx * sy * cz;
        R.T[1][0] = -cy * sz;
        R.T[1][1] = cx * cz - sx * sy * sz;
        R.T[1][2] = sx * cz + cx * sy * sz;
        R.T[2][0] = sy;
        R.T[2][1] = -sx * cy;
        R.T[2][2] = cx * cy;
    } else {
        double ax = rot[0], ay = rot[1], az = rot[2];
        // Rot3::set divides the axis by its norm, HomTra3 once more; a stored
        // axis has had the first division already
        const int divisions = mode == MLGPU_ROT_AXIS_STORED ? 1 : 2;
        for (int k = 0; k < divisions; k++) {
            const double nn = std::sqrt(ax * ax + ay * ay + az * az);
            ax /= nn; ay /= nn; az /= nn;
        }
        const double s = std::sin(rot[3]), c = std::cos(rot[3]);
        R.T[0][0] = c + ax * ax * (1. - c);
        R.T[0][1] = ax * ay * (1. - c) - az * s;
        R.T[0][2] = ax * az * (1. - c) + ay * s;
        R.T[1][0] = ay * ax * (1. - c) + az * s;
        R.T[1][1] = c + ay * ay * (1. - c);
        R.T[1][2] = ay * az * (1. - c) - ax * s;
        R.T[2][0] = az * ax * (1. - c) - ay * s;
        R.T[2][1] = az * ay * (1. - c) + ax * s;
        R.T[2][2] = c + az * az * (1. - c);
    }
    M34 Tr = unit();
    Tr.T[0][3] = pos[0]; Tr.T[1][3] = pos[1]; Tr.T[2][3] = pos[2];
    return compose(Tr, R);
}

// frame2world of every description frame exactly as merlict computes it.
// Frame::init_tree_based_on_mother_child_relations (merlict/Frame.cpp:125-147)
// first moves the children of any frame with more than 16 children into
// octant frames centred at their bounding centre (Frame::cluster_children,
// Frame.cpp:190-255; bound::bounding_sphere_center, Frames.cpp:39-124); the
// leaves' frame2world products then run through those octants.  The kernels
// solve each leaf in its object frame with these transforms, so the poses
// are reproduced bit for bit (only the transforms; the GPU traversal uses its
// own sphere tree).
class ReferencePoses {
 public:
    ReferencePoses(const mlgpu_scene_desc& d, const std::vector<double>& radius) {
        const int32_t n = d.n_frames;
        f_.resize(n);
        for (int32_t i = 0; i < n; i++) {
            F& f = f_[i];
            f.mother = d.mother[i];
            for (int k = 0; k < 3; k++) f.pos[k] = d.pos[3 * i + k];
            f.mode = d.rot_mode[i];
            for (int k = 0; k < 4; k++) f.rot[k] = d.rot[4 * i + k];
            f.t2m = pose(f.pos, f.mode, f.rot);
            f.radius = radius[i];
            if (f.mother >= 0) f_[f.mother].children.push_back(i);
        }
        cluster(0);
        t2w(0);
    }
    const M34& frame2world(int32_t i) const { return f_[i].t2w; }

    // Rank of every childless description frame in merlict's candidate order:
    // the depth-first walk over the clustered children lists
    // (RayAndFrame.cpp:137-149), whose first minimum wins exact ties of
    // distance (RayAndFrame.cpp:168-172).
    std::vector<int32_t> candidate_rank(int32_t n_desc) const {
        std::vector<int32_t> rank(n_desc, -1);
        int32_t next = 0;
        std::vector<int32_t> todo{0};
        while (!todo.empty()) {
            const int32_t fi = todo.back();
            todo.pop_back();
            const std::vector<int32_t>& kids = f_[fi].children;
            if (kids.empty()) {
                if (fi < n_desc) rank[fi] = next++;
                continue;
            }
            for (auto it = kids.rbegin(); it != kids.rend(); ++it) todo.push_back(*it);
        }
        return rank;
    }

 private:
    struct F {
        int32_t mother;
        double pos[3];
        int mode;
        double rot[4];
        M34 t2m, t2w;
        double radius;
        std::vector<int32_t> children;
    };
    std::vector<F> f_;

    static double norm3(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

    void mean(const std::vector<int32_t>& idx, double* out) const {
        double s[3] = {0., 0., 0.};
        for (int32_t i : idx)
            for (int k = 0; k < 3; k++) s[k] = s[k] + f_[i].pos[k];
        for (int k = 0; k < 3; k++) out[k] = s[k] / static_cast<double>(idx.size());
    }
    double radius_about(const std::vector<int32_t>& idx, const double* c) const {
        double r = 0.0;
        for (int32_t i : idx) {
            const double d[3] = {c[0] - f_[i].pos[0], c[1] - f_[i].pos[1], c[2] - f_[i].pos[2]};
            const double need = norm3(d) + f_[i].radius;
            if (need > r) r = need;
        }
        return r;
    }
    void alternative(const std::vector<int32_t>& idx, double* out) const {
        if (idx.size() == 1) { for (int k = 0; k < 3; k++) out[k] = f_[idx[0]].pos[k]; return; }
        int32_t a = -1, b = -1;
        double maximum = 0.0;
        for (size_t i = 0; i < idx.size(); i++)
            for (size_t j = i; j < idx.size(); j++) {
                if (i == j) continue;
                const F& p = f_[idx[i]];
                const F& q = f_[idx[j]];
                const double d[3] = {p.pos[0] - q.pos[0], p.pos[1] - q.pos[1], p.pos[2] - q.pos[2]};
                const double dist = norm3(d) + p.radius + q.radius;
                if (dist > maximum) { maximum = dist; a = idx[i]; b = idx[j]; }
            }
        const double zero[3] = {0., 0., 0.};
        const double* ap = a >= 0 ? f_[a].pos : zero;
        const double* bp = b >= 0 ? f_[b].pos : zero;
        const double ar = a >= 0 ? f_[a].radius : 0.0;
        double dir[3] = {bp[0] - ap[0], bp[1] - ap[1], bp[2] - ap[2]};
        const double nn = norm3(dir);
        for (int k = 0; k < 3; k++) dir[k] = dir[k] / nn;
        for (int k = 0; k < 3; k++) out[k] = ap[k] + dir[k] * (0.5 * maximum - ar);
    }
    void centre(const std::vector<int32_t>& idx, double* out) const {
        mean(idx, out);
        if (idx.size() < 16) {
            double alt[3];
            alternative(idx, alt);
            if (radius_about(idx, out) > radius_about(idx, alt))
                for (int k = 0; k < 3; k++) out[k] = alt[k];
        }
    }
    bool too_close(const std::vector<int32_t>& idx) const {
        if (idx.size() < 2) return false;
        double m[3];
        mean(idx, m);
        double u[3] = {0., 0., 0.};
        for (int32_t i : idx) {
            const double r[3] = {f_[i].pos[0] - m[0], f_[i].pos[1] - m[1], f_[i].pos[2] - m[2]};
            u[0] = u[0] + r[0] * r[0];
            u[1] = u[1] + r[1] * r[1];
            u[2] = u[2] + r[2] * r[2];
        }
        for (int k = 0; k < 3; k++) u[k] = u[k] / static_cast<double>(idx.size());
        return std::sqrt(norm3(u)) < 1e-6;
    }
    void cluster(int32_t fi) {
        if (f_[fi].children.size() > 16) {
            const std::vector<int32_t> old = f_[fi].children;
            std::vector<int32_t> oct[8];
            for (int32_t c : old) {
                const double* p = f_[c].pos;
                oct[4 * (p[0] >= 0.) + 2 * (p[1] >= 0.) + (p[2] >= 0.)].push_back(c);
            }
            f_[fi].children.clear();
            for (int k = 0; k < 8; k++) {
                if (too_close(oct[k])) {
                    for (int32_t c : oct[k]) { f_[c].mother = fi; f_[fi].children.push_back(c); }
                } else if (oct[k].size() == 1) {
                    f_[oct[k][0]].mother = fi;
                    f_[fi].children.push_back(oct[k][0]);
                } else if (oct[k].size() > 1) {
                    double c[3];
                    centre(oct[k], c);
                    F o;
                    o.mother = fi;
                    for (int q = 0; q < 3; q++) o.pos[q] = c[q];
                    o.mode = MLGPU_ROT_XYZ;
                    for (int q = 0; q < 4; q++) o.rot[q] = 0.0;
                    o.t2m = pose(o.pos, o.mode, o.rot);
                    o.radius = 0.0;
                    const int32_t oi = static_cast<int32_t>(f_.size());
                    f_.push_back(o);
                    f_[fi].children.push_back(oi);
                    for (int32_t ch : oct[k]) {
                        F& cf = f_[ch];
                        for (int q = 0; q < 3; q++) cf.pos[q] = cf.pos[q] - c[q];
                        cf.t2m = pose(cf.pos, cf.mode, cf.rot);
                        cf.mother = oi;
                        f_[oi].children.push_back(ch);
                    }
                }
            }
        }
        const std::vector<int32_t> kids = f_[fi].children;
        for (int32_t c : kids) cluster(c);
    }
    void t2w(int32_t fi) {
        F& f = f_[fi];
        f.t2w = f.mother >= 0 ? compose(f_[f.mother].t2w, f.t2m) : f.t2m;
        const std::vector<int32_t> kids = f_[fi].children;
        for (int32_t c : kids) t2w(c);
    }
};

// Bounding radius each primitive reports about its own origin and the
// kernel parameters of the leaf (primitive post_initialize_* methods).
double leaf_params(int kind, const double* p, double* out) {
    for (int i = 0; i < 8; i++) out[i] = 0.0;
    switch (kind) {
    case MLGPU_KIND_SPHERE_CAP_HEX: {
        const double R = p[0], r = p[1], rm = r > R ? R : r;
        const double radius = std::hypot(rm, R - std::sqrt(R * R - rm * rm));
        out[0] = R; out[1] = r * std::cos(M_PI / 6.0); out[7] = radius;
        return radius;
    }
    case MLGPU_KIND_SPHERE_CAP_CYL: {
        const double R = p[0], r = p[1], rm = r > R ? R : r;
        const double radius = std::hypot(rm, R - std::sqrt(R * R - rm * rm));
        out[0] = R; out[1] = r * r; out[7] = radius;
        return radius;
    }
    case MLGPU_KIND_SPHERE_CAP_RECT: {
        const double hx = 0.5 * p[1], hy = 0.5 * p[2];
        const double radius = std::hypot(hx, hy);
        out[0] = p[0]; out[1] = hx; out[2] = hy; out[7] = radius;
        return radius;
    }
    case MLGPU_KIND_PLANE: {
        const double hx = 0.5 * p[0], hy = 0.5 * p[1];
        out[1] = hx; out[2] = hy;
        return std::hypot(hx, hy);
    }
    case MLGPU_KIND_DISC:
        out[1] = p[0] * p[0];
        return std::sqrt(out[1]);
    case MLGPU_KIND_HEX_PLANE:
        out[1] = p[0] * std::cos(M_PI / 6.0);
        return p[0];
    case MLGPU_KIND_ANNULUS:
        out[1] = p[0] * p[0]; out[2] = p[1] * p[1];
        return std::sqrt(out[1]);
    case MLGPU_KIND_PLANE_DUAL_SPHERICAL: {
        const double h = p[0], w = p[1];
        const double dist = (w * w - h * h) / (2.0 * h);
        const double cr = dist / 2.0 + h / 2.0;
        out[1] = dist / 2.0; out[2] = cr * cr;
        return w * 0.5;
    }
    case MLGPU_KIND_CYLINDER:
        out[0] = p[0]; out[1] = p[1] / 2.0;
        return std::hypot(0.5 * p[1], p[0]);
    case MLGPU_KIND_TRIANGLE: {
        for (int i = 0; i < 6; i++) out[i] = p[i];
        double m = 0.0;
        for (int i = 0; i < 3; i++) m = std::max(m, std::sqrt(p[2 * i] * p[2 * i] + p[2 * i + 1] * p[2 * i + 1]));
        return m;
    }
    case MLGPU_KIND_SPHERE:
        out[0] = p[0];
        return p[0];
    case MLGPU_KIND_ELLIPTICAL_CAP_HEX: {
        // EllipticalCapWithHexagonalBound.cpp:10-52: the radii of curvature,
        // the outer hex radius restricted to the z one before the bound
        // takes it, the enclosing sphere of the restricted hexagon
        const double mean = p[0] * 0.5;
        const double X = mean * p[1], Y = mean / p[1], Z = 0.5 * (X + Y);
        const double r = p[3] > Z ? Z : p[3];
        const double radius = std::hypot(r, Z - std::sqrt(Z * Z - r * r));
        out[0] = Z;
        out[1] = r * std::cos(M_PI / 6.0);
        out[2] = 1.0 / (X * X);
        out[3] = 1.0 / (Y * Y);
        out[4] = 1.0 / (Z * Z);
        out[5] = r;
        out[7] = radius;
        return radius;
    }
    default:
        return 0.0;
    }
}

// Axis-aligned box in the world frame.
struct Box {
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL};
    double hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    void grow(const double* p) {
        for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
    }
    void grow(const Box& b) {
        for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
    }
    bool empty() const { return lo[0] > hi[0]; }
    double area() const {
        if (empty()) return 0.0;
        const double x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
        return 2.0 * (x * y + y * z + z * x);
    }
    double centre(int k) const { return 0.5 * (lo[k] + hi[k]); }
};

// Object-frame box holding every point where a leaf can report a hit.
Box local_box(int kind, const double* p, const double* prm, double radius) {
    Box b;
    double e[3] = {0, 0, 0}, z0 = 0.0, z1 = 0.0;
    switch (kind) {
    case MLGPU_KIND_SPHERE_CAP_HEX:
    case MLGPU_KIND_SPHERE_CAP_CYL:
    case MLGPU_KIND_SPHERE_CAP_RECT: {
        // hits satisfy |p| <= bounding radius and lie on the sphere
        // x^2 + y^2 + z^2 = 2 R z, so 0 <= z = |p|^2 / (2 R) <= r^2 / (2 R)
        // x, y: the hexagon's corners (+-r_o, 0), (+-r_o / 2, +-r_i), the
        // rectangle's half widths, the circle's radius
        const double ex = kind == MLGPU_KIND_SPHERE_CAP_RECT ? prm[1] : p[1];
        const double ey = kind == MLGPU_KIND_SPHERE_CAP_RECT ? prm[2]
                          : kind == MLGPU_KIND_SPHERE_CAP_HEX ? prm[1] : p[1];
        e[0] = std::min(radius, ex);
        e[1] = std::min(radius, ey);
        z0 = 0.0;
        z1 = std::min(radius, radius * radius / (2.0 * p[0]));
        break;
    }
    case MLGPU_KIND_PLANE: e[0] = prm[1]; e[1] = prm[2]; break;
    case MLGPU_KIND_HEX_PLANE:
        e[0] = radius;  // corners (+-r_o, 0), (+-r_o / 2, +-r_i)
        e[1] = prm[1];
        break;
    case MLGPU_KIND_DISC:
    case MLGPU_KIND_ANNULUS:
    case MLGPU_KIND_PLANE_DUAL_SPHERICAL:
        e[0] = e[1] = radius;
        break;
    case MLGPU_KIND_TRIANGLE: {
        double lo[3] = {std::min({p[0], p[2], p[4]}), std::min({p[1], p[3], p[5]}), 0.0};
        double hi[3] = {std::max({p[0], p[2], p[4]}), std::max({p[1], p[3], p[5]}), 0.0};
        b.grow(lo);
        b.grow(hi);
        return b;
    }
    case MLGPU_KIND_ELLIPTICAL_CAP_HEX:
        // hits lie within the enclosing sphere, inside the hexagon (corners
        // (+-r_o, 0), (+-r_o / 2, +-r_i)) and on the ellipsoid, whose points
        // have 0 <= z <= 2 C
        e[0] = std::min(radius, prm[5]);
        e[1] = std::min(radius, prm[1]);
        z0 = 0.0;
        z1 = std::min(radius, 2.0 * prm[0]);
        break;
    case MLGPU_KIND_CYLINDER: e[0] = e[1] = p[0]; z0 = -0.5 * p[1]; z1 = 0.5 * p[1]; break;
    case MLGPU_KIND_SPHERE: e[0] = e[1] = p[0]; z0 = -p[0]; z1 = p[0]; break;
    default: return b;  // a frame without surface is never hit
    }
    const double lo[3] = {-e[0], -e[1], z0}, hi[3] = {e[0], e[1], z1};
    b.grow(lo);
    b.grow(hi);
    return b;
}

// Corners of a polygon in the object frame's plane z = 0 that holds every
// point where a flat leaf can report a hit; the number of corners, 0 for a
// leaf that is not flat.  Triangle: its corners; plane: the rectangle;
// hexagon: HexagonalPrismZ's corners (+-r_o, 0), (+-r_o / 2, +-r_i)
// (HexagonalPrismZ.cpp:14-29); disc and annulus: the 16-gon around the
// outer circle.
int flat_outline(int kind, const double* p, double (*v)[2]) {
    switch (kind) {
    case MLGPU_KIND_TRIANGLE:
        for (int c = 0; c < 3; c++) { v[c][0] = p[2 * c]; v[c][1] = p[2 * c + 1]; }
        return 3;
    case MLGPU_KIND_PLANE: {
        const double hx = 0.5 * p[0], hy = 0.5 * p[1];
        const double xs[4] = {-hx, hx, hx, -hx}, ys[4] = {-hy, -hy, hy, hy};
        for (int c = 0; c < 4; c++) { v[c][0] = xs[c]; v[c][1] = ys[c]; }
        return 4;
    }
    case MLGPU_KIND_HEX_PLANE: {
        const double ro = p[0], ri = p[0] * std::cos(M_PI / 6.0);
        const double xs[6] = {ro, 0.5 * ro, -0.5 * ro, -ro, -0.5 * ro, 0.5 * ro};
        const double ys[6] = {0.0, ri, ri, 0.0, -ri, -ri};
        for (int c = 0; c < 6; c++) { v[c][0] = xs[c]; v[c][1] = ys[c]; }
        return 6;
    }
    case MLGPU_KIND_DISC:
    case MLGPU_KIND_ANNULUS: {
        const double r = p[0] / std::cos(M_PI / 16.0);
        for (int c = 0; c < 16; c++) {
            v[c][0] = r * std::cos(c * M_PI / 8.0);
            v[c][1] = r * std::sin(c * M_PI / 8.0);
        }
        return 16;
    }
    default:
        return 0;
    }
}

struct BNode {
    Box box;
    int32_t leaf_frame = -1;       // description frame of a leaf
    std::vector<int32_t> kids;
};

class Builder {
 public:
    Builder(const mlgpu_scene_desc& d, HostScene* hs) : d_(d), hs_(hs) {}

    void run() {
        const int32_t n = d_.n_frames;
        t2w_.resize(n);
        children_.assign(n, {});
        std::vector<double> radius(n);
        for (int32_t i = 0; i < n; i++) {
            double prm[8];
            radius[i] = leaf_params(d_.kind[i], d_.params + MLGPU_NUM_PARAMS * i, prm);
            if (d_.mother[i] >= 0) children_[d_.mother[i]].push_back(i);
        }
        {
            const ReferencePoses rp(d_, radius);
            for (int32_t i = 0; i < n; i++) t2w_[i] = rp.frame2world(i);
            rank_ = rp.candidate_rank(n);
        }
        scope_.assign(n, 0);
        scope_[0] = 1;
        for (int32_t i = 0; i < n; i++)
            if (d_.allowed_frame && d_.allowed_frame[i] >= 0) scope_[d_.allowed_frame[i]] = 1;
        build_functions();
        bnodes_.reserve(2 * n);
        const int32_t world = build_scope(0);
        if (internal(world)) emit_pair(world, 1);
        // scopes in frame order; a scope whose tree is a single leaf (or
        // nothing) gets a node of its own to start at
        hs_->scope_of_frame.assign(n, -1);
        for (int32_t f = 0; f < n; f++) {
            if (!scope_[f]) continue;
            const int32_t r = scope_bnode_.at(f);
            int32_t entry;
            if (internal(r)) {
                auto it = pair_of_.find(r);
                entry = it != pair_of_.end() ? it->second : emit_pair(r, 1);
            } else {
                entry = new_pair(1);
                set_child(entry, 0, r, child_code(r, 2));
                set_child(entry, 1, -1, kNoChild);
            }
            hs_->scope_of_frame[f] = static_cast<int32_t>(hs_->scopes.size());
            Scope sc{};
            sc.entry = entry;
            hs_->scopes.push_back(sc);
            scope_box_.push_back(bnodes_[r].box);
        }
        hs_->root_scope = hs_->scope_of_frame[0];
        if (hs_->pairs.size() >= (size_t{1} << 29) || hs_->geo.size() >= (size_t{1} << 30))
            throw std::invalid_argument("scene has too many surfaces for the node codes");
        if (hs_->tree_depth > kMaxTreeDepth)
            throw std::invalid_argument("scene tree is " + std::to_string(hs_->tree_depth) +
                                        " levels deep, more than the traversal stack holds (" +
                                        std::to_string(kMaxTreeDepth) + ")");
        // leaf physics needs the scopes
        for (size_t l = 0; l < hs_->phys.size(); l++) {
            LeafPhys& ph = hs_->phys[l];
            const int32_t fr = ph.frame;
            const int32_t a = d_.allowed_frame ? d_.allowed_frame[fr] : -1;
            ph.allowed_scope = a >= 0 ? hs_->scope_of_frame[a] : -1;
            ph.root_scope = hs_->root_scope;
        }
        hot_first();
        inflate();
    }

 private:
    const mlgpu_scene_desc& d_;
    HostScene* hs_;
    std::vector<M34> t2w_;
    std::vector<int32_t> rank_;
    std::vector<std::vector<int32_t>> children_;
    std::vector<char> scope_;
    std::vector<BNode> bnodes_;
    std::unordered_map<int32_t, int32_t> scope_bnode_;
    std::unordered_map<int32_t, int32_t> pair_of_;   // internal BNode -> PairNode
    std::unordered_map<int32_t, int32_t> leaf_of_;   // frame -> leaf
    std::vector<std::array<Box, 2>> pair_box_;       // exact child boxes per PairNode
    std::vector<Box> scope_box_;
    std::vector<std::array<double, 15>> leaf_obb_;   // rotation, centre, half extents

    int32_t make_leaf(int32_t f) {
        BNode b;
        double prm[8];
        const double* p = d_.params + MLGPU_NUM_PARAMS * f;
        const double radius = leaf_params(d_.kind[f], p, prm);
        const Box lb = local_box(d_.kind[f], p, prm, radius);
        double outline[16][2];
        const int nv = flat_outline(d_.kind[f], p, outline);
        if (nv > 0) {
            // a flat leaf: the world box of its outline, much tighter than
            // that of its object-frame box when the leaf is slanted
            const M34& m = t2w_[f];
            for (int c = 0; c < nv; c++) {
                double w[3];
                for (int i = 0; i < 3; i++) w[i] = m.T[i][0] * outline[c][0] + m.T[i][1] * outline[c][1] + m.T[i][3];
                b.box.grow(w);
            }
        } else if (!lb.empty()) {
            const M34& m = t2w_[f];
            for (int c = 0; c < 8; c++) {
                const double q[3] = {(c & 1) ? lb.hi[0] : lb.lo[0], (c & 2) ? lb.hi[1] : lb.lo[1],
                                     (c & 4) ? lb.hi[2] : lb.lo[2]};
                double w[3];
                for (int i = 0; i < 3; i++) w[i] = m.T[i][0] * q[0] + m.T[i][1] * q[1] + m.T[i][2] * q[2] + m.T[i][3];
                b.box.grow(w);
            }
        }
        b.leaf_frame = f;
        bnodes_.push_back(b);
        return static_cast<int32_t>(bnodes_.size()) - 1;
    }

    // the leaves under frame f, each as an item of f's tree; a restriction
    // frame below f gets a tree of its own as well, for the propagation
    // restricted to it, while f's tree holds its leaves one by one (so the
    // boxes of f's tree are not bounded by the restricted sub-assemblies)
    void collect(int32_t f, std::vector<int32_t>* items) {
        for (int32_t c : children_[f]) {
            if (scope_[c]) build_scope(c);
            if (children_[c].empty()) {
                // a frame without a surface never reports a hit (Frame.cpp:172-179)
                if (d_.kind[c] != MLGPU_KIND_FRAME) items->push_back(make_leaf(c));
            } else {
                collect(c, items);
            }
        }
    }

    int32_t build_scope(int32_t f) {
        std::vector<int32_t> items;
        if (children_[f].empty()) {
            items.push_back(make_leaf(f));
        } else {
            collect(f, &items);
        }
        int32_t root;
        if (items.empty()) {
            BNode b;  // empty box: never entered
            bnodes_.push_back(b);
            root = static_cast<int32_t>(bnodes_.size()) - 1;
        } else {
            root = bvh(items, 0, items.size());
        }
        scope_bnode_[f] = root;
        return root;
    }

    // binary tree over items[lo, hi) with the surface-area heuristic, binned
    // along the three axes of the centroid bounds
    int32_t bvh(std::vector<int32_t>& items, size_t lo, size_t hi, int depth = 0) {
        if (hi - lo == 1) return items[lo];
        Box cb, all;
        for (size_t i = lo; i < hi; i++) {
            const Box& b = bnodes_[items[i]].box;
            all.grow(b);
            if (!b.empty()) {
                const double c[3] = {b.centre(0), b.centre(1), b.centre(2)};
                cb.grow(c);
            }
        }
        constexpr int kBins = 32;
        int best_axis = -1, best_split = 0;
        double best_cost = HUGE_VAL;
        // deep down the tree, balanced splits bound the depth
        const bool balanced = depth >= kBalancedFromDepth;
        for (int axis = 0; axis < 3 && !cb.empty() && !balanced; axis++) {
            const double ext = cb.hi[axis] - cb.lo[axis];
            if (!(ext > 0.0)) continue;
            Box bins[kBins];
            int count[kBins] = {0};
            for (size_t i = lo; i < hi; i++) {
                const Box& b = bnodes_[items[i]].box;
                const double c = b.empty() ? cb.lo[axis] : b.centre(axis);
                int k = static_cast<int>((c - cb.lo[axis]) / ext * kBins);
                k = std::min(std::max(k, 0), kBins - 1);
                bins[k].grow(b);
                count[k]++;
            }
            Box left[kBins];
            int nl[kBins];
            Box acc;
            int n = 0;
            for (int k = 0; k < kBins; k++) { acc.grow(bins[k]); n += count[k]; left[k] = acc; nl[k] = n; }
            acc = Box();
            n = 0;
            for (int k = kBins - 1; k > 0; k--) {
                acc.grow(bins[k]);
                n += count[k];
                if (nl[k - 1] == 0 || n == 0) continue;
                const double cost = left[k - 1].area() * nl[k - 1] + acc.area() * n;
                if (cost < best_cost) { best_cost = cost; best_axis = axis; best_split = k; }
            }
        }
        size_t mid;
        if (balanced && !cb.empty()) {
            int axis = 0;
            for (int k = 1; k < 3; k++)
                if (cb.hi[k] - cb.lo[k] > cb.hi[axis] - cb.lo[axis]) axis = k;
            mid = lo + (hi - lo) / 2;
            std::nth_element(items.begin() + lo, items.begin() + mid, items.begin() + hi, [&](int32_t a, int32_t b) {
                const Box& ba = bnodes_[a].box;
                const Box& bb = bnodes_[b].box;
                const double ca = ba.empty() ? cb.lo[axis] : ba.centre(axis);
                const double cbb = bb.empty() ? cb.lo[axis] : bb.centre(axis);
                return ca < cbb;
            });
        } else if (best_axis < 0) {
            mid = lo + (hi - lo) / 2;  // coincident centroids: any split
        } else {
            const int axis = best_axis;
            const double ext = cb.hi[axis] - cb.lo[axis];
            auto it = std::partition(items.begin() + lo, items.begin() + hi, [&](int32_t q) {
                const Box& b = bnodes_[q].box;
                const double c = b.empty() ? cb.lo[axis] : b.centre(axis);
                int k = static_cast<int>((c - cb.lo[axis]) / ext * kBins);
                k = std::min(std::max(k, 0), kBins - 1);
                return k < best_split;
            });
            mid = static_cast<size_t>(it - items.begin());
            if (mid == lo || mid == hi) mid = lo + (hi - lo) / 2;
        }
        const int32_t left = bvh(items, lo, mid, depth + 1);
        const int32_t right = bvh(items, mid, hi, depth + 1);
        BNode b;
        b.box = bnodes_[left].box;
        b.box.grow(bnodes_[right].box);
        b.kids = {left, right};
        bnodes_.push_back(b);
        return static_cast<int32_t>(bnodes_.size()) - 1;
    }

    static constexpr int kBalancedFromDepth = 16;

    bool internal(int32_t bi) const { return !bnodes_[bi].kids.empty(); }

    int32_t new_pair(int depth) {
        hs_->pairs.push_back(PairNode{});
        pair_box_.push_back({Box(), Box()});
        hs_->tree_depth = std::max(hs_->tree_depth, depth);
        return static_cast<int32_t>(hs_->pairs.size()) - 1;
    }

    void set_child(int32_t pair, int k, int32_t bi, int32_t code) {
        PairNode& p = hs_->pairs[pair];
        (k == 0 ? p.c0 : p.c1) = code;
        pair_box_[pair][k] = (code == kNoChild || bi < 0) ? Box() : bnodes_[bi].box;
    }

    // code of BNode bi as a child: a leaf, nothing, or its (emitted) node
    int32_t child_code(int32_t bi, int depth) {
        const BNode& b = bnodes_[bi];
        if (b.leaf_frame >= 0) {
            if (d_.kind[b.leaf_frame] == MLGPU_KIND_FRAME) return kNoChild;  // never reports a hit
            auto it = leaf_of_.find(b.leaf_frame);
            const int32_t leaf = it != leaf_of_.end() ? it->second : (leaf_of_[b.leaf_frame] = add_leaf(b.leaf_frame));
            return -1 - leaf;
        }
        if (b.kids.empty()) return kNoChild;
        return emit_pair(bi, depth);
    }

    // depth-first emission: a node's subtree follows it in memory
    int32_t emit_pair(int32_t bi, int depth) {
        const int32_t idx = new_pair(depth);
        pair_of_[bi] = idx;
        const std::vector<int32_t> kids = bnodes_[bi].kids;
        for (int k = 0; k < 2; k++) set_child(idx, k, kids[k], child_code(kids[k], depth + 1));
        return idx;
    }

    int32_t add_leaf(int32_t f) {
        LeafGeo g;
        const M34& m = t2w_[f];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) g.rot[3 * i + j] = m.T[i][j];
        g.tx = m.T[0][3]; g.ty = m.T[1][3]; g.tz = m.T[2][3];
        // the constant part of HomTra3::position_inverse (merlict/HomTra3.cpp:202-216)
        g.ti[0] = m.T[0][0] * m.T[0][3] + m.T[1][0] * m.T[1][3] + m.T[2][0] * m.T[2][3];
        g.ti[1] = m.T[0][1] * m.T[0][3] + m.T[1][1] * m.T[1][3] + m.T[2][1] * m.T[2][3];
        g.ti[2] = m.T[0][2] * m.T[0][3] + m.T[1][2] * m.T[1][3] + m.T[2][2] * m.T[2][3];
        leaf_params(d_.kind[f], d_.params + MLGPU_NUM_PARAMS * f, g.p);
        g.kind = d_.kind[f];
        g.rank = rank_[f];
        LeafPhys ph;
        for (int s = 0; s < MLGPU_NUM_SURFACE_SLOTS; s++) {
            int32_t fn = d_.surface[MLGPU_NUM_SURFACE_SLOTS * f + s];
            if (fn < 0) {
                if (s <= MLGPU_INNER_REFLECTION) fn = def_refl_;
                else if (s <= MLGPU_INNER_REFRACTION) fn = def_refr_;
                else fn = def_abs_;
            }
            if (fn > kMaxFunctions) throw std::invalid_argument("scene has more functions than a leaf can index");
            ph.surf[s] = static_cast<int16_t>(fn);
        }
        ph.allowed_scope = -1;
        ph.root_scope = -1;
        ph.frame = f;
        ph.sensor = d_.sensor_id ? d_.sensor_id[f] : -1;
        ph.pad = 0;
        hs_->geo.push_back(g);
        hs_->phys.push_back(ph);
        {
            // object-frame box of the leaf, as centre and half extents in the world
            double prm[8];
            const double* p = d_.params + MLGPU_NUM_PARAMS * f;
            const double radius = leaf_params(d_.kind[f], p, prm);
            const Box lb = local_box(d_.kind[f], p, prm, radius);
            std::array<double, 15> ob{};
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) ob[3 * i + j] = m.T[j][i];
            if (!lb.empty()) {
                const double cl[3] = {lb.centre(0), lb.centre(1), lb.centre(2)};
                for (int i = 0; i < 3; i++) {
                    ob[9 + i] = m.T[i][0] * cl[0] + m.T[i][1] * cl[1] + m.T[i][2] * cl[2] + m.T[i][3];
                    ob[12 + i] = 0.5 * (lb.hi[i] - lb.lo[i]);
                }
            } else {
                for (int i = 0; i < 3; i++) ob[12 + i] = -1.0;  // never entered
            }
            leaf_obb_.push_back(ob);
        }
        hs_->leaf_frame.push_back(f);
        return static_cast<int32_t>(hs_->geo.size()) - 1;
    }

    int32_t def_refl_ = -1, def_refr_ = -1, def_abs_ = -1;

    void add_function(const double* x, const double* y, int n, int flags) {
        struct P { double x, y, slope; };
        std::vector<P> pts(n);
        for (int i = 0; i < n; i++) {
            pts[i].x = x[i];
            pts[i].y = y[i];
            pts[i].slope = (i != n - 1) ? (y[i] - y[i + 1]) / (x[i] - x[i + 1]) : 0.0;
        }
        std::sort(pts.begin(), pts.end(), [](const P& a, const P& b) { return a.x < b.x; });
        for (int i = 0; i + 1 < n; i++)
            if (pts[i].x == pts[i + 1].x) throw std::invalid_argument("function has a duplicate argument");
        FuncHead h;
        h.lower = pts.front().x;
        h.upper = pts.back().x;
        h.start = static_cast<int32_t>(hs_->points.size());
        h.n = n;
        h.flags = flags;
        h.pad = 0;
        for (const P& p : pts) {
            FuncPoint fp;
            fp.x = p.x; fp.y = p.y; fp.slope = p.slope; fp.pad = 0.0;
            hs_->points.push_back(fp);
        }
        // bucket index: buckets no wider than the narrowest segment (at most
        // 4096 per function; wider buckets only cost more steps), each
        // pointing at the segment of its lower edge
        double narrowest = INFINITY;
        for (int i = 0; i + 1 < n; i++) narrowest = std::min(narrowest, pts[i + 1].x - pts[i].x);
        const double range = h.upper - h.lower;
        int32_t nb = 1;
        if (n > 2 && std::isfinite(range) && range > 0.0)
            nb = static_cast<int32_t>(std::min(4096.0, std::max(1.0, std::ceil(range / narrowest))));
        h.n_buckets = nb;
        h.scale = (range > 0.0 && std::isfinite(range)) ? nb / range : 0.0;
        h.bucket_start = static_cast<int32_t>(hs_->buckets.size());
        int32_t seg = 0;
        for (int32_t b = 0; b < nb; b++) {
            const double edge = h.lower + b / h.scale;
            while (seg + 2 < n && pts[seg + 1].x <= edge) seg++;
            hs_->buckets.push_back(n >= 2 ? seg : 0);
        }
        hs_->funcs.push_back(h);
    }

    void build_functions() {
        for (int32_t f = 0; f < d_.n_functions; f++) {
            const int32_t a = d_.func_start[f], b = d_.func_start[f + 1];
            add_function(d_.func_x + a, d_.func_y + a, b - a, 0);
        }
        // merlict's defaults (merlict/SurfaceEntity.cpp:12-33)
        const double x[2] = {200e-9, 1200e-9};
        const double r0[2] = {0.0, 0.0}, r1[2] = {1.0, 1.0};
        const double inf[2] = {INFINITY, INFINITY};
        def_refl_ = d_.n_functions;
        add_function(x, r0, 2, FUNC_DEFAULT_REFLECTION);
        def_refr_ = d_.n_functions + 1;
        add_function(x, r1, 2, FUNC_DEFAULT_REFRACTION);
        def_abs_ = d_.n_functions + 2;
        // Func1 slope of an infinite table is NaN, so merlict's default
        // absorption evaluates to NaN and never absorbs; flagged explicitly.
        add_function(x, inf, 2, FUNC_DEFAULT_ABSORPTION);
        hs_->default_refraction = def_refr_;
        // HexagonalPrismZ unit vectors V, W = (+-sin(2 pi/3), cos(2 pi/3)) (HexagonalPrismZ.h:10-16)
        hs_->hex_sin = std::sin(2. / 3. * M_PI);
        hs_->hex_cos = std::cos(2. / 3. * M_PI);
    }

    // Moves the top kHotNodes nodes of the world scope's tree (breadth first)
    // to the front of the node array, where the traversal kernels copy them
    // into LDS; every other node keeps its order.
    void hot_first() {
        const size_t m = hs_->pairs.size();
        std::vector<int32_t> order;
        order.reserve(m);
        std::vector<char> taken(m, 0);
        std::deque<int32_t> q{hs_->scopes[hs_->root_scope].entry};
        while (!q.empty() && order.size() < static_cast<size_t>(kHotNodes)) {
            const int32_t c = q.front();
            q.pop_front();
            if (c < 0 || taken[c]) continue;
            taken[c] = 1;
            order.push_back(c);
            q.push_back(hs_->pairs[c].c0);
            q.push_back(hs_->pairs[c].c1);
        }
        hs_->hot_count = static_cast<int32_t>(order.size());
        for (size_t i = 0; i < m; i++)
            if (!taken[i]) order.push_back(static_cast<int32_t>(i));
        std::vector<int32_t> new_of(m);
        for (size_t k = 0; k < m; k++) new_of[order[k]] = static_cast<int32_t>(k);
        std::vector<PairNode> pairs(m);
        std::vector<std::array<Box, 2>> boxes(m);
        for (size_t k = 0; k < m; k++) {
            pairs[k] = hs_->pairs[order[k]];
            boxes[k] = pair_box_[order[k]];
            if (pairs[k].c0 >= 0) pairs[k].c0 = new_of[pairs[k].c0];
            if (pairs[k].c1 >= 0) pairs[k].c1 = new_of[pairs[k].c1];
        }
        for (Scope& sc : hs_->scopes) sc.entry = new_of[sc.entry];
        hs_->pairs.swap(pairs);
        pair_box_.swap(boxes);
        for (auto& kv : pair_of_) kv.second = new_of[kv.second];
    }

    // IEEE half nearest to v (|v| <= 1, ties to even) and its value.
    static uint16_t to_half(const double v) {
        const uint16_t sign = std::signbit(v) ? 0x8000u : 0u;
        const double a = std::fabs(v);
        if (a == 0.0) return sign;
        int e;
        std::frexp(a, &e);  // a in [2^(e-1), 2^e)
        int E = std::max(e - 1, -14);
        double q = std::nearbyint(a / std::ldexp(1.0, E - 10));
        if (q >= 2048.0) {
            q = 1024.0;
            E++;
        }
        if (q < 1024.0) return static_cast<uint16_t>(sign | static_cast<uint16_t>(q));  // subnormal
        if (E > 15) throw std::logic_error("leaf box rotation entry out of half range");
        return static_cast<uint16_t>(sign | static_cast<uint16_t>((E + 15) << 10) | static_cast<uint16_t>(q - 1024.0));
    }
    static double from_half(const uint16_t h) {
        const int e = (h >> 10) & 31;
        const double mant = h & 1023;
        const double mag = e == 0 ? std::ldexp(mant, -24) : std::ldexp(1024.0 + mant, e - 25);
        return (h & 0x8000u) ? -mag : mag;
    }

    // The node's grid and its children's faces on it (PairNode): per axis
    // the origin is the children's lowest face rounded down to a float with
    // 15 mantissa bits, the step the least power of two that spans their
    // highest face in 254 steps; faces are rounded outwards on the grid and
    // moved one more step where the kernel's fmaf would round them back
    // inside the float box.
    static void quantize_pair(const float (&lo)[2][3], const float (&hi)[2][3], const std::array<bool, 2> live,
                              PairNode& p) {
        auto bits = [](float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; };
        auto from = [](uint32_t u) { float x; std::memcpy(&x, &u, 4); return x; };
        uint8_t q[12] = {};
        for (int a = 0; a < 3; a++) {
            float ulo = HUGE_VALF, uhi = -HUGE_VALF;
            for (int c = 0; c < 2; c++)
                if (live[c]) {
                    ulo = std::min(ulo, lo[c][a]);
                    uhi = std::max(uhi, hi[c][a]);
                }
            if (!(ulo <= uhi)) ulo = uhi = 0.0f;  // no child: any grid
            // origin: ulo with its low 8 mantissa bits cleared, rounded down
            uint32_t ob = bits(ulo);
            if ((ob & 0xFFu) != 0u) ob = (ob & 0x80000000u) ? (ob | 0xFFu) + 1u : ob;
            ob &= 0xFFFFFF00u;
            const float o = from(ob);
            if (!(static_cast<double>(o) <= ulo)) throw std::logic_error("pair node grid origin above its boxes");
            // step 2^k: o + 254 * 2^k >= uhi
            int k = -126;
            const double span = static_cast<double>(uhi) - static_cast<double>(o);
            if (span > 0.0) {
                int ex;
                std::frexp(span / 254.0, &ex);  // span / 254 < 2^ex
                k = std::max(-126, ex - 1);
                while (static_cast<double>(o) + 254.0 * std::ldexp(1.0, k) < uhi) k++;
            }
            if (k > 127) throw std::invalid_argument("scene box too large for the node grid");
            const float step = from(static_cast<uint32_t>(k + 127) << 23);
            p.g[a] = ob | static_cast<uint32_t>(k + 127);
            for (int c = 0; c < 2; c++) {
                if (!live[c]) continue;
                double fl = std::floor((static_cast<double>(lo[c][a]) - o) / step);
                double fh = std::ceil((static_cast<double>(hi[c][a]) - o) / step);
                int kl = static_cast<int>(std::min(255.0, std::max(0.0, fl)));
                int kh = static_cast<int>(std::min(255.0, std::max(0.0, fh)));
                while (kl > 0 && std::fmaf(static_cast<float>(kl), step, o) > lo[c][a]) kl--;
                while (kh < 255 && std::fmaf(static_cast<float>(kh), step, o) < hi[c][a]) kh++;
                if (std::fmaf(static_cast<float>(kl), step, o) > lo[c][a] ||
                    std::fmaf(static_cast<float>(kh), step, o) < hi[c][a])
                    throw std::logic_error("pair node face off its grid");
                q[6 * c + a] = static_cast<uint8_t>(kl);
                q[6 * c + 3 + a] = static_cast<uint8_t>(kh);
            }
        }
        for (int w = 0; w < 3; w++)
            p.q[w] = static_cast<uint32_t>(q[4 * w]) | static_cast<uint32_t>(q[4 * w + 1]) << 8 |
                     static_cast<uint32_t>(q[4 * w + 2]) << 16 | static_cast<uint32_t>(q[4 * w + 3]) << 24;
    }

    // Round boxes outwards to float and inflate them so the FP32 slab test
    // stays conservative: the ray origin is rounded to float and the slab
    // arithmetic rounds, both bounded by a few ulp of the scene's extent.
    void inflate() {
        double extent = 1.0;
        for (const Box& b : scope_box_)
            if (!b.empty())
                for (int k = 0; k < 3; k++) extent = std::max({extent, std::fabs(b.lo[k]), std::fabs(b.hi[k])});
        const double eps = 4e-6 * extent;
        hs_->eps = eps;
        auto put = [eps](const Box& b, float* lo, float* hi) {
            if (b.empty()) {
                // nothing inside: a point box (its code says there is no child)
                for (int k = 0; k < 3; k++) lo[k] = hi[k] = 0.0f;
                return;
            }
            for (int k = 0; k < 3; k++) {
                float l = static_cast<float>(b.lo[k] - eps), h = static_cast<float>(b.hi[k] + eps);
                if (static_cast<double>(l) > b.lo[k] - eps) l = std::nextafter(l, -HUGE_VALF);
                if (static_cast<double>(h) < b.hi[k] + eps) h = std::nextafter(h, HUGE_VALF);
                lo[k] = l;
                hi[k] = h;
            }
        };
        for (size_t i = 0; i < hs_->pairs.size(); i++) {
            PairNode& p = hs_->pairs[i];
            float lo[2][3], hi[2][3];
            put(pair_box_[i][0], lo[0], hi[0]);
            put(pair_box_[i][1], lo[1], hi[1]);
            quantize_pair(lo, hi, {!pair_box_[i][0].empty(), !pair_box_[i][1].empty()}, p);
        }
        for (size_t i = 0; i < hs_->scopes.size(); i++) put(scope_box_[i], hs_->scopes[i].lo, hs_->scopes[i].hi);
        // oriented leaf boxes: half extents grow by the same bound, rounded up
        hs_->boxes.resize(leaf_obb_.size());
        for (size_t i = 0; i < leaf_obb_.size(); i++) {
            const auto& ob = leaf_obb_[i];
            LeafBox& lb = hs_->boxes[i];
            uint16_t mh[10] = {};
            double row_err[3] = {0.0, 0.0, 0.0};
            for (int k = 0; k < 9; k++) {
                mh[k] = to_half(ob[k]);
                row_err[k / 3] = std::max(row_err[k / 3], std::fabs(from_half(mh[k]) - ob[k]));
            }
            for (int w = 0; w < 5; w++) lb.m[w] = static_cast<uint32_t>(mh[2 * w]) | static_cast<uint32_t>(mh[2 * w + 1]) << 16;
            // |(m~ - m)(x - c)|_k <= row_err_k |x - c|_1 <= row_err_k sqrt(3) |h|_2 for x in the box
            double diag = 0.0;
            for (int k = 0; k < 3; k++)
                if (ob[12 + k] >= 0.0) diag += (ob[12 + k] + eps) * (ob[12 + k] + eps);
            diag = std::sqrt(3.0 * diag) * (1.0 + 1e-6);
            for (int k = 0; k < 3; k++) {
                lb.c[k] = static_cast<float>(ob[9 + k]);
                if (ob[12 + k] < 0.0) {
                    lb.h[k] = -1.0f;
                    continue;
                }
                const double h = ob[12 + k] + eps + row_err[k] * diag;
                float hf = static_cast<float>(h);
                if (static_cast<double>(hf) < h) hf = std::nextafter(hf, HUGE_VALF);
                lb.h[k] = hf;
            }
            lb.pad = 0u;
        }
    }
};

}  // namespace

void build_host_scene(const mlgpu_scene_desc& d, HostScene* hs) {
    validate_desc(d);
    Builder b(d, hs);
    b.run();
}

void pose_matrix(const double* pos, int mode, const double* rot, double* out) {
    const M34 m = pose(pos, mode, rot);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++) out[4 * r + c] = m.T[r][c];
}

void frame_poses(const mlgpu_scene_desc& d, double* out) {
    validate_desc(d);
    std::vector<double> radius(d.n_frames);
    for (int32_t i = 0; i < d.n_frames; i++) {
        double prm[8];
        radius[i] = leaf_params(d.kind[i], d.params + MLGPU_NUM_PARAMS * i, prm);
    }
    const ReferencePoses rp(d, radius);
    for (int32_t i = 0; i < d.n_frames; i++) {
        const M34& m = rp.frame2world(i);
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 4; c++) out[12 * i + 4 * r + c] = m.T[r][c];
    }
}

}  // namespace mlgpu
