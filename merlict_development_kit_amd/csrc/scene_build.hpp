// Host representation of a flattened scene before upload.
#pragma once

#include <array>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "device_scene.hpp"
#include "scene.hpp"

namespace mlgpu {

struct HostScene {
    std::vector<PairNode> pairs;
    std::vector<Scope> scopes;
    int32_t hot_count = 0;                           // nodes moved to the front (kHotNodes at most)
    int32_t tree_depth = 0;                          // longest chain of pair nodes
    std::vector<LeafGeo> geo;
    std::vector<LeafBox> boxes;
    std::vector<LeafPhys> phys;
    std::vector<int32_t> leaf_frame;
    std::vector<FuncHead> funcs;
    std::vector<FuncPoint> points;
    std::vector<int32_t> buckets;
    std::vector<int32_t> scope_of_frame;  // per description frame, -1 if not a scope
    int32_t root_scope = -1;
    int32_t default_refraction = -1;
    double eps = 0.0;
    double hex_sin = 0.0, hex_cos = 0.0;
};

// Flatten a description (throws std::invalid_argument on bad input).
void build_host_scene(const mlgpu_scene_desc& d, HostScene* hs);

// frame2world of every frame (12 doubles each, row-major 3 x 4), as merlict
// computes them after clustering (throws std::invalid_argument on bad input).
void frame_poses(const mlgpu_scene_desc& d, double* out);
// HomTra3::set_transformation(Rot3, Vec3) of one pose (merlict/HomTra3.cpp:
// 43-112): out holds the row-major 3 x 4 transform.
void pose_matrix(const double* pos, int mode, const double* rot, double* out);

}  // namespace mlgpu
