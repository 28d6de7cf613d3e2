// Device layout of a flattened scene (shared by the host builder and the
// HIP kernels).
//
// Traversal hierarchy: a bounding-volume tree stored in depth-first order
// with skip indices, so a subtree is the contiguous range [i, skip[i]) and
// traversal needs no stack.  Every frame that can be a propagation root (the
// world and each restriction frame, merlict/SurfaceEntity.cpp:59-65) owns a
// node whose range holds exactly its leaves; below such a frame the library
// builds its own binary tree of axis-aligned boxes over the leaves (surface-
// area heuristic) instead of merlict's bounding spheres and octant clustering
// (merlict/Frame.cpp:190-255, RayAndFrame.cpp:16-107).  Any conservative
// hierarchy yields the same closest hit, because merlict keeps the first
// minimum over all candidates (RayAndFrame.cpp:160-173).
#pragma once

#include <cstdint>

namespace mlgpu {

// 32 B per node: world-frame box (inflated by the float rounding bound)
// and (skip, leaf index or -1).
struct alignas(16) Node {
    float lo[3];
    int32_t skip;
    float hi[3];
    int32_t leaf;
};

enum LeafKind : int32_t {
    LK_CAP_HEX = 1,
    LK_CAP_CYL = 2,
    LK_CAP_RECT = 3,
    LK_PLANE = 4,
    LK_DISC = 5,
    LK_HEX_PLANE = 6,
    LK_ANNULUS = 7,
    LK_DUAL_SPHERICAL = 8,
    LK_CYLINDER = 9,
    LK_TRIANGLE = 10,
    LK_SPHERE = 11,
};

// Geometry of a leaf (208 B), read by every leaf test.
//   rot    frame2world rotation, row major (HomTra3 T[i][j], j < 3)
//   t      frame2world translation (T[i][3])
//   ti     R^T t, the constant of HomTra3::position_inverse
//   p      kind parameters, see the kernels
struct alignas(16) LeafGeo {
    double rot[9];
    double tx, ty, tz;
    double ti[3];
    double p[8];
    int32_t kind;
    int32_t pad0;
};

// Physics of a leaf, read once per bounce (48 B).
struct alignas(16) LeafPhys {
    int32_t surf[6];       // function index per surface slot (defaults resolved)
    int32_t allowed_node;  // restriction scope node, -1 none
    int32_t root_node;     // scope node of the tree root
    int32_t frame;         // description frame index
    int32_t sensor;        // sensor id or -1
    int32_t pad0, pad1;
};

// Func1 table: points sorted by x with the slope merlict computes before
// sorting (merlict/function/Func1.cpp:25-36).
struct alignas(16) FuncHead {
    double lower, upper;  // valid arguments: lower <= x < upper
    int32_t start, n;
    int32_t flags;        // FUNC_DEFAULT_* identity of merlict's defaults
    int32_t pad;
};

enum FuncFlags : int32_t {
    FUNC_DEFAULT_REFLECTION = 1,
    FUNC_DEFAULT_REFRACTION = 2,
    FUNC_DEFAULT_ABSORPTION = 4,
};

struct alignas(16) FuncPoint {
    double x;
    double y;
    double slope;
    double pad;
};

struct DeviceScene {
    const Node* nodes;
    const LeafGeo* geo;
    const LeafPhys* phys;
    const FuncHead* funcs;
    const FuncPoint* points;
    int32_t num_nodes;
    int32_t num_leaves;
    int32_t root_node;           // node of frame 0
    int32_t default_refraction;  // function index of merlict's DEFAULT_REFRACTION
    double hex_sin, hex_cos;     // HexagonalPrismZ unit-vector components
    float eps;                   // box inflation (absolute, metres)
};

}  // namespace mlgpu
