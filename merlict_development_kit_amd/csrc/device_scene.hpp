// Device layout of a flattened scene (shared by the host builder and the
// HIP kernels).
//
// Traversal hierarchy: a binary tree of axis-aligned boxes whose nodes
// store the boxes of both children, so one 32-byte fetch tests two boxes;
// traversal visits the nearer child first and keeps the farther one on a
// short per-lane stack.  Every frame that can be a propagation root (the
// world and each restriction frame, merlict/SurfaceEntity.cpp:59-65) is a
// scope: a subtree that holds exactly its leaves.  Below a scope the
// library builds its own tree (surface-area heuristic) instead of merlict's
// bounding spheres and octant clustering (merlict/Frame.cpp:190-255,
// RayAndFrame.cpp:16-107).  Any conservative hierarchy yields the same
// closest hit, because merlict keeps the first minimum over all candidates
// (RayAndFrame.cpp:160-173).
#pragma once

#include <cstdint>
#include <climits>

namespace mlgpu {

// Child codes of a PairNode: >= 0 another PairNode, kNoChild nothing, any
// other negative value the leaf -1 - code.
constexpr int32_t kNoChild = INT32_MIN;

// 32 B per node: the world-frame boxes of both children (inflated by the
// float rounding bound), quantised to one byte per face on a grid of the
// node's own, and their codes.  Axis a of the grid has origin o_a, a float
// whose low 8 mantissa bits are zero, kept in the high 24 bits of g[a], and
// step 2^(e_a - 127), e_a in the low 8 bits; the face of grid index k lies
// at fmaf(k, step, o) (one rounding).  Byte 6 c + 3 f + a of q is child c's
// lower (f = 0) or upper (f = 1) face on axis a.  The builder rounds lower
// faces down and upper faces up past that rounding, so a quantised box
// holds its child's float box; half the bytes of float faces, which is half
// of what a node visit costs the vector memory pipeline.
struct alignas(32) PairNode {
    uint32_t g[3];
    uint32_t q[3];
    int32_t c0;
    int32_t c1;
};

// A propagation root: the box of all its leaves and the node to start at.
struct alignas(16) Scope {
    float lo[3];
    int32_t entry;
    float hi[3];
    int32_t pad;
};

// Deepest node chain a scene may have: the traversal stack holds at most
// one entry per level.
constexpr int kMaxTreeDepth = 28;

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

// Geometry of a leaf (192 B), read by every leaf test.
//   rot    frame2world rotation, row major (HomTra3 T[i][j], j < 3)
//   t      frame2world translation (T[i][3])
//   ti     R^T t, the constant of HomTra3::position_inverse
//   p      kind parameters, see the kernels
//   rank   position in merlict's candidate order, which breaks exact ties
struct alignas(16) LeafGeo {
    double rot[9];
    double tx, ty, tz;
    double ti[3];
    double p[8];
    int32_t kind;
    int32_t rank;
};

// Float oriented box of a leaf (48 B): a point x of the world lies in it when
// |m (x - c)|_k <= h_k for k = x, y, z.  It holds every point where the leaf
// can report a hit, inflated like the tree's boxes and, because the rotation
// m is kept in half precision, by the most that rounding can move a point of
// the box (the row's largest rounding error times the box's diagonal), so a
// ray that misses it (or meets it only behind the best hit so far) needs no
// FP64 solve.
struct alignas(16) LeafBox {
    float c[3];     // world centre
    float h[3];     // half extents
    uint32_t m[5];  // world -> object rotation, row major: 9 IEEE halves, two per word, low half first
    uint32_t pad;
};

// Physics of a leaf, read once per bounce (32 B).
struct alignas(16) LeafPhys {
    int16_t surf[6];       // function index per surface slot (defaults resolved; a scene has at most 32,767)
    int32_t allowed_scope; // scope of the restriction frame, -1 none
    int32_t root_scope;    // scope of the tree root
    int32_t frame;         // description frame index
    int32_t sensor;        // sensor id or -1
    int32_t pad;
};
constexpr int32_t kMaxFunctions = 32767;

// Func1 table: points sorted by x with the slope merlict computes before
// sorting (merlict/function/Func1.cpp:25-36), and a bucket index over
// [lower, upper): bucket b = floor((x - lower) * scale) holds the segment
// that contains the bucket's lower edge, and buckets are no wider than the
// narrowest segment, so the segment of x is found in a step or two from
// there instead of by a binary search.
struct alignas(16) FuncHead {
    double lower, upper;  // valid arguments: lower <= x < upper
    double scale;         // buckets per unit of x
    int32_t start, n;
    int32_t flags;        // FUNC_DEFAULT_* identity of merlict's defaults
    int32_t bucket_start; // first bucket in DeviceScene::buckets
    int32_t n_buckets;
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
    const PairNode* pairs;
    const Scope* scopes;
    const LeafGeo* geo;
    const LeafBox* boxes;
    const LeafPhys* phys;
    const FuncHead* funcs;
    const FuncPoint* points;
    const int32_t* buckets;  // segment index per bucket of every function
    int32_t num_pairs;
    int32_t num_leaves;
    int32_t root_scope;          // scope of frame 0
    int32_t default_refraction;  // function index of merlict's DEFAULT_REFRACTION
    double hex_sin, hex_cos;     // HexagonalPrismZ unit-vector components
    float eps;                   // box inflation (absolute, metres)
    int32_t num_funcs, num_points, num_buckets;  // lengths of the function tables
    // The first hot_count nodes are the top of the world scope's tree,
    // breadth first.  A kernel that copies them into LDS sets num_hot to
    // hot_count; as uploaded, num_hot is 0.
    int32_t hot_count;
    int32_t num_hot;
};

// Nodes at the top of the world scope's tree that the traversal kernels keep
// in LDS (4 KB).
constexpr int kHotNodes = 128;

}  // namespace mlgpu
