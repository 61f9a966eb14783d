// Shared host/device data layout of the pvx engine (gfx950).
//
// HBM layout of a scene:
//   nodes[]   BVH2 node, 64 B: two FP32 child boxes + two child links.
//   prims[]   primitive record, 64 B, in BVH leaf order.  Triangles and
//             spheres are stored inline (no indirection on the hot path);
//             boxes, tori and CSG roots refer to their shape arrays.
//   infinite[] objects tested outside the BVH (planes, unbounded CSG).
// Child link encoding: >= 0 inner node index; < 0 leaf, ~link = first prim
// << 3 | (count - 1), count <= 8.
#ifndef PVX_TYPES_H
#define PVX_TYPES_H

#include <stdint.h>

#include "../../include/pvx.h"

namespace pvx {

// core/configcore.h constants used on the path.
constexpr double kEpsilon = 1.0e-10;
constexpr double kBoundHuge = 2.0e10;
constexpr double kSmallTolerance = 1.0e-3;
constexpr double kMaxDistance = 1.0e7;
constexpr double kMinIsectDepth = 1.0e-4;
constexpr double kShadowTolerance = 1.0e-3; // SHADOW_TOLERANCE (render/trace.cpp:81)
// Internal ray-kind bit: a light-ray launch whose caller only needs to know
// whether something blocks the light (every material opaque).
constexpr uint32_t kRayOcclusion = 0x100u;

enum PrimKind : uint32_t {
    kPrimTriangle = 0,
    kPrimSphere = 1,
    kPrimBox = 2,
    kPrimTorus = 3,
    kPrimCsg = 4,
    kPrimPlane = 5,
    kPrimTriangle64 = 6, // stand-alone triangle, by reference (FP64 record)
    kPrimCone = 7,       // cone / cylinder, by reference
    kPrimDisc = 8,       // disc, by reference
    kPrimQuadric = 9,    // quadric (outside the BVH)
    kPrimPolygon = 10,   // polygon, by reference
    kPrimRevolution = 11,// lathe or sor, by reference
};

struct alignas(16) Node {
    float lo0x, hi0x, lo0y, hi0y; // child 0 x/y slabs
    float lo1x, hi1x, lo1y, hi1y; // child 1 x/y slabs
    float lo0z, hi0z, lo1z, hi1z; // z slabs of both children
    int32_t c0, c1, pad0, pad1;
};
static_assert(sizeof(Node) == 64, "node is one 64 B record");

// Pixel of position t in tile order: tiles of tw x th pixels in row-major
// tile order, row-major inside each tile; the last tile column and row hold
// only the pixels inside the W x H image.  (The reference renders blocks of
// 32 x 32 pixels, backend/scene/view.cpp:258-261, 810; smaller tiles here
// keep the 64 rays of a wavefront close together on the screen.)
#ifdef __HIPCC__
__host__ __device__ inline void tile_pixel(uint32_t W, uint32_t H, uint32_t tw, uint32_t th, uint32_t t,
                                           uint32_t &x, uint32_t &y) {
    const uint32_t row = W * th;   // pixels in a full row of tiles
    const uint32_t tr = t / row;   // tile row (all before it are full)
    const uint32_t k = t - tr * row;
    const uint32_t thr = H - tr * th < th ? H - tr * th : th; // rows in this tile row
    const uint32_t full = W / tw, per = tw * thr;
    const uint32_t tx = k / per;
    if (tx < full) {
        const uint32_t o = k - tx * per;
        x = tx * tw + o % tw;
        y = tr * th + o / tw;
    } else { // the narrower last tile of the row
        const uint32_t o = k - full * per, wr = W - full * tw;
        x = full * tw + o % wr;
        y = tr * th + o / wr;
    }
}
#endif

// 4-wide node, 128 B (triangle / sphere scenes): the boxes of up to four
// children as six float4 rows (x lo, x hi, y lo, y hi, z lo, z hi), their
// links (same encoding as Node; INT32_MIN = no child in this slot).
struct alignas(16) Node4 {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t c[4];
    int32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "wide node is 128 B");

// kd-tree node (the reference's BSP tree, core/bounding/bsptree.h:64-115),
// 8 B.  info bit 31: leaf; leaf: count in bits 0..30, u.first = offset into
// the primitive list; split: axis in bits 29..30, left child in bits 0..28
// (right = left + 1), u.plane = split plane.
constexpr uint32_t kKdMaxDepth = 64;
struct KdNode {
    uint32_t info;
    union {
        float plane;
        uint32_t first;
    } u;
};
static_assert(sizeof(KdNode) == 8, "kd node is 8 B");

struct KdView {
    const KdNode *nodes;
    const uint32_t *list; // leaf primitive lists (indices into prims)
    float bmin[3], bmax[3];
};

// Mesh triangle after Compute_Mesh_Triangle (core/shape/mesh.cpp:838-945):
// FP32 face normal (Normals[Normal_Ind]), FP32 Distance, vertices P1..P3 in
// post-swap order, dominant axis.
struct TriRec {
    float n[3];
    float dist;
    float p1[3];
    float p2[3];
    float p3[3];
    int32_t prim; // triangle index in the mesh2 face list
};
struct SphRec {
    double c[3];
    double r;
    double pad[3];
};
struct RefRec {
    uint32_t index; // index into the kind's shape array
    uint32_t pad[13];
};

struct alignas(16) PrimRec {
    union {
        TriRec tri;
        SphRec sph;
        RefRec ref;
    };
    uint32_t kind_axis; // kind | dominant_axis << 4 | (mesh placement + 1) << 8 (mesh triangles)
    int32_t obj;        // object id
};
static_assert(sizeof(PrimRec) == 64, "prim record is 64 B");

// Transformed shapes (FP64): TRANSFORM matrix + inverse, row-vector convention.
struct Xform {
    double m[16];
    double inv[16];
};

struct BoxRec {
    double c1[3], c2[3];
    int32_t xf;    // -1 = none
    uint32_t flags;
};

// Triangle / SmoothTriangle after Compute_Triangle and
// Compute_Smooth_Triangle (core/shape/triangle.cpp:134-420): FP64 unit
// normal and Distance, P1..P3 in post-swap order, dominant axis; smooth
// ones also N1..N3 (post-swap), Perp and vAxis.  degenerate: never hit.
struct TriangleRec {
    double n[3];
    double dist;
    double p1[3], p2[3], p3[3];
    double n1[3], n2[3], n3[3];
    double perp[3];
    int32_t axis, vaxis, smooth, degenerate;
};

struct TorusRec {
    double R, r;
    int32_t xf;    // always valid (identity if the scene gave none)
    uint32_t sturm;
    uint32_t flags;
    uint32_t pad;
};

// Cone / cylinder after Compute_Cone_Data / Compute_Cylinder_Data
// (core/shape/cone.cpp:715-840): the canonical shape is the frustum of
// x^2 + y^2 = z^2 for z in [dist, 1] (a cone) or the unit cylinder
// x^2 + y^2 = 1 for z in [0, 1], placed by xforms[xf].
constexpr uint32_t kConeCylinder = 1u; // CYLINDER_FLAG
constexpr uint32_t kConeClosed = 2u;   // CLOSED_FLAG (end caps)
struct ConeRec {
    double dist;
    int32_t xf;
    uint32_t flags;
};

// Disc after Parse_Disc / Compute_Disc (core/shape/disc.cpp): the plane
// z = 0 of xforms[xf], squared radii, the stored (world) normal.
struct DiscRec {
    double n[3];
    double ir2, or2;
    int32_t xf;
    uint32_t pad;
};

// Quadric (core/shape/quadric.cpp) in world space after its transforms:
// QA..QJ of quadric.cpp:80-92 (QA/QE/QH square, QB/QC/QF mixed xy/xz/yz,
// QD/QG/QI linear, QJ constant).
struct QuadricRec {
    double qa, qe, qh, qb, qc, qf, qd, qg, qi, qj;
};

// Polygon after Compute_Polygon (core/shape/polygon.cpp:646-805): the
// points in its own u,v plane (poly_pts, 2 doubles each, from `first`),
// xforms[xf].inv maps the world to that plane, S_Normal in world space.
struct PolygonRec {
    double n[3];
    int32_t xf;
    uint32_t first, count;
    uint32_t degenerate;
    uint32_t pad;
};

// Lathe / sor after Compute_Lathe / Compute_Sor and Create_BCyl
// (core/shape/lathe.cpp:488-760, sor.cpp:920-1180,
// core/bounding/boundingcylinder.cpp:408-560): the overall cylinder, the
// squared base and cap radii (sor), the segment range, the placement
// (always valid: identity if the scene gave none) and flags = kind
// (PVX_LATHE_* / PVX_SOR_CUBIC) | kRevClosed | kRevSturm.
constexpr int kRevMaxSegments = 64; // segments per lathe / sor on the device
constexpr uint32_t kRevClosed = 8u;
constexpr uint32_t kRevSturm = 16u;
struct RevRec {
    double radius1, radius2, height1, height2;
    double base_r2, cap_r2;
    uint32_t seg0, nseg;
    int32_t xf;
    uint32_t flags;
};
// One spline segment: x (radius) and y (height) cubic coefficients (a sor
// keeps its r^2(h) cubic in [0]) and the indices of its bounding cylinder's
// radii (squared) and heights in the scene-wide rev_radius / rev_height.
struct RevSeg {
    double a[2], b[2], c[2], d[2];
    uint32_t r1, r2, h1, h2;
};

struct PlaneRec {
    double n[3];
    double d;
};

// CSG node / child tables (device copies of pvx_csg_node / pvx_object) plus
// parent links so a leaf hit can be filtered bottom-up.
struct CsgNode {
    uint32_t op, first, count;
    int32_t parent_slot; // slot of this node in its parent's children, -1 at root
};
struct CsgChild {
    uint32_t type, index, flags;
    uint32_t node; // CSG node that owns this slot
};

// Smooth-normal data of a mesh triangle (Mesh_Triangle_Struct fields used by
// Mesh::Normal / Smooth_Mesh_Normal, core/shape/mesh.cpp:380-420, 2366-2394),
// indices into the scene-wide vertex / normal arrays.
struct ShadeTri {
    float perp[3];
    int32_t p1, p2, p3;   // post-swap vertex indices
    int32_t n1, n2, n3;   // post-swap smooth normal indices
    int32_t fn;           // face normal index
    uint32_t vaxis_smooth; // vAxis | smooth << 4
};

// Device copies of the shading description.
struct MatRec {
    float pigment[3];
    float att;            // opacity 1 - filter - transmit
    float emission[3];    // emission + ambient * ambient_light (classic lighting)
    float ambient_only[3];// emission alone (radiosity lighting)
    float diffuse, brilliance, phong, phong_size;
    float rmin[3], rmax[3];
    float falloff;
    float rweight;        // max(WeightMax(rmax), WeightMax(rmin))
    uint32_t reflective;  // any reflection component non-zero
    float fil[3];         // TransmittedColour: pigment * filter + transmit
    float trans;          // min(1, |grey(fil)|): remaining translucency
    float fmax;           // WeightMaxAbs(fil)
    float ior;            // interior IOR (every object has an interior; default 1)
    uint32_t opaque;      // filter == transmit == 0 (OPAQUE_FLAG)
};
struct LightRec {
    double pos[3];
    double axis1[3];
    double axis2[3];
    float colour[3];
    uint32_t n1, n2;      // grid (1x1 for point lights)
    uint32_t first;       // first shadow-sample slot of this light within a hit
};

// A shaded point lit by one light: the light rays to the light's grid points
// start here (light rays made inside the traversal kernels).
struct LightGroup {
    double ip[3];
    uint32_t hit;
    uint32_t light; // index into the scene's lights
};

// An object tested outside the BVH.
struct InfRec {
    uint32_t kind; // kPrimPlane / kPrimCsg / kPrimQuadric
    uint32_t index;
    int32_t obj;
    uint32_t pad;
};

// Everything a kernel needs, passed by value.
struct SceneView {
    const Node *nodes;
    const PrimRec *prims;
    const InfRec *inf;
    const uint32_t *obj_flags;
    const PlaneRec *planes;
    const BoxRec *boxes;
    const TorusRec *tori;
    const ConeRec *cones;
    const DiscRec *discs;
    const QuadricRec *quadrics;
    const PolygonRec *polygons;
    const double *poly_pts;
    const Xform *xforms;
    const double *sph; // sphere array (cx, cy, cz, r) per sphere, for CSG leaves
    const CsgNode *csg;
    const CsgChild *csg_children;
    const uint32_t *csg_leaves;      // leaf slots of each CSG root, contiguous
    const uint32_t *csg_leaf_range;  // 2 per CSG node: [begin, end) into csg_leaves
    const ShadeTri *shade_tris;      // per mesh triangle (global index)
    const uint32_t *obj_tri_base;    // per object: first ShadeTri of a mesh object
    const int32_t *obj_mesh_xf;      // per object: a mesh's placement (xforms index) or -1
    const float *mesh_vtx;           // 3 per vertex (all meshes)
    const float *mesh_nrm;           // 3 per normal (all meshes: vertex + face normals)
    const TriangleRec *triangles;    // stand-alone triangles (by triangle index)
    const RevRec *revs;              // lathes and sors
    const RevSeg *rev_segs;
    const double *rev_radius;        // bounding cylinder radii (squared)
    const double *rev_height;        // bounding cylinder heights
    const Node4 *nodes4;             // 4-wide tree over the same leaves (wide scenes)
    int32_t root4;     // root link of the 4-wide tree
    uint32_t wide;     // 1: triangle/sphere scene traversed through nodes4
    int32_t root;      // root link (node index or leaf link); INT32_MIN = empty
    uint32_t n_inf;
    uint32_t any_flags;
    uint32_t has_complex; // 0: mesh triangles/spheres only, 1: + boxes, stand-alone triangles, 2: + tori or CSG,
                          // 3: + lathes or sors
};

} // namespace pvx

#endif
