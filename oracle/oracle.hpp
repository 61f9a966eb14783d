// FP64 CPU ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A restatement, in this project's own code, of the POV-Ray 3.8 algorithms on
// the closest-hit / per-shape / radiosity hot path.  It is the checker for the
// HIP engine: only tests, the smoke check and the bench's CPU-baseline leg may
// load it.  The product library never links or calls it.
//
// Every function cites the reference file:line it follows (paths relative to
// the POV-Ray source tree).  Floating-point types follow the reference: DBL
// (double) for geometry, SNGL (float) where the reference stores FP32
// (BoundingBox, mesh vertices/normals/Perp/Distance).  Build with
// -ffp-contract=off so that no FMA contraction changes rounding.
#ifndef PVX_ORACLE_HPP
#define PVX_ORACLE_HPP

#include <cmath>
#include <cstdint>
#include <vector>

#include "../include/pvx.h"

namespace orc {

// core/configcore.h:153-201 (HUGE_VAL is <cmath>'s +inf: <cmath> is included
// before configcore.h's #ifndef HUGE_VAL fallback).
constexpr double kEpsilon = 1.0e-10;
constexpr double kBoundHuge = 2.0e10;
constexpr double kSmallTolerance = 1.0e-3;
constexpr double kMaxDistance = 1.0e7;
constexpr double kMinIsectDepth = 1.0e-4;

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    double &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(double s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double length_sqr(V3 a) { return dot(a, a); }
inline double length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalized(V3 a) { return a / length(a); }

struct Ray {
    V3 o, d;
    V3 eval(double t) const { return o + d * t; }
};

// core/math/matrix.h TRANSFORM: matrix + inverse, row-vector convention.
struct Transform {
    double m[4][4];
    double inv[4][4];
};
V3 mtrans_point(const double (*m)[4], V3 p);     // matrix.cpp:415-427
V3 mtrans_direction(const double (*m)[4], V3 p); // matrix.cpp:455-466
V3 minv_trans_normal(const double (*m)[4], V3 p);// matrix.cpp MInvTransNormal

// core/bounding/boundingbox.h:83-108 BoundingBox (lowerLeft/size, FP32).
struct FBox {
    float ll[3] = {0, 0, 0};
    float size[3] = {0, 0, 0};
};
FBox make_bbox_from_min_max(V3 mins, V3 maxs); // boundingbox.h:114-118 (Vector3d)

// One entry of an IStack (core/coretypes.h Intersection, reduced to what the
// hot path reports).
struct Isect {
    double depth;
    int32_t prim;   // triangle index / CSG leaf slot
    int32_t side;   // box side (Intersection::i1); lathe segment; sor part | segment << 2
    int32_t leaf;   // CSG: leaf shape (type<<24 | index), else -1
    double d1 = 0;  // Intersection::d1 (lathe: spline parameter of the hit)
};

// ---------------------------------------------------------------------------
// BBox tree (core/bounding/boundingbox.cpp)
// ---------------------------------------------------------------------------
struct BNode {
    FBox bbox;
    bool infinite = false;
    std::vector<int32_t> kids; // empty => leaf
    int32_t element = -1;      // leaf payload (object id / triangle index)
};

class BBoxTree {
  public:
    std::vector<BNode> nodes;
    int32_t root = -1;
    // Build_BBox_Tree (boundingbox.cpp:262-323) over leaf boxes; infinite
    // elements go into the root's first child.
    void build(const std::vector<FBox> &finite_boxes, const std::vector<int32_t> &finite_ids,
               const std::vector<FBox> &infinite_boxes, const std::vector<int32_t> &infinite_ids);
    int depth() const;
};

// Rayinfo (boundingbox.h:182-216)
struct RayInfo {
    float origin[3];
    float inv_dir[3];
    bool nonzero[3];
    bool positive[3];
    explicit RayInfo(const Ray &r);
};

// BBoxPriorityQueue (boundingbox.cpp:80-147): binary min-heap, element 0 unused.
class PQueue {
  public:
    struct Elem {
        double depth;
        int32_t node;
    };
    std::vector<Elem> q{Elem{0, -1}};
    void clear() { q.resize(1); }
    bool empty() const { return q.size() == 1; }
    void insert(double depth, int32_t node);
    bool remove_min(double &depth, int32_t &node);
};

// Check_And_Enqueue (boundingbox.cpp:541-648)
void check_and_enqueue(PQueue &q, const BNode &node, int32_t node_id, const RayInfo &ri);

// Intersect_BBox_Dir (object.cpp:1074-1108) on an object's FP32 box.
bool intersect_bbox_dir(const FBox &b, const Ray &r, float mind, float maxd);

// ---------------------------------------------------------------------------
// shapes
// ---------------------------------------------------------------------------
// Sphere::Intersect (sphere.cpp:214-241)
bool sphere_intersect(const Ray &r, V3 centre, double radius2, double *d1, double *d2);

struct MeshTri {                       // Mesh_Triangle_Struct (mesh.h:88-107)
    float perp[3];
    float distance;
    int32_t normal_ind;
    int32_t p1, p2, p3;
    int32_t n1, n2, n3;
    uint8_t smooth, dominant_axis, v_axis;
};

struct Mesh {
    std::vector<float> vertices;  // 3 * nv (MeshVector, FP32)
    std::vector<float> normals;   // 3 * (n_normals + n_triangles)
    std::vector<MeshTri> tris;
    std::vector<uint8_t> valid;   // Compute_Mesh_Triangle result (degenerate => 0)
    BBoxTree tree;
    FBox bbox;
    V3 vert(int i) const {
        return {vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]};
    }
};
void mesh_prepare(Mesh &m, const pvx_mesh_desc &d);                // Parse_Mesh2 tail + Build_Mesh_BBox_Tree
bool mesh_triangle_intersect(const Mesh &m, const MeshTri &t, const Ray &r, double *depth); // mesh.cpp:1040-1136
// Mesh::Intersect (mesh.cpp:179-230): tr places the mesh (NULL: none); the
// depths pushed are world depths (test_hit, mesh.cpp:1218-1243)
void mesh_all_intersections(const Mesh &m, const Transform *tr, const Ray &r, std::vector<Isect> &stack,
                            bool use_tree);
bool mesh_inside(const Mesh &m, V3 p);
V3 mesh_normal(const Mesh &m, int tri, V3 ipoint);                   // Smooth_Mesh_Normal, mesh space
V3 mesh_normal_world(const Mesh &m, const Transform *tr, int tri, V3 ipoint); // Mesh::Normal (mesh.cpp:342-380)
FBox recompute_bbox(const FBox &b, const Transform *tr);             // Recompute_BBox

// Box::Intersect (box.cpp:172-505)
bool box_intersect(const Ray &r, const Transform *tr, V3 c1, V3 c2, double *d1, double *d2,
                   int *s1, int *s2);
bool box_inside(V3 p, const Transform *tr, V3 c1, V3 c2, bool inverted);
// Torus (torus.cpp)
bool torus_all_intersections(const Ray &r, const Transform *tr, double R, double rr, bool sturm,
                             std::vector<Isect> &stack, int32_t slot);
FBox box_bbox(V3 c1, V3 c2, const Transform *tr);
FBox torus_bbox(double R, double rr, const Transform *tr);
bool torus_inside(V3 p, const Transform *tr, double R, double rr, bool inverted);

// TRANSFORM builders (core/math/matrix.cpp): Compose_Transforms
// (matrix.cpp:788-793) and Compute_Coordinate_Transform (matrix.cpp:880-906).
void compose(Transform &t, const Transform &add);
void coordinate_transform(Transform &t, V3 origin, V3 &up, double radius, double len);

// Disc (disc.cpp): placement, normal and radii after Parse_Disc /
// Compute_Disc and the object's transform.
struct Disc {
    Transform tr;
    V3 normal;
    double iradius2 = 0.0, oradius2 = 1.0;
};
Disc disc_prepare(const double *center, const double *normal, const double *radius, const Transform *user);
bool disc_intersect(const Disc &c, const Ray &r, double *depth);
bool disc_inside(const Disc &c, V3 p, bool inverted);
FBox disc_bbox(const Disc &c);

// Polygon (polygon.cpp): points in the polygon's own u,v plane, the
// world-to-plane transform (inverse only), S_Normal, the bounding box.
struct PolygonO {
    Transform tr;
    std::vector<double> pts;
    int number = 0;
    bool degenerate = false;
    V3 normal;
    FBox bbox;
};
PolygonO polygon_prepare(const double *pts, int number, const Transform *user);
bool polygon_intersect(const PolygonO &g, const Ray &r, double *depth);

// Quadric (quadric.cpp): coefficients A..J after the object's transform.
struct Quadric {
    double c[10];
    FBox bbox; // Compute_BBox at parse time, moved by the transform
};
Quadric quadric_prepare(const double *coef, const Transform *user);
FBox quadric_bbox(const double *coef, const Transform *user);
int quadric_intersect(const Quadric &q, const Ray &r, double *depth); // up to 2 hits
bool quadric_inside(const Quadric &q, V3 p, bool inverted);
V3 quadric_normal(const Quadric &q, V3 p);

// Cone / cylinder (cone.cpp): the shape's placement and flags after
// Compute_Cone_Data / Compute_Cylinder_Data and the object's transform.
struct Cone {
    Transform tr;
    double dist = 0.0;
    bool cylinder = false, closed = true;
};
Cone cone_prepare(const double *apex, const double *base, const double *radius, uint8_t flags,
                  const Transform *user);
int cone_intersect(const Cone &c, const Ray &r, double *depth, int *part); // up to 4 hits
bool cone_inside(const Cone &c, V3 p, bool inverted);
V3 cone_normal(const Cone &c, int part, V3 ip);
FBox cone_bbox(const Cone &c);

// polynomialsolver.cpp:1585-1732
int solve_polynomial(int n, const double *c0, double *r, bool sturm, double epsilon);

// Lathe (lathe.cpp) and surface of revolution (sor.cpp) with their bounding
// cylinders (core/bounding/boundingcylinder.cpp): spline segments, the
// bounding cylinder lists, the overall cylinder and the placement.
struct Revolution {
    bool sor = false;
    int spline = 0;              // PVX_LATHE_* (PVX_SOR_CUBIC for a sor)
    bool closed = false, sturm = false;
    std::vector<double> seg;     // 8 per segment: A.x A.y B.x B.y C.x C.y D.x D.y (sor: A B C D in x)
    std::vector<double> radius;  // BCYL radius (squared), height lists
    std::vector<double> height;
    std::vector<int32_t> entry;  // 4 per segment: r1 r2 h1 h2
    double radius1 = 0, radius2 = 0, height1 = 0, height2 = 0;
    double base_r2 = 0, cap_r2 = 0;
    Transform tr;
    FBox bbox;
    int segments() const { return (int)(seg.size() / 8); }
};
Revolution lathe_prepare(const double *pts, int n, int spline, bool sturm, const Transform *user);
Revolution sor_prepare(const double *pts, int n, bool open, bool sturm, const Transform *user);
// All_Intersections; `child` is IS_CHILD_OBJECT (no early exit once the
// next bounding interval starts beyond the best hit)
bool revolution_intersect(const Revolution &s, const Ray &r, bool child, std::vector<Isect> &st, int32_t slot);
bool revolution_inside(const Revolution &s, V3 p, bool inverted);
V3 revolution_normal(const Revolution &s, const Isect &h, V3 ip);

// ---------------------------------------------------------------------------
// scene (SceneData::objects + shape storage + bounding slabs)
// ---------------------------------------------------------------------------
struct Object {
    uint32_t type, index, flags;
    FBox bbox;
    bool infinite = false;
};

// A stand-alone triangle after Compute_Triangle (triangle.cpp:134-420).
struct Tri {
    V3 p1, p2, p3, n1, n2, n3, normal, perp;
    double distance = 0;
    int axis = 0, vaxis = 0;
    bool smooth = false, degenerate = false;
};
Tri make_tri(const double *P, const double *N, bool smooth);
bool tri_intersect(const Tri &t, const Ray &r, double *depth);
V3 tri_normal(const Tri &t, V3 ip);
FBox tri_bbox(const double *P);

struct Scene {
    std::vector<Object> objects;
    std::vector<Tri> tris;
    std::vector<double> tri_p; // 9 per triangle, as given
    std::vector<V3> sph_c;
    std::vector<double> sph_r;
    std::vector<V3> pl_n;
    std::vector<double> pl_d;
    std::vector<V3> box_c1, box_c2;
    std::vector<int32_t> box_tr;
    std::vector<double> tor_R, tor_r;
    std::vector<int32_t> tor_tr;
    std::vector<uint8_t> tor_sturm;
    std::vector<Cone> cones;
    std::vector<Disc> discs;
    std::vector<Quadric> quadrics;
    std::vector<PolygonO> polygons;
    std::vector<Revolution> revs;    // lathes and sors
    std::vector<Transform> trs;
    std::vector<Mesh> meshes;
    std::vector<int32_t> mesh_tr;     // each mesh's placement (trs index) or -1
    std::vector<pvx_csg_node> csg;
    std::vector<pvx_object> csg_children;
    BBoxTree tree;

    const Transform *tr(int32_t i) const { return i < 0 ? nullptr : &trs[i]; }
    FBox shape_bbox(uint32_t type, uint32_t index, uint32_t flags, bool *infinite) const;
    // `child`: the shape is a CSG operand (IS_CHILD_OBJECT)
    bool all_intersections(uint32_t type, uint32_t index, uint32_t flags, const Ray &r,
                           std::vector<Isect> &st, int32_t slot, bool child = false) const;
    bool inside(uint32_t type, uint32_t index, uint32_t flags, V3 p) const;
    bool find_object(uint32_t obj, const Ray &r, bool shadow, Isect &out) const;
};

} // namespace orc

#endif
