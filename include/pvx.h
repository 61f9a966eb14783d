/*
 * pvx.h — C-ABI of the MI355X ray-intersection engine for POV-Ray's hot path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes (0 = ok),
 * no exceptions across the boundary, explicit stream argument (a hipStream_t
 * passed as void*; NULL = the default stream).  Ray/hit/point buffers passed
 * to the trace entry points live in DEVICE memory and are owned by the caller.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the POV-Ray 3.8 source tree).
 *
 * Streams: the per-ray entry points (pvx_trace_closest[_counted],
 * pvx_all_intersections, pvx_inside) may run on one scene from several
 * streams at once — each stream gets its own traversal-stack area, allocated
 * on its first launch.  The frame-level calls (pvx_render, pvx_render_samples,
 * pvx_radiosity_render) keep per-scene work buffers, and a batch keeps its
 * chunk table and stack area: calls on one scene's frames, or on one batch,
 * must not overlap in time.
 */
#ifndef PVX_H
#define PVX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    PVX_OK = 0,
    PVX_EINVAL = 1,       /* bad argument (null pointer, bad index, bad size)   */
    PVX_ENOMEM = 2,       /* host or device allocation failed                   */
    PVX_EDEVICE = 3,      /* a HIP call failed                                  */
    PVX_EUNSUPPORTED = 4, /* feature not available in this build                */
    PVX_ESTATE = 5        /* handle in the wrong state (e.g. no scene uploaded) */
};

/* Opaque scene handle: device-resident geometry + acceleration structure. */
typedef struct pvx_scene pvx_scene;

/* ---- object model ------------------------------------------------------ */
/* Shape kinds: the subset of core/shape on the hot path (sphere.cpp,
 * box.cpp, torus.cpp, plane.cpp, mesh.cpp, csg.cpp, triangle.cpp, cone.cpp,
 * disc.cpp, quadric.cpp, polygon.cpp, lathe.cpp, sor.cpp). */
enum {
    PVX_SPHERE = 1,
    PVX_BOX = 2,
    PVX_TORUS = 3,
    PVX_PLANE = 4,
    PVX_MESH = 5,
    PVX_CSG = 6,
    PVX_TRIANGLE = 7,        /* stand-alone triangle (triangle.cpp, Triangle)              */
    PVX_SMOOTH_TRIANGLE = 8, /* stand-alone smooth triangle (triangle.cpp, SmoothTriangle) */
    PVX_CONE = 9,            /* cone or cylinder (cone.cpp, Cone)                          */
    PVX_DISC = 10,           /* disc, optionally with a hole (disc.cpp, Disc)              */
    PVX_QUADRIC = 11,        /* quadric surface (quadric.cpp, Quadric)                     */
    PVX_POLYGON = 12,        /* planar polygon (polygon.cpp, Polygon; not a CSG operand)   */
    PVX_LATHE = 13,          /* lathe (lathe.cpp, Lathe)                                   */
    PVX_SOR = 14             /* surface of revolution (sor.cpp, Sor)                       */
};

/* pvx_scene_desc.revolution_kind: the spline a lathe is built from
 * (Parse_Lathe's linear_spline / quadratic_spline / cubic_spline /
 * bezier_spline), or PVX_SOR_CUBIC for a sor (Compute_Sor's cubic in the
 * height through each pair of points) */
enum {
    PVX_LATHE_LINEAR = 0,
    PVX_LATHE_QUADRATIC = 1,
    PVX_LATHE_CUBIC = 2,
    PVX_LATHE_BEZIER = 3,
    PVX_SOR_CUBIC = 4
};

/* pvx_scene_desc.revolution_flags bits */
enum {
    PVX_REV_STURM = 1u << 0, /* `sturm`: Sturm sequences for the polynomial roots (STURM_FLAG) */
    PVX_REV_OPEN = 1u << 1   /* sor `open`: no base and cap discs (CLOSED_FLAG cleared)        */
};

/* pvx_scene_desc.cone_flags bits */
enum {
    PVX_CONE_OPEN = 1u << 0,    /* `open`: no end caps (CLOSED_FLAG cleared)               */
    PVX_CONE_CYLINDER = 1u << 1 /* a `cylinder` statement (Parse_Cylinder): one radius     */
};

/* Object flags: the bits of ObjectBase::Flags that the hot path reads
 * (core/scene/object.h: NO_SHADOW_FLAG, NO_IMAGE_FLAG, NO_REFLECTION_FLAG,
 * NO_RADIOSITY_FLAG, INVERTED_FLAG). */
enum {
    PVX_FLAG_NO_SHADOW = 1u << 0,
    PVX_FLAG_NO_IMAGE = 1u << 1,
    PVX_FLAG_NO_REFLECTION = 1u << 2,
    PVX_FLAG_NO_RADIOSITY = 1u << 3,
    PVX_FLAG_INVERTED = 1u << 4
};

/* Ray kinds: the Ray::IsImageRay/IsReflectionRay/IsRadiosityRay tests read
 * by NoSomethingFlagRayObjectCondition (core/render/trace.cpp:84-95), plus
 * SHADOW for the NoShadowFlagRayObjectCondition of shadow rays. */
enum {
    PVX_RAY_PRIMARY = 0,
    PVX_RAY_IMAGE = 1,
    PVX_RAY_REFLECTION = 2,
    PVX_RAY_RADIOSITY = 4,
    PVX_RAY_SHADOW = 8
};

/* CSG operators (core/shape/csg.h).  A difference is expressed as an
 * intersection whose children 2..n carry PVX_FLAG_INVERTED, as the parser
 * builds it (parser/parser.cpp:2223-2225). */
enum {
    PVX_CSG_UNION = 1,
    PVX_CSG_INTERSECTION = 2,
    PVX_CSG_MERGE = 3
};

/* One entry of the object table: shape kind + index into that kind's array.
 * Top-level objects are listed in scene order (SceneData::objects); the
 * position in the table is the object id reported in hits. */
typedef struct pvx_object {
    uint32_t type;  /* PVX_SPHERE ... PVX_CSG                     */
    uint32_t index; /* index into the per-kind arrays below       */
    uint32_t flags; /* PVX_FLAG_* bits                            */
    uint32_t pad;
} pvx_object;

/* CSG node: operator + a run of children in pvx_scene_desc.csg_children.
 * A child of type PVX_CSG refers to another node (nested CSG). */
typedef struct pvx_csg_node {
    uint32_t op;          /* PVX_CSG_*                          */
    uint32_t first_child; /* index into csg_children           */
    uint32_t n_children;
    uint32_t pad;
} pvx_csg_node;

/* mesh2 data as Parse_Mesh2 receives it (parser/parser.cpp:4048-4600):
 * FP32 vertex_vectors, face_indices, optional per-vertex normal_vectors
 * (normal indices = face indices, the "number_of_normals == number_of_vertices"
 * rule at parser.cpp:4494-4509).  Vertices are in the mesh's own space
 * (world space unless pvx_scene_desc.mesh_transform places the mesh). */
typedef struct pvx_mesh_desc {
    uint32_t n_vertices;
    uint32_t n_triangles;
    uint32_t n_normals;      /* 0 or n_vertices                          */
    uint32_t pad;
    const float *vertices;   /* 3 * n_vertices                           */
    const int32_t *indices;  /* 3 * n_triangles                          */
    const float *normals;    /* 3 * n_normals, or NULL                   */
} pvx_mesh_desc;

/* Scene description.  All arrays are HOST memory and are copied.
 * Transforms follow POV-Ray's TRANSFORM (core/math/matrix.h): two 4x4 FP64
 * matrices, `matrix` then `inverse`, row-vector convention
 * (p' = p.x*m[0][i] + p.y*m[1][i] + p.z*m[2][i] + m[3][i]); 32 doubles each. */
typedef struct pvx_scene_desc {
    uint32_t n_objects;
    const pvx_object *objects;

    uint32_t n_spheres;
    const double *sphere_center; /* 3n */
    const double *sphere_radius; /* n  */

    uint32_t n_planes;
    const double *plane_normal;   /* 3n, unit length                        */
    const double *plane_distance; /* n: N.P + D = 0 (Plane::Distance)       */

    uint32_t n_boxes;
    const double *box_corner1;     /* 3n (Corner1 <= Corner2 per axis)       */
    const double *box_corner2;     /* 3n                                     */
    const int32_t *box_transform;  /* n, index into transforms or -1         */

    uint32_t n_tori;
    const double *torus_major;     /* n (MajorRadius)                        */
    const double *torus_minor;     /* n (MinorRadius)                        */
    const int32_t *torus_transform;/* n, index into transforms or -1         */
    const uint8_t *torus_sturm;    /* n, 1 = `sturm` keyword (STURM_FLAG)    */

    uint32_t n_transforms;
    const double *transforms;      /* 32 * n_transforms                      */

    uint32_t n_meshes;
    const pvx_mesh_desc *meshes;

    uint32_t n_csg;
    const pvx_csg_node *csg;
    uint32_t n_csg_children;
    const pvx_object *csg_children;

    /* stand-alone triangles (FP64, world space; not CSG children): an object
     * of type PVX_TRIANGLE or PVX_SMOOTH_TRIANGLE refers to triangle `index`;
     * a triangle is referred to as one of the two kinds only */
    uint32_t n_triangles;
    const double *triangle_vertices; /* 9n: P1, P2, P3                          */
    const double *triangle_normals;  /* 9n: N1, N2, N3 (smooth ones; may be NULL
                                        when no smooth triangle is listed)      */

    /* cones and cylinders as Parse_Cone / Parse_Cylinder read them
     * (parser/parser.cpp:2255-2315): the first point and radius are the
     * shape's `apex`, the second its `base`; a cylinder's one radius is
     * cone_radius[2i].  The shape's own placement (Compute_Cone_Data /
     * Compute_Cylinder_Data, cone.cpp:715-840) is computed from these, then
     * the optional transform is composed after it (Cone::Transform). */
    uint32_t n_cones;
    const double *cone_apex;        /* 3n: first point                          */
    const double *cone_base;        /* 3n: second point                         */
    const double *cone_radius;      /* 2n: first radius, second radius          */
    const uint8_t *cone_flags;      /* n: PVX_CONE_* bits, or NULL (closed cones) */
    const int32_t *cone_transform;  /* n: index into transforms or -1; or NULL  */

    /* discs as Parse_Disc reads them: centre, normal (normalised here),
     * radius and hole radius; the placement (Compute_Disc) is computed from
     * them, the optional transform composed after it (Disc::Transform). */
    uint32_t n_discs;
    const double *disc_center;      /* 3n                                       */
    const double *disc_normal;      /* 3n                                       */
    const double *disc_radius;      /* 2n: radius, hole radius (0: none)        */
    const int32_t *disc_transform;  /* n: index into transforms or -1; or NULL  */

    /* quadrics as Parse_Quadric reads them: <A,B,C> (Square_Terms),
     * <D,E,F> (Mixed_Terms), <G,H,I> (Terms), J (Constant), i.e. the
     * surface A x^2 + B y^2 + C z^2 + D xy + E xz + F yz + G x + H y + I z + J
     * = 0; the optional transform is applied to the coefficients
     * (Quadric::Transform).  Quadrics are tested outside the BVH, like planes. */
    uint32_t n_quadrics;
    const double *quadric_coef;     /* 10n                                      */
    const int32_t *quadric_transform; /* n: index into transforms or -1; or NULL */

    /* polygons as Parse_Polygon reads them: polygon i has the points
     * polygon_points[3 * polygon_first[i] .. 3 * polygon_first[i + 1]) in
     * statement order (sub-polygons closed by repeating their first point;
     * an open last one is closed as the parser does); the optional transform
     * is composed after the polygon's own placement (Polygon::Transform). */
    uint32_t n_polygons;
    const uint32_t *polygon_first;  /* n + 1 offsets, in points                 */
    const double *polygon_points;   /* 3 per point                              */
    const int32_t *polygon_transform; /* n: index into transforms or -1; or NULL */

    /* the placement of each mesh (Mesh::Transform, core/shape/mesh.cpp:568-580):
     * its vertices stay in the mesh's own space; a ray is moved into that space
     * and its direction normalised (Mesh::Intersect, mesh.cpp:179-230), hit
     * depths are divided by the direction's length (test_hit, mesh.cpp:1218-1243),
     * normals leave through MTransNormal (Mesh::Normal, mesh.cpp:342-380) and
     * the object box is the mesh box through Recompute_BBox */
    const int32_t *mesh_transform;  /* n_meshes: index into transforms or -1; or NULL */

    /* lathes and surfaces of revolution as Parse_Lathe (parser/parser.cpp:
     * 2812-2995) and Parse_Sor (parser.cpp:5628-5714) read them: shape i has
     * the 2-D points revolution_points[2 * revolution_first[i] ..
     * 2 * revolution_first[i + 1]) (x the radius, y the height) and the spline
     * revolution_kind[i]; an object of type PVX_LATHE refers to a shape with a
     * PVX_LATHE_* kind, one of type PVX_SOR to a PVX_SOR_CUBIC one.  The
     * segments and bounding cylinders (Compute_Lathe, Compute_Sor, Create_BCyl)
     * are computed from the points; the optional transform places the shape
     * (Lathe::Transform / Sor::Transform).  At most 64 segments per shape. */
    uint32_t n_revolutions;
    const uint32_t *revolution_first;   /* n + 1 offsets, in points                */
    const double *revolution_points;    /* 2 per point                             */
    const uint8_t *revolution_kind;     /* n: PVX_LATHE_* or PVX_SOR_CUBIC         */
    const uint8_t *revolution_flags;    /* n: PVX_REV_* bits, or NULL              */
    const int32_t *revolution_transform;/* n: index into transforms or -1; or NULL */
} pvx_scene_desc;

/* ---- rays and hits ----------------------------------------------------- */
/* FP64 ray as core/render/ray.h BasicRay (Origin, Direction); 48 bytes. */
typedef struct pvx_ray {
    double o[3];
    double d[3];
} pvx_ray;

/* Closest hit: Intersection::Depth, object id, primitive id; 16 bytes.
 * obj = -1 on a miss (t is then the ray's max depth, BOUND_HUGE = 2e10 if
 * none: the default Intersection::Depth, core/coretypes.h:384-387).
 * prim = triangle index for meshes (face order of the mesh2), the leaf's
 * position in csg_children for CSG hits, 0 otherwise. */
typedef struct pvx_hit {
    double t;
    int32_t obj;
    int32_t prim;
} pvx_hit;

/* Acceleration structure the closest-hit calls use. */
enum {
    PVX_ACCEL_BVH = 0, /* flattened BVH (4-wide for triangle/sphere scenes); default  */
    PVX_ACCEL_BSP = 1  /* kd-tree with the cost model of BSPTree (+BM2,
                          core/bounding/bsptree.cpp:276-705); same hits, slower      */
};

/* Who builds the BVH (PVX_ACCEL_BVH).  Both trees give the same hits. */
enum {
    PVX_BUILDER_HOST_SAH = 0, /* binned SAH on the host; default                    */
    PVX_BUILDER_GPU = 1       /* linear BVH (Morton order) built on the GPU: a much
                                 faster build, a somewhat slower tree to traverse  */
};

typedef struct pvx_build_opts {
    uint32_t max_leaf_size; /* triangles per BVH leaf (0 = default 4) */
    uint32_t sah_bins;      /* binned-SAH bins (0 = default 16)        */
    uint32_t accel;         /* PVX_ACCEL_BVH / PVX_ACCEL_BSP          */
    uint32_t bsp_max_depth; /* BSP tree depth limit (0 = 64, at most 64) */
    uint32_t builder;       /* PVX_BUILDER_HOST_SAH / PVX_BUILDER_GPU (BVH only) */
    uint32_t reserved[3];
} pvx_build_opts;

/* Counts describe the tree the closest-hit kernels walk: the kd-tree (8 B
 * nodes; n_prims = leaf list entries) for PVX_ACCEL_BSP, else the 4-wide tree
 * (128 B nodes) for triangle / sphere scenes and the binary tree (64 B). */
typedef struct pvx_scene_stats {
    uint64_t n_nodes;        /* BVH nodes                               */
    uint64_t n_prims;        /* BVH primitive references                */
    uint64_t node_bytes;     /* bytes of one BVH node record            */
    uint64_t tri_bytes;      /* bytes of one triangle record            */
    uint64_t device_bytes;   /* total device bytes held by the scene    */
    uint32_t bvh_depth;      /* max BVH depth                           */
    uint32_t n_infinite;     /* objects tested outside the BVH (planes) */
    double build_ms;         /* scene preparation + tree build, wall    */
    double bvh_build_ms;     /* the tree build alone (host or GPU), wall */
} pvx_scene_stats;

/* ---- library ----------------------------------------------------------- */
/* Last error message of the calling thread (empty if none). */
const char *pvx_last_error(void);

/* Device count / device properties check; returns PVX_EDEVICE without a GPU. */
int pvx_init(int device);

/* ---- scene ------------------------------------------------------------- */
/* Upload a scene and build its acceleration structure.
 * Replaces parse-time Compute_Mesh_Triangle (core/shape/mesh.cpp:838-945)
 * and BoundingTask::Run -> Build_Bounding_Slabs / Build_Mesh_BBox_Tree
 * (backend/bounding/boundingtask.cpp:168-210,
 *  core/bounding/boundingbox.cpp:325-430, core/shape/mesh.cpp:1376-1409). */
int pvx_scene_create(const pvx_scene_desc *desc, const pvx_build_opts *opts,
                     pvx_scene **out);
int pvx_scene_destroy(pvx_scene *scene);
int pvx_scene_get_stats(const pvx_scene *scene, pvx_scene_stats *out);

/* ---- hot path ---------------------------------------------------------- */
/* Closest hit for n rays (device buffers).
 * Replaces bool Trace::FindIntersection(Intersection&, const Ray&,
 *   const RayObjectCondition& pre, const RayObjectCondition& post)
 *   (core/render/trace.h:255-258, trace.cpp:285-344) with
 *   pre = NoSomethingFlagRayObjectCondition for `ray_kind`
 *   (NoShadowFlag for PVX_RAY_SHADOW), post = TrueRayObjectCondition, and
 *   bestisect.Depth seeded from max_depth (trace.cpp:157-158); max_depth may
 *   be NULL (HUGE_VAL for every ray). */
int pvx_trace_closest(pvx_scene *scene, const pvx_ray *rays,
                      const double *max_depth, uint32_t ray_kind, size_t n,
                      pvx_hit *hits, void *stream);

/* pvx_trace_closest plus per-ray work counters: node_visits[i] = BVH nodes
 * fetched, prim_tests[i] = primitive records tested.  These are the device
 * counterparts of the reference's RenderStatistics counters nChecked /
 * Ray_<Shape>_Tests (core/support/statisticids.h:120-268); the bench turns
 * them into algorithmic bytes per ray. */
int pvx_trace_closest_counted(pvx_scene *scene, const pvx_ray *rays,
                              const double *max_depth, uint32_t ray_kind, size_t n,
                              pvx_hit *hits, uint32_t *node_visits,
                              uint32_t *prim_tests, void *stream);

/* pvx_trace_closest without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and checks as pvx_trace_closest; the
 * traversal-stack spill area is a fixed 32 MiB device array of the library
 * (one per device, held until the process ends) instead of the per-stream
 * area pvx_trace_closest allocates on a stream's first launch, so the
 * persistent grid is capped at what it holds (819 blocks of 256 lanes).
 * Every call on one device shares that array: calls that overlap in time on
 * one device (other streams, other threads) race and must not be issued.
 * The scene must live on the current device (PVX_EINVAL otherwise). */
int pvx_trace_closest_noalloc(pvx_scene *scene, const pvx_ray *rays,
                              const double *max_depth, uint32_t ray_kind, size_t n,
                              pvx_hit *hits, void *stream);

/* Traversal variants (every one gives the same hits). */
enum {
    PVX_TRAVERSAL_DEFAULT = 0,   /* by scene content and ray kind                     */
    PVX_TRAVERSAL_WIDE_WW = 1,   /* 4-wide tree, while-while (triangle/sphere scenes) */
    PVX_TRAVERSAL_WIDE_STEP = 2, /* 4-wide tree, one node or leaf per pass            */
    PVX_TRAVERSAL_BINARY = 3,    /* binary tree (scenes without tori / CSG)           */
    /* pvx_render's light rays in triangle/sphere scenes whose materials are all
     * opaque (chosen by pvx_render_autotune, reported in its result): the rays
     * are made in the kernel from the shaded point and the light's grid and
     * only tested for occlusion, with the 4-wide tree walked ... */
    PVX_TRAVERSAL_PACKET = 4,    /* ... by each wavefront as one packet (one shared
                                    stack, ballots of the lanes' box tests)          */
    PVX_TRAVERSAL_FUSED = 5,     /* ... while-while per lane (the default)          */
    PVX_TRAVERSAL_REFILL = 6,    /* ... while-while per lane, a wavefront refilling
                                    its free lanes with the next unstarted rays once
                                    16 or fewer of its lanes still hold rays         */
    PVX_TRAVERSAL_LIGHT_BINARY = 7, /* ... through the binary tree, per lane      */
    PVX_TRAVERSAL_BINARY_WW = 8, /* binary tree, while-while (triangle/sphere scenes) */
    PVX_TRAVERSAL_REFILL_BINARY = 9 /* pvx_render's light rays as REFILL, through the
                                       binary tree                                  */
};

typedef struct pvx_tune_result {
    uint32_t variant;      /* chosen PVX_TRAVERSAL_*                        */
    uint32_t grid_blocks;  /* chosen persistent grid (blocks of 256 lanes)  */
    double best_ms;        /* its time per launch on the sample              */
    double default_ms;     /* the untuned choice's time on the same sample   */
    uint32_t tile_w;       /* pvx_render_autotune's camera rays: the chosen  */
    uint32_t tile_h;       /* pixel tile of the ray queue (0 x 0 otherwise;
                              tile_w 0 with tile_h 1: row-major)             */
} pvx_tune_result;

/* Autotuner: times the traversal variants this scene admits on persistent
 * grids of one and two times the resident blocks, over the caller's sample
 * rays (device), and keeps the fastest for later closest-hit calls on rays
 * of the same class (PVX_RAY_SHADOW, or every other kind).  No reference
 * counterpart (the reference has one traversal).  Results of later calls are
 * unchanged; only their speed is. */
int pvx_scene_autotune(pvx_scene *scene, const pvx_ray *rays, size_t n, uint32_t ray_kind,
                       pvx_tune_result *out, void *stream);

/* Sets the variant and grid for rays of ray_kind's class directly (e.g. a
 * choice saved from an earlier pvx_scene_autotune); variant
 * PVX_TRAVERSAL_DEFAULT with grid_blocks 0 restores the untuned behaviour.
 * PVX_EINVAL for a variant the scene does not admit or grid_blocks above
 * twice the resident blocks.  Not to be called while launches on the scene
 * are being issued from another thread. */
int pvx_scene_set_traversal(pvx_scene *scene, uint32_t ray_kind, uint32_t variant, uint32_t grid_blocks);

/* Pixel order of pvx_render's ray queue: tiles of tile_w x tile_h pixels
 * (tile order as pvx_camera_rays_tiled), or row-major with tile_w 0 and
 * tile_h 1.  8 x 8 by default; pvx_render_autotune picks among 8 x 8,
 * 16 x 4, 32 x 2 and row-major (reported in its tune[0]).  The image does
 * not depend on the order.  No counterpart in the reference, whose render
 * blocks are 32 x 32 pixels (backend/scene/view.cpp:258-261, 810). */
int pvx_scene_set_render_tile(pvx_scene *scene, uint32_t tile_w, uint32_t tile_h);

/* Per-object intersection list for n (ray, object) pairs.
 * Replaces virtual bool ObjectBase::All_Intersections(const Ray&, IStack&,
 *   TraceThreadData*) (core/scene/object.h:238) for the kinds above.
 * Writes up to max_hits depths per ray into depths[i*max_hits + k]
 * (sorted ascending) and prims likewise; counts[i] = number pushed
 * (may exceed max_hits: the list is then truncated). */
int pvx_all_intersections(pvx_scene *scene, const pvx_ray *rays,
                          const uint32_t *obj, size_t n, uint32_t max_hits,
                          double *depths, int32_t *prims, uint32_t *counts,
                          void *stream);

/* Point membership for n (point, object) pairs.
 * Replaces virtual bool ObjectBase::Inside(const Vector3d&, TraceThreadData*)
 *   (core/scene/object.h:240) as called through Inside_Object
 *   (core/scene/object.cpp:346-355). points: 3n doubles; out: n bytes 0/1. */
int pvx_inside(pvx_scene *scene, const double *points, const uint32_t *obj,
               size_t n, uint8_t *out, void *stream);

/* ---- scene batches ----------------------------------------------------- */
/* Many scenes traced in one launch.  The batch keeps a device table of the
 * scenes' device-resident BVHs and primitive pools; each wavefront takes up
 * to 64 rays of one scene, so scenes of uneven size share one persistent
 * grid.  The scenes stay owned by the caller and must outlive the batch.
 * No reference counterpart (the reference renders one scene per process):
 * for each scene the results are exactly those of pvx_trace_closest, i.e.
 * Trace::FindIntersection (core/render/trace.cpp:285-344). */
typedef struct pvx_batch pvx_batch;
int pvx_batch_create(pvx_scene *const *scenes, uint32_t n_scenes, pvx_batch **out);
int pvx_batch_destroy(pvx_batch *batch);

/* Closest hit for the rays of every scene of the batch.  rays, max_depth
 * (may be NULL) and hits are device buffers holding the scenes' rays one
 * after another; ray_first (HOST memory, n_scenes + 1 non-decreasing
 * entries, ray_first[0] = 0) gives scene s the rays
 * [ray_first[s], ray_first[s+1]); a scene may have none.  scene_clock
 * (optional device buffer of n_scenes + 1 entries) receives [0] = the
 * earliest wavefront start and [1 + s] = when scene s's last wavefront
 * finished (0 if it had no rays), in ticks of pvx_clock_rate. */
int pvx_batch_trace_closest(pvx_batch *batch, const pvx_ray *rays, const uint64_t *ray_first,
                            const double *max_depth, uint32_t ray_kind, pvx_hit *hits,
                            uint64_t *scene_clock, void *stream);

/* pvx_batch_trace_closest plus the per-ray work counters of
 * pvx_trace_closest_counted. */
int pvx_batch_trace_closest_counted(pvx_batch *batch, const pvx_ray *rays,
                                    const uint64_t *ray_first, const double *max_depth,
                                    uint32_t ray_kind, pvx_hit *hits, uint32_t *node_visits,
                                    uint32_t *prim_tests, uint64_t *scene_clock, void *stream);

/* Ticks per second of the device wall clock that scene_clock counts. */
int pvx_clock_rate(double *hz);

/* ---- camera ------------------------------------------------------------ */
/* Perspective camera as TracePixel::CreateCameraRay builds it
 * (core/render/tracepixel.cpp:357-372, 918-926): for pixel (x, y) sampled at
 * (x + jx, y + jy), dir = direction + ((x+jx)/w - 0.5)*right
 *                            + (0.5 - (y+jy)/h)*up, normalised. */
typedef struct pvx_camera {
    double location[3];
    double direction[3];
    double right[3];
    double up[3];
    uint32_t width;
    uint32_t height;
    double jitter_x; /* sample offset inside the pixel (0.5 = centre) */
    double jitter_y;
} pvx_camera;

/* Writes width*height rays (row-major, y down) into device buffer `rays`. */
int pvx_camera_rays(const pvx_camera *cam, pvx_ray *rays, void *stream);

/* The same rays in tile order: tiles of tile_w x tile_h pixels, the tiles
 * row-major and the pixels row-major inside each (the last tile column and
 * row hold only the pixels inside the image), and, when `pixel` is not
 * NULL, each ray's pixel y * width + x.  The reference traces pixels in
 * blocks of 32 x 32 (backend/scene/view.cpp:258-261, 810); 8 x 8 tiles keep
 * the 64 rays of a wavefront together on the screen (closest hit on the
 * 1.3M-triangle frame: 0.672 ms against 0.754 ms in row-major order). */
int pvx_camera_rays_tiled(const pvx_camera *cam, uint32_t tile_w, uint32_t tile_h, pvx_ray *rays,
                          uint32_t *pixel, void *stream);

/* ---- shading (plain finishes) ------------------------------------------ */
/* Surface description of one top-level object: pigment { rgb filter
 * transmit }, finish { ambient diffuse brilliance phong phong_size
 * reflection }, interior { ior } (core/material/texture.h FINISH, pigment,
 * interior; defaults as Create_Finish, core/material/texture.cpp:236-270). */
typedef struct pvx_material {
    float pigment[3];
    float filter;            /* pigment filter (0 = opaque)                   */
    float transmit;          /* pigment transmit (0 = opaque)                 */
    float ambient[3];        /* finish ambient, times the global ambient_light
                              * (classic ambient: ignored when radiosity is on) */
    float emission[3];       /* finish emission (default 0)                  */
    float diffuse;           /* default 0.6                                  */
    float brilliance;        /* default 1.0                                  */
    float phong;             /* default 0.0                                  */
    float phong_size;        /* default 40.0                                 */
    float reflection_min[3]; /* `reflection r` sets min = max = r            */
    float reflection_max[3];
    float reflection_falloff;/* default 1.0                                  */
    float ior;               /* interior ior (default 1.0; <= 0 means 1.0)    */
} pvx_material;

enum { PVX_LIGHT_POINT = 1, PVX_LIGHT_AREA = 2 };

/* light_source { position, color [area_light axis1, axis2, size1, size2] },
 * no jitter, default adaptive (core/lighting/lightsource.cpp:375-400). */
typedef struct pvx_light {
    uint32_t kind;
    uint32_t size1, size2;   /* area light grid (>= 1)                       */
    uint32_t pad;
    double position[3];
    double axis1[3];
    double axis2[3];
    float colour[3];
    float pad2;
} pvx_light;

typedef struct pvx_shading_desc {
    uint32_t n_materials;          /* == scene objects                      */
    const pvx_material *materials; /* host memory, copied                   */
    uint32_t n_lights;
    const pvx_light *lights;
    float ambient_light[3];        /* global_settings ambient_light (1,1,1)  */
    float background[3];           /* background colour (black)              */
    uint32_t max_trace_level;      /* default 5                              */
    float adc_bailout;             /* default 1/255                          */
} pvx_shading_desc;

typedef struct pvx_render_stats {
    uint64_t closest_rays;   /* camera + reflection + refraction rays traced  */
    uint64_t shadow_rays;    /* light-visibility rays traced                  */
    uint32_t bounces;        /* wavefront iterations run                      */
    uint32_t pad;
    double trace_ms;         /* GPU time in closest-hit launches              */
    double shadow_ms;        /* GPU time in shadow-ray launches (light rays
                                made in the kernel run on a second stream,
                                beside the next bounces: the two then overlap
                                and may add up to more than total_ms)        */
    double total_ms;         /* GPU time of the whole frame                   */
} pvx_render_stats;

/* Attach surface and light descriptions to a scene (replaces the textures,
 * interiors and light list the parser attaches to SceneData). */
int pvx_scene_set_shading(pvx_scene *scene, const pvx_shading_desc *desc);

/* Render one frame, 1 sample per pixel at the pixel centre, as
 * TracePixel -> Trace::TraceRay -> ComputeTextureColour ->
 * ComputeLightedTexture (core/render/trace.cpp:135-228, 457-586, 739-1179)
 * for plain finishes: classic ambient, diffuse and Phong per light
 * (ComputeOneDiffuseLight, trace.cpp:1637-1729) with shadow rays filtered by
 * transparent occluders (TracePointLightShadowRay / TraceAreaLightShadowRay /
 * ComputeShadowColour, trace.cpp:1951-2420), reflection (ComputeReflection,
 * trace.cpp:1264-1321) and refraction through the ray's interior list with
 * total internal reflection (ComputeRefraction / TraceRefractionRay,
 * trace.cpp:1323-1483) up to max_trace_level, ADC bailout.  Runs as a
 * wavefront: camera rays, then per bounce one closest-hit launch, one shade
 * launch, light-ray launches and the placement of the next bounce's rays.
 * rgb: device buffer of 3*width*height floats (linear RGB, row-major). */
int pvx_render(pvx_scene *scene, const pvx_camera *cam, float *rgb, pvx_render_stats *stats,
               void *stream);

/* Supersampled frame: n_samples rays per pixel at the given offsets from the
 * pixel corner (host array, 2 doubles each; cam->jitter_* is ignored), the
 * pixel being the mean of its samples.  Non-adaptive supersampling with
 * threshold 0 and no jitter (TraceTask::NonAdaptiveSupersamplingM1 /
 * SupersampleOnePixel, backend/render/tracetask.cpp:521-600, 838-890) is the
 * centre plus an aa_depth x aa_depth grid: offsets (0.5, 0.5) and
 * 0.5 + (-range .. range step 1/aa_depth), range = 0.5 - 0.5/aa_depth. */
int pvx_render_samples(pvx_scene *scene, const pvx_camera *cam, uint32_t n_samples, const double *offsets,
                       float *rgb, pvx_render_stats *stats, void *stream);

/* pvx_render that first runs pvx_scene_autotune on the frame's own camera
 * rays, generated in each candidate pixel order of the ray queue (8 x 8,
 * 16 x 4, 32 x 2 tiles and row-major; the fastest {tile, walk, grid} is
 * kept, tune[0] with its tile), and on its first wavefront of light rays
 * (tune[1], all zero when the frame has none); the image is the one
 * pvx_render gives.  Where light
 * rays are made in the kernel (PVX_TRAVERSAL_PACKET / FUSED / REFILL), the
 * walk and grid are timed on the light rays of camera-ray hits (reported in
 * tune[1]) and again on those of the first bounce, each set keeping its own
 * choice for later frames.  tune may be NULL.  stats->total_ms includes the
 * tuning launches; trace_ms and shadow_ms do not. */
int pvx_render_autotune(pvx_scene *scene, const pvx_camera *cam, float *rgb, pvx_tune_result *tune,
                        pvx_render_stats *stats, void *stream);

/* ---- radiosity ---------------------------------------------------------- */
enum { PVX_RADIOSITY_PLAIN = 0, PVX_RADIOSITY_LEARNED = 1 };

/* global_settings { radiosity { ... } } (SceneRadiositySettings defaults,
 * core/lighting/radiosity.h:141-166) plus the surrogate controls. */
typedef struct pvx_radiosity_settings {
    uint32_t count;            /* rays per gather at depth 0 (default 35)       */
    uint32_t recursion_limit;  /* default 2                                     */
    uint32_t nearest_count;    /* default 5                                     */
    uint32_t mode;             /* PVX_RADIOSITY_PLAIN / PVX_RADIOSITY_LEARNED   */
    float error_bound;         /* default 1.8                                   */
    float minimum_reuse;       /* default 0.015                                 */
    float maximum_reuse;       /* default 0.2                                   */
    float pretrace_start;      /* default 0.08                                  */
    float pretrace_end;        /* default 0.04                                  */
    float brightness;          /* default 1.0                                   */
    float adc_bailout;         /* default 0.01                                  */
    float probe_threshold;     /* learned: max mean probe error per tile in     
                                * displayed colour, diffuse reflectance x
                                * irradiance error (0.02)                      */
    uint32_t train_samples;    /* learned: plain training gathers per network   */
    uint32_t train_steps;      /* learned: optimiser steps per network          */
    uint32_t probe_tile;       /* learned: probe tile edge in pixels (32)       */
    uint32_t probes_per_tile;  /* learned: plain gathers probed per tile (2)    */
    uint32_t train_batch;      /* learned: samples per optimiser step, a
                                * multiple of 256 (0: 2048); the learning
                                * rate grows with its square root            */
} pvx_radiosity_settings;

typedef struct pvx_radiosity_stats {
    double frame_ms;           /* whole frame (GPU events)                      */
    double pretrace_ms;
    double final_ms;
    double train_ms;
    uint64_t gathers[2];       /* plain gathers at recursion depth 0 / 1        */
    uint64_t gather_rays;      /* rays traced by plain gathers                  */
    uint64_t records[2];       /* irradiance-cache records at depth 0 / 1       */
    uint64_t final_misses;     /* final-pass points with no reusable record     */
    uint64_t surrogate_queries;/* misses answered by the surrogate              */
    uint64_t fallbacks;        /* misses sent back to the plain gather          */
    uint64_t probes;           /* plain gathers spent on probing                */
    float probe_error;         /* mean abs probe error (surrogate vs gather)    */
    float train_loss;          /* final training loss                           */
} pvx_radiosity_stats;

/* Render one frame lit by radiosity: pretrace passes fill the irradiance
 * cache, the final pass reuses it and gathers where nothing is reusable
 * (RadiosityFunction::ComputeAmbient / GatherLight, RadiosityCache
 * FindReusableBlock / AverageNearBlock / AddBlock,
 * core/lighting/radiosity.cpp:385-743, 1194-1769).  Surfaces:
 * pigment * (emission + diffuse * irradiance).  LEARNED mode has no
 * pretrace: one batch of spread-out plain gathers (train_samples at depth 0
 * and as many at depth 1) trains a depth-1 and a depth-0 MLP irradiance
 * surrogate; the final pass reuses this frame's records where they reach and
 * takes the depth-0 network's answer at the misses, except in tiles whose
 * held-out probe gathers disagree (probe_threshold), which fall back to the
 * plain gather. */
/* Radiosity cache files (`save_file` / `load_file`, the .rca format of
 * RadiosityCache::Load and ot_write_block, core/lighting/radiosity.cpp:909-1000,
 * core/support/octree.cpp:1345-1370): one text line per record,
 *   C<depth+1> x y z <normal hex> r g b <harmonic mean> <nearest> <to-nearest hex>.
 * pvx_radiosity_save writes the records of the scene's last radiosity frame;
 * pvx_radiosity_load reads records that seed every later frame on the scene
 * (path NULL drops them); lines that are not 11-field "C" records are
 * skipped, as the reference does.  Host paths; n_* may be NULL. */
int pvx_radiosity_save(pvx_scene *scene, const char *path, uint64_t *n_written);
int pvx_radiosity_load(pvx_scene *scene, const char *path, uint64_t *n_read);

int pvx_radiosity_render(pvx_scene *scene, const pvx_camera *cam,
                         const pvx_radiosity_settings *settings, float *rgb,
                         pvx_radiosity_stats *stats, void *stream);

/* Opens a per-query irradiance cache on the scene, as constructing a
 * RadiosityFunction over the shared RadiosityCache does
 * (core/lighting/radiosity.h:257-300, radiosity.cpp:177-260): settings as
 * above (mode PVX_RADIOSITY_PLAIN only), cam->location is the camera
 * position that scales the reuse distances (GatherLight's to_eye).  Clears
 * the cache (records loaded by pvx_radiosity_load seed it); the cache then
 * grows with every pvx_radiosity_compute_ambient call until the next
 * pvx_radiosity_begin or pvx_radiosity_render on the scene. */
int pvx_radiosity_begin(pvx_scene *scene, const pvx_camera *cam,
                        const pvx_radiosity_settings *settings, void *stream);

/* Ambient colour for n surface points.
 * Replaces virtual void Trace::RadiosityFunctor::ComputeAmbient(
 *   const Vector3d& ipoint, const Vector3d& raw_normal,
 *   const Vector3d& layer_normal, double brilliance, MathColour& ambient_colour,
 *   double weight, TraceTicket& ticket)  (core/render/trace.h:210-216),
 * implemented by RadiosityFunction::ComputeAmbient
 * (core/lighting/radiosity.h:287, radiosity.cpp:385-461), batched:
 * ipoint, raw_normal: 3n doubles (device); layer_normal: 3n doubles or NULL
 * (= raw_normal; a layer normal not `similar` to its raw normal, i.e. a
 * perturbed normal, gives PVX_EUNSUPPORTED); weight: n doubles or NULL
 * (1.0); depth: ticket.radiosityRecursionDepth (< recursion_limit);
 * final_trace: RadiosityFunction::isFinalTrace.  brilliance is not an
 * argument: with the default `brilliance off` ComputeAmbient uses 1.0.
 * Each point reuses the cache where FindReusableBlock finds enough weight
 * and gathers otherwise (GatherLight), filing the new records; the points of
 * one call are gathered in rounds (a spread-out subset first), so later
 * points of the call reuse earlier ones' records as in the sequential code.
 * ambient_rgb: 3n floats (device), times `brightness` (gray_threshold 0).
 * stats (may be NULL): gathers and rays of this call, cache sizes after it. */
int pvx_radiosity_compute_ambient(pvx_scene *scene, size_t n, const double *ipoint,
                                  const double *raw_normal, const double *layer_normal,
                                  const double *weight, uint32_t depth, uint32_t final_trace,
                                  float *ambient_rgb, pvx_radiosity_stats *stats, void *stream);

/* Radiosity frames of every scene of a batch.  No reference counterpart
 * (the reference renders one scene per process): for each scene the image
 * and statistics are exactly those of pvx_radiosity_render with its camera
 * and settings (cams, settings: n_scenes host structs; rgb: n_scenes device
 * buffers of 3*width*height floats).  The frames run on n_streams HIP
 * streams at once (1..16), one host thread each, largest frames first; the
 * scenes' BVHs stay device-resident side by side.  Learned-mode training
 * launches (cooperative kernels) are serialised across the streams.
 * scene_stats (n_scenes, may be NULL) receives each frame's statistics,
 * scene_done_ms (n_scenes, may be NULL) the host wall time from the batch
 * start to each scene's last kernel. */
typedef struct pvx_batch_radiosity_stats {
    double batch_ms;            /* host wall clock of the whole batch            */
    double gpu_frame_ms;        /* sum of the frames' own GPU event times        */
    uint64_t pixels;            /* camera rays, one per pixel                    */
    uint64_t gather_rays;       /* rays traced by plain gathers                  */
    uint64_t surrogate_queries; /* final-pass misses answered by the surrogate   */
    uint64_t fallbacks;         /* misses sent back to the plain gather          */
    uint64_t final_misses;      /* final-pass points with no reusable record     */
    uint32_t streams;
    uint32_t pad;
} pvx_batch_radiosity_stats;

int pvx_batch_radiosity_render(pvx_batch *batch, const pvx_camera *cams,
                               const pvx_radiosity_settings *settings, float *const *rgb,
                               uint32_t n_streams, pvx_radiosity_stats *scene_stats,
                               double *scene_done_ms, pvx_batch_radiosity_stats *stats);

/* ---- irradiance surrogate ------------------------------------------------ */
/* No reference counterpart: a small MLP that stands in for
 * RadiosityFunction::GatherLight (core/lighting/radiosity.cpp:509-743) at
 * cache misses.  Input per point: 6 floats (position normalised to [0,1]^3,
 * unit normal); 16-frequency sin/cos features per input (pi 2^(k/3) x,
 * k = 0..15, a quarter of that for the normal) -> 3 ReLU layers of width 64 -> RGB irradiance.  BF16 MFMA with
 * FP32 accumulation; FP32 master weights and Adam (beta 0.9 / 0.99).
 * Parameter vector (FP32, PVX_SURROGATE_PARAMS): W1[192][64], b1[64],
 * W2[64][64], b2[64], W3[64][64], b3[64], W4[64][3], b4[3] (row = input). */
enum { PVX_SURROGATE_PARAMS = 20867 };
typedef struct pvx_surrogate pvx_surrogate;

/* params: host array of PVX_SURROGATE_PARAMS floats, or NULL for a seeded
 * He-uniform initialisation. */
int pvx_surrogate_create(const float *params, uint64_t seed, pvx_surrogate **out);
void pvx_surrogate_destroy(pvx_surrogate *s);
int pvx_surrogate_get_params(pvx_surrogate *s, float *params /* host */);
/* Loss L = (1/n) sum_p |y_p - target_p|^2 and dL/dparams (host outputs) over
 * the n points (device x: n*6, target: n*3). */
int pvx_surrogate_grad(pvx_surrogate *s, const float *x, const float *target, size_t n, float *loss,
                       float *grad, void *stream);
/* `steps` Adam steps on seeded mini-batches of `batch` points (multiple of
 * 256) drawn from the n samples; learning rate decays from lr to lr/10. */
int pvx_surrogate_train(pvx_surrogate *s, const float *x, const float *target, size_t n, uint32_t steps,
                        uint32_t batch, float lr, float *final_loss, void *stream);
/* out (device, n*3) = network(x). */
int pvx_surrogate_infer(pvx_surrogate *s, const float *x, size_t n, float *out, void *stream);

/* ---- utility ----------------------------------------------------------- */
/* Plain device-to-device copy kernel (16 B per lane per step).  No reference
 * counterpart: it is the known-size workload used to calibrate hardware byte
 * counters (FETCH_SIZE / WRITE_SIZE) against the trace kernels.  nbytes must
 * be a multiple of 16 and both pointers 16 B aligned. */
int pvx_copy(void *dst, const void *src, size_t nbytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PVX_H */
