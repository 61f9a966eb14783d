// Host-side scene preparation: pvx_scene_desc -> device-ready arrays.
#ifndef PVX_SCENE_PREP_H
#define PVX_SCENE_PREP_H

#include <string>
#include <vector>

#include "pvx_types.h"

namespace pvx {

struct HostScene {
    std::vector<PrimRec> prims; // BVH leaf order
    std::vector<Node> nodes;
    int32_t root = INT32_MIN;
    uint32_t depth = 0;
    std::vector<KdNode> kd_nodes; // kd-tree (PVX_ACCEL_BSP): nodes, leaf lists, bounds
    std::vector<uint32_t> kd_list;
    float kd_bmin[3] = {0, 0, 0}, kd_bmax[3] = {0, 0, 0};
    uint32_t kd_depth = 0;
    bool use_kd = false;
    std::vector<Node4> nodes4; // 4-wide tree (triangle / sphere scenes)
    int32_t root4 = INT32_MIN;
    uint32_t depth4 = 0;
    std::vector<InfRec> inf;
    std::vector<uint32_t> obj_flags;
    std::vector<uint32_t> obj_type;
    std::vector<PlaneRec> planes;
    std::vector<BoxRec> boxes;
    std::vector<TriangleRec> triangles;
    std::vector<TorusRec> tori;
    std::vector<ConeRec> cones;
    std::vector<DiscRec> discs;
    std::vector<QuadricRec> quadrics;
    std::vector<PolygonRec> polygons;
    std::vector<double> poly_pts;
    std::vector<RevRec> revs;            // lathes and sors
    std::vector<RevSeg> rev_segs;
    std::vector<double> rev_radius, rev_height; // bounding cylinder lists (radii squared)
    std::vector<Xform> xforms;
    std::vector<double> sph; // 4 per sphere: centre, radius
    std::vector<CsgNode> csg;
    std::vector<CsgChild> csg_children;
    std::vector<uint32_t> csg_leaves;
    std::vector<uint32_t> csg_leaf_range;
    std::vector<ShadeTri> shade_tris;
    std::vector<uint32_t> obj_tri_base;
    std::vector<int32_t> obj_mesh_xf; // per object: a mesh's placement (xforms index) or -1
    std::vector<float> mesh_vtx, mesh_nrm;
    uint32_t any_flags = 0;
    uint32_t has_complex = 0; // 0: triangles/spheres, 1: + boxes, 2: + tori or CSG, 3: + lathes or sors
    uint64_t n_triangles = 0;
    double build_ms = 0.0;
    double bvh_ms = 0.0;         // the tree build alone
    bool gpu_build = false;      // PVX_BUILDER_GPU: prims in input order, tree built on the device
    std::vector<float> prim_lo, prim_hi; // their FP32 bounds (GPU build)
};

// Validates the description and builds every host array plus the BVH.
// Returns a PVX_* status; `err` explains a failure.
int prepare_scene(const pvx_scene_desc &d, const pvx_build_opts *opts, HostScene &out,
                  std::string &err);

} // namespace pvx

#endif
