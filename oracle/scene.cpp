// FP64 CPU ORACLE — TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Scene-level restatement: object table, Find_Intersection per object
// (core/scene/object.cpp:118-162 and the postcondition variant 165-224),
// Trace::FindIntersection methods 0 and 1 (core/render/trace.cpp:285-344),
// Intersect_BBox_Tree (core/bounding/boundingbox.cpp:485-539), ray
// pre/postconditions (trace.cpp:84-95, 1941-1949), and a C API for tests.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <thread>

#include "oracle.hpp"

namespace orc {

namespace {
const double kPlaneDepthTol = 1.0e-6;  // plane.cpp:62
const double kSphereDepthTol = 1.0e-6; // sphere.cpp:63
const double kBoxDepthTol = 1.0e-6;    // box.cpp:67
const double kInfiniteVolume = kBoundHuge; // parser.cpp:9286 INFINITE_VOLUME
} // namespace

// Compute_BBox per kind: Sphere (sphere.cpp:646-658), Plane (plane.cpp),
// Mesh (mesh.cpp), Box (box.cpp:947-960), Torus (torus.cpp), CSG
// (csg.cpp:872-990); INFINITE_FLAG when the box volume exceeds
// INFINITE_VOLUME (parser.cpp:9284-9292).
FBox Scene::shape_bbox(uint32_t type, uint32_t index, uint32_t flags, bool *infinite) const {
    FBox b;
    switch (type) {
    case PVX_SPHERE: {
        V3 c = sph_c[index];
        double r = sph_r[index];
        for (int k = 0; k < 3; k++) {
            b.ll[k] = (float)(c[k] - r);
            b.size[k] = (float)(2.0 * r);
        }
        break;
    }
    case PVX_PLANE:
        for (int k = 0; k < 3; k++) {
            b.ll[k] = (float)(-kBoundHuge / 2);
            b.size[k] = (float)kBoundHuge;
        }
        break;
    case PVX_MESH:
        b = meshes[index].bbox;
        break;
    case PVX_BOX:
        b = box_bbox(box_c1[index], box_c2[index], tr(box_tr[index]));
        break;
    case PVX_TORUS:
        b = torus_bbox(tor_R[index], tor_r[index], tr(tor_tr[index]));
        break;
    case PVX_CONE:
        b = cone_bbox(cones[index]);
        break;
    case PVX_DISC:
        b = disc_bbox(discs[index]);
        break;
    case PVX_POLYGON:
        b = polygons[index].bbox;
        break;
    case PVX_QUADRIC: // Quadric::Compute_BBox at parse time (see quadric.cpp)
        b = quadrics[index].bbox;
        break;
    case PVX_LATHE: // Lathe::Compute_BBox (lathe.cpp:455-461)
    case PVX_SOR:   // Sor::Compute_BBox (sor.cpp:705-711)
        b = revs[index].bbox;
        break;
    case PVX_TRIANGLE:
    case PVX_SMOOTH_TRIANGLE:
        b = tri_bbox(&tri_p[9 * (size_t)index]);
        break;
    case PVX_CSG: {
        // CSG::Compute_BBox runs when the CSG is parsed, before any later
        // Invert() (which keeps the box): undo this node's inversion to get
        // the operator and child flags it had then.
        const pvx_csg_node &n = csg[index];
        bool inv = (flags & PVX_FLAG_INVERTED) != 0;
        bool isect = inv ? (n.op == PVX_CSG_MERGE) : (n.op == PVX_CSG_INTERSECTION);
        FBox dflt; // ObjectBase default: (-BOUND_HUGE/2, BOUND_HUGE) per axis
        for (int k = 0; k < 3; k++) {
            dflt.ll[k] = (float)(-kBoundHuge / 2);
            dflt.size[k] = (float)kBoundHuge;
        }
        V3 nmin, nmax;
        if (isect) {
            nmin = V3(-kBoundHuge, -kBoundHuge, -kBoundHuge);
            nmax = V3(kBoundHuge, kBoundHuge, kBoundHuge);
        } else {
            nmin = V3(kBoundHuge, kBoundHuge, kBoundHuge);
            nmax = V3(-kBoundHuge, -kBoundHuge, -kBoundHuge);
        }
        for (uint32_t k = 0; k < n.n_children; k++) {
            const pvx_object &c = csg_children[n.first_child + k];
            bool c_inv = ((c.flags & PVX_FLAG_INVERTED) != 0) != inv;
            // csg.cpp:884-906: inverted children do not bound an intersection
            if (isect && c_inv) continue;
            bool cinf;
            FBox cb = shape_bbox(c.type, c.index, c.flags, &cinf);
            V3 lo(cb.ll[0], cb.ll[1], cb.ll[2]);
            V3 hi = lo + V3(cb.size[0], cb.size[1], cb.size[2]);
            for (int a = 0; a < 3; a++) {
                if (isect) {
                    nmin[a] = std::max(nmin[a], lo[a]);
                    nmax[a] = std::min(nmax[a], hi[a]);
                } else {
                    nmin[a] = std::min(nmin[a], lo[a]);
                    nmax[a] = std::max(nmax[a], hi[a]);
                }
            }
        }
        b = dflt;
        if (!(nmin.x > nmax.x || nmin.y > nmax.y || nmin.z > nmax.z)) {
            double nv = (nmax.x - nmin.x) * (nmax.y - nmin.y) * (nmax.z - nmin.z);
            double ov = (double)dflt.size[0] * dflt.size[1] * dflt.size[2];
            if (nv < ov) {
                b = make_bbox_from_min_max(nmin, nmax);
                if (b.size[0] > 1.0e6 || b.size[1] > 1.0e6 || b.size[2] > 1.0e6) b = dflt;
            }
        }
        break;
    }
    }
    double vol = (double)b.size[0] * b.size[1] * b.size[2];
    *infinite = vol > kInfiniteVolume;
    (void)flags;
    return b;
}

// ObjectBase::All_Intersections dispatch; `slot` is reported as the hit's
// primitive id for non-mesh shapes (the leaf's csg_children position inside
// a CSG, 0 at top level).
bool Scene::all_intersections(uint32_t type, uint32_t index, uint32_t flags, const Ray &r,
                              std::vector<Isect> &st, int32_t slot, bool child) const {
    switch (type) {
    case PVX_SPHERE: { // Sphere::All_Intersections (sphere.cpp:92-178), spherical mode
        double d1, d2;
        bool found = false;
        double rad = sph_r[index];
        if (sphere_intersect(r, sph_c[index], rad * rad, &d1, &d2)) {
            if (d1 > kSphereDepthTol && d1 < kMaxDistance) {
                st.push_back(Isect{d1, slot, 0, slot});
                found = true;
            }
            if (d2 > kSphereDepthTol && d2 < kMaxDistance) {
                st.push_back(Isect{d2, slot, 0, slot});
                found = true;
            }
        }
        return found;
    }
    case PVX_PLANE: { // Plane::Intersect (plane.cpp:110-170), untransformed
        V3 n = pl_n[index];
        double ndd = dot(n, r.d);
        if (std::fabs(ndd) < kEpsilon) return false;
        double ndo = dot(n, r.o);
        double depth = -(ndo + pl_d[index]) / ndd;
        if (depth >= kPlaneDepthTol && depth <= kMaxDistance) {
            st.push_back(Isect{depth, slot, 0, slot});
            return true;
        }
        return false;
    }
    case PVX_MESH: {
        size_t before = st.size();
        mesh_all_intersections(meshes[index], tr(mesh_tr[index]), r, st, true);
        return st.size() > before;
    }
    case PVX_BOX: { // Box::All_Intersections (box.cpp:100-150)
        double d1, d2;
        int s1, s2;
        bool found = false;
        if (box_intersect(r, tr(box_tr[index]), box_c1[index], box_c2[index], &d1, &d2, &s1, &s2)) {
            if (d1 > kBoxDepthTol) {
                st.push_back(Isect{d1, slot, s1, slot});
                found = true;
            }
            st.push_back(Isect{d2, slot, s2, slot});
            found = true;
        }
        return found;
    }
    case PVX_TORUS:
        return torus_all_intersections(r, tr(tor_tr[index]), tor_R[index], tor_r[index],
                                       tor_sturm[index] != 0, st, slot);
    case PVX_CONE: { // Cone::All_Intersections (cone.cpp:104-129), no clipped_by
        double dep[4];
        int part[4];
        const int n = cone_intersect(cones[index], r, dep, part);
        for (int k = 0; k < n; k++) st.push_back(Isect{dep[k], slot, part[k], slot});
        return n > 0;
    }
    case PVX_POLYGON: { // Polygon::All_Intersections (polygon.cpp:131-160)
        double dep;
        if (!polygon_intersect(polygons[index], r, &dep)) return false;
        st.push_back(Isect{dep, slot, 0, slot});
        return true;
    }
    case PVX_QUADRIC: { // Quadric::All_Intersections (quadric.cpp:100-135), no clipped_by
        double dep[2];
        const int n = quadric_intersect(quadrics[index], r, dep);
        for (int k = 0; k < n; k++) st.push_back(Isect{dep[k], slot, 0, slot});
        return n > 0;
    }
    case PVX_LATHE: // Lathe::All_Intersections (lathe.cpp:176-187), no clipped_by
    case PVX_SOR:   // Sor::All_Intersections (sor.cpp:193-205), no clipped_by
        return revolution_intersect(revs[index], r, child, st, slot);
    case PVX_DISC: { // Disc::All_Intersections (disc.cpp), no clipped_by
        double dep;
        if (!disc_intersect(discs[index], r, &dep)) return false;
        st.push_back(Isect{dep, slot, 0, slot});
        return true;
    }
    case PVX_TRIANGLE:
    case PVX_SMOOTH_TRIANGLE: { // Triangle::All_Intersections (triangle.cpp:426-445)
        double depth;
        if (!tri_intersect(tris[index], r, &depth)) return false;
        st.push_back(Isect{depth, slot, 0, slot});
        return true;
    }
    case PVX_CSG: {
        const pvx_csg_node &n = csg[index];
        bool found = false;
        if (n.op == PVX_CSG_UNION) { // CSGUnion::All_Intersections (csg.cpp:128-200)
            for (uint32_t k = 0; k < n.n_children; k++) {
                const pvx_object &c = csg_children[n.first_child + k];
                if (all_intersections(c.type, c.index, c.flags, r, st, (int32_t)(n.first_child + k), true))
                    found = true;
            }
            return found;
        }
        // CSGIntersection (csg.cpp:219-277) / CSGMerge (csg.cpp:307-375):
        // each child hit is kept if it is inside all (intersection) / none
        // (merge) of the other siblings, tested at the hit point.
        bool is_merge = (n.op == PVX_CSG_MERGE);
        std::vector<Isect> local;
        for (uint32_t k = 0; k < n.n_children; k++) {
            const pvx_object &c = csg_children[n.first_child + k];
            local.clear();
            if (!all_intersections(c.type, c.index, c.flags, r, local, (int32_t)(n.first_child + k), true))
                continue;
            while (!local.empty()) {
                Isect top = local.back();
                V3 ip = r.eval(top.depth);
                bool keep = true;
                for (uint32_t j = 0; j < n.n_children && keep; j++) {
                    if (j == k) continue;
                    const pvx_object &sib = csg_children[n.first_child + j];
                    bool in = inside(sib.type, sib.index, sib.flags, ip);
                    if (is_merge ? in : !in) keep = false;
                }
                if (keep) {
                    st.push_back(top);
                    found = true;
                }
                local.pop_back();
            }
        }
        return found;
    }
    }
    return false;
}

// ObjectBase::Inside dispatch (through Inside_Object, object.cpp:346-355).
bool Scene::inside(uint32_t type, uint32_t index, uint32_t flags, V3 p) const {
    bool inv = (flags & PVX_FLAG_INVERTED) != 0;
    switch (type) {
    case PVX_SPHERE: { // Sphere::Inside (sphere.cpp:271-303)
        double oc2 = length_sqr(sph_c[index] - p);
        double r2 = sph_r[index] * sph_r[index];
        return inv ? (oc2 > r2) : (oc2 < r2);
    }
    case PVX_PLANE: { // Plane::Inside (plane.cpp:192-215)
        double t = dot(p, pl_n[index]);
        return inv ? !((t + pl_d[index]) < kEpsilon) : ((t + pl_d[index]) < kEpsilon);
    }
    case PVX_MESH:
    case PVX_POLYGON: // Polygon::Inside (polygon.cpp:269-272): never inside
    case PVX_TRIANGLE: // Triangle::Inside (triangle.cpp:608-611): never inside
    case PVX_SMOOTH_TRIANGLE:
        return inv;
    case PVX_BOX:
        return box_inside(p, tr(box_tr[index]), box_c1[index], box_c2[index], inv);
    case PVX_TORUS:
        return torus_inside(p, tr(tor_tr[index]), tor_R[index], tor_r[index], inv);
    case PVX_CONE:
        return cone_inside(cones[index], p, inv);
    case PVX_DISC:
        return disc_inside(discs[index], p, inv);
    case PVX_QUADRIC:
        return quadric_inside(quadrics[index], p, inv);
    case PVX_LATHE:
    case PVX_SOR:
        return revolution_inside(revs[index], p, inv);
    case PVX_CSG: {
        // CSGUnion/CSGMerge::Inside (csg.cpp:405-424), CSGIntersection::Inside
        // (csg.cpp:440-454); the CSG's own INVERTED flag is carried by its
        // children after Invert (csg.cpp:629-645).
        const pvx_csg_node &n = csg[index];
        for (uint32_t k = 0; k < n.n_children; k++) {
            const pvx_object &c = csg_children[n.first_child + k];
            bool in = inside(c.type, c.index, c.flags, p);
            if (n.op == PVX_CSG_INTERSECTION) {
                if (!in) return false;
            } else if (in) {
                return true;
            }
        }
        return n.op == PVX_CSG_INTERSECTION;
    }
    }
    return false;
}

// Find_Intersection (object.cpp:118-224) / Trace::FindIntersection(object,...)
// (trace.cpp:346-450): the FP32 object-box pretest of ObjectBase::Intersect_BBox
// (object.cpp:917-942, 1074-1108) for every shape that does not override it
// (Sphere, Box, Plane, Triangle and Poly do: sphere.h:92, box.h:106,
// plane.h:94, triangle.h:99, polynomial.h:104), then the IStack scan from the
// top.
bool Scene::find_object(uint32_t obj, const Ray &r, bool shadow, Isect &out) const {
    const Object &o = objects[obj];
    double closest = HUGE_VAL;
    if (o.type == PVX_MESH || o.type == PVX_TORUS || o.type == PVX_CSG || o.type == PVX_CONE ||
        o.type == PVX_DISC || o.type == PVX_POLYGON || o.type == PVX_QUADRIC || o.type == PVX_LATHE ||
        o.type == PVX_SOR) {
        if (!intersect_bbox_dir(o.bbox, r, (float)kMinIsectDepth, (float)closest)) return false;
    }
    std::vector<Isect> st;
    if (!all_intersections(o.type, o.index, o.flags, r, st, 0)) return false;
    bool found = false;
    while (!st.empty()) {
        const Isect &top = st.back();
        double d = top.depth;
        if (d < closest && d >= kMinIsectDepth && (!shadow || d > kSmallTolerance)) {
            out = top;
            closest = d;
            found = true;
        }
        st.pop_back();
    }
    return found;
}

static bool precondition(uint32_t flags, uint32_t kind) {
    if (kind == PVX_RAY_SHADOW) return !(flags & PVX_FLAG_NO_SHADOW);
    if ((kind & PVX_RAY_IMAGE) && (flags & PVX_FLAG_NO_IMAGE)) return false;
    if ((kind & PVX_RAY_REFLECTION) && (flags & PVX_FLAG_NO_REFLECTION)) return false;
    if ((kind & PVX_RAY_RADIOSITY) && (flags & PVX_FLAG_NO_RADIOSITY)) return false;
    return true;
}


static Scene *scene_create(const pvx_scene_desc &d) {
    std::unique_ptr<Scene> holder(new Scene()); // freed if preparation throws
    Scene *s = holder.get();
    for (uint32_t i = 0; i < d.n_spheres; i++) {
        s->sph_c.push_back(V3(d.sphere_center[3 * i], d.sphere_center[3 * i + 1],
                              d.sphere_center[3 * i + 2]));
        s->sph_r.push_back(d.sphere_radius[i]);
    }
    for (uint32_t i = 0; i < d.n_planes; i++) {
        s->pl_n.push_back(V3(d.plane_normal[3 * i], d.plane_normal[3 * i + 1],
                             d.plane_normal[3 * i + 2]));
        s->pl_d.push_back(d.plane_distance[i]);
    }
    s->meshes.resize(d.n_meshes);
    for (uint32_t i = 0; i < d.n_meshes; i++) mesh_prepare(s->meshes[i], d.meshes[i]);
    for (uint32_t i = 0; i < d.n_transforms; i++) {
        Transform t;
        std::memcpy(t.m, d.transforms + 32 * (size_t)i, 16 * sizeof(double));
        std::memcpy(t.inv, d.transforms + 32 * (size_t)i + 16, 16 * sizeof(double));
        s->trs.push_back(t);
    }
    s->mesh_tr.assign(d.n_meshes, -1);
    for (uint32_t i = 0; i < d.n_meshes; i++) { // Mesh::Transform: Recompute_BBox of the mesh box
        const int32_t t = d.mesh_transform ? d.mesh_transform[i] : -1;
        if (t < -1 || t >= (int32_t)d.n_transforms) throw std::runtime_error("mesh transform index out of range");
        s->mesh_tr[i] = t;
        if (t >= 0) s->meshes[i].bbox = recompute_bbox(s->meshes[i].bbox, &s->trs[t]);
    }
    for (uint32_t i = 0; i < d.n_boxes; i++) {
        s->box_c1.push_back(V3(d.box_corner1[3 * i], d.box_corner1[3 * i + 1], d.box_corner1[3 * i + 2]));
        s->box_c2.push_back(V3(d.box_corner2[3 * i], d.box_corner2[3 * i + 1], d.box_corner2[3 * i + 2]));
        s->box_tr.push_back(d.box_transform ? d.box_transform[i] : -1);
    }
    for (uint32_t i = 0; i < d.n_tori; i++) {
        s->tor_R.push_back(d.torus_major[i]);
        s->tor_r.push_back(d.torus_minor[i]);
        s->tor_tr.push_back(d.torus_transform ? d.torus_transform[i] : -1);
        s->tor_sturm.push_back(d.torus_sturm ? d.torus_sturm[i] : 0);
    }
    auto check_tr = [&](int32_t t) {
        if (t >= (int32_t)d.n_transforms) throw std::runtime_error("transform index out of range");
    };
    for (uint32_t i = 0; i < d.n_cones; i++) {
        const int32_t t = d.cone_transform ? d.cone_transform[i] : -1;
        check_tr(t);
        s->cones.push_back(cone_prepare(d.cone_apex + 3 * (size_t)i, d.cone_base + 3 * (size_t)i,
                                        d.cone_radius + 2 * (size_t)i, d.cone_flags ? d.cone_flags[i] : 0,
                                        t >= 0 ? &s->trs[t] : nullptr));
    }
    for (uint32_t i = 0; i < d.n_discs; i++) {
        const int32_t t = d.disc_transform ? d.disc_transform[i] : -1;
        check_tr(t);
        s->discs.push_back(disc_prepare(d.disc_center + 3 * (size_t)i, d.disc_normal + 3 * (size_t)i,
                                        d.disc_radius + 2 * (size_t)i, t >= 0 ? &s->trs[t] : nullptr));
    }
    for (uint32_t i = 0; i < d.n_quadrics; i++) {
        const int32_t t = d.quadric_transform ? d.quadric_transform[i] : -1;
        check_tr(t);
        s->quadrics.push_back(quadric_prepare(d.quadric_coef + 10 * (size_t)i, t >= 0 ? &s->trs[t] : nullptr));
    }
    for (uint32_t i = 0; i < d.n_polygons; i++) {
        const int32_t t = d.polygon_transform ? d.polygon_transform[i] : -1;
        check_tr(t);
        const uint32_t b = d.polygon_first[i], e = d.polygon_first[i + 1];
        if (e < b + 3) throw std::runtime_error("polygon with fewer than three points");
        s->polygons.push_back(polygon_prepare(d.polygon_points + 3 * (size_t)b, (int)(e - b),
                                              t >= 0 ? &s->trs[t] : nullptr));
    }
    for (uint32_t i = 0; i < d.n_revolutions; i++) {
        const int32_t t = d.revolution_transform ? d.revolution_transform[i] : -1;
        check_tr(t);
        const uint32_t b = d.revolution_first[i], e = d.revolution_first[i + 1];
        if (e < b) throw std::runtime_error("revolution point offsets decrease");
        const uint8_t kind = d.revolution_kind[i], fl = d.revolution_flags ? d.revolution_flags[i] : 0;
        const Transform *user = t >= 0 ? &s->trs[t] : nullptr;
        const double *pts = d.revolution_points + 2 * (size_t)b;
        if (kind == PVX_SOR_CUBIC)
            s->revs.push_back(sor_prepare(pts, (int)(e - b), (fl & PVX_REV_OPEN) != 0, (fl & PVX_REV_STURM) != 0, user));
        else
            s->revs.push_back(lathe_prepare(pts, (int)(e - b), kind, (fl & PVX_REV_STURM) != 0, user));
        if (s->revs.back().segments() > 64) throw std::runtime_error("more than 64 segments in a lathe or sor");
    }
    if (d.n_triangles) s->tri_p.assign(d.triangle_vertices, d.triangle_vertices + 9 * (size_t)d.n_triangles);
    s->tris.resize(d.n_triangles);
    for (uint32_t i = 0; i < d.n_objects; i++) {
        const pvx_object &o = d.objects[i];
        if ((o.type == PVX_TRIANGLE || o.type == PVX_SMOOTH_TRIANGLE) && o.index < d.n_triangles) {
            bool sm = o.type == PVX_SMOOTH_TRIANGLE;
            s->tris[o.index] = make_tri(d.triangle_vertices + 9 * (size_t)o.index,
                                        sm ? d.triangle_normals + 9 * (size_t)o.index : nullptr, sm);
        }
    }
    s->csg.assign(d.csg, d.csg + d.n_csg);
    s->csg_children.assign(d.csg_children, d.csg_children + d.n_csg_children);
    auto check_rev = [&](const pvx_object &o) {
        if (o.type != PVX_LATHE && o.type != PVX_SOR) return;
        if (o.index >= s->revs.size() || s->revs[o.index].sor != (o.type == PVX_SOR))
            throw std::runtime_error("a lathe or sor object refers to a missing shape or one of the other kind");
    };
    for (uint32_t i = 0; i < d.n_objects; i++) check_rev(d.objects[i]);
    for (uint32_t i = 0; i < d.n_csg_children; i++) check_rev(d.csg_children[i]);
    std::vector<FBox> fb, ib;
    std::vector<int32_t> fi, ii;
    for (uint32_t i = 0; i < d.n_objects; i++) {
        Object o;
        o.type = d.objects[i].type;
        o.index = d.objects[i].index;
        o.flags = d.objects[i].flags;
        o.bbox = s->shape_bbox(o.type, o.index, o.flags, &o.infinite);
        s->objects.push_back(o);
        if (o.infinite) {
            ib.push_back(o.bbox);
            ii.push_back((int32_t)i);
        } else {
            fb.push_back(o.bbox);
            fi.push_back((int32_t)i);
        }
    }
    // Build_Bounding_Slabs (boundingbox.cpp:325-430); BoundingTask uses it
    // when objects >= Bounding_Threshold (3), else brute force
    // (boundingtask.cpp:168-210).
    s->tree.build(fb, fi, ib, ii);
    return holder.release();
}

// Trace::FindIntersection (trace.cpp:285-344), method 0 or 1.
static bool find_intersection(const Scene &s, const Ray &r, double max_depth, uint32_t kind,
                              int method, Isect &best, int32_t &best_obj) {
    best.depth = (max_depth >= kEpsilon) ? max_depth : kBoundHuge; // trace.cpp:157-158
    best_obj = -1;
    bool shadow = (kind == PVX_RAY_SHADOW);
    bool found = false;
    if (method == 1 && s.tree.root >= 0) {
        // Intersect_BBox_Tree (boundingbox.cpp:485-539)
        RayInfo ri(r);
        PQueue q;
        check_and_enqueue(q, s.tree.nodes[s.tree.root], s.tree.root, ri);
        double d;
        int32_t node;
        while (!q.empty()) {
            q.remove_min(d, node);
            if (d > best.depth) break;
            const BNode &n = s.tree.nodes[node];
            if (!n.kids.empty()) {
                for (int32_t k : n.kids) check_and_enqueue(q, s.tree.nodes[k], k, ri);
            } else {
                uint32_t obj = (uint32_t)n.element;
                if (!precondition(s.objects[obj].flags, kind)) continue;
                Isect is;
                if (s.find_object(obj, r, shadow, is) && is.depth < best.depth) {
                    best = is;
                    best_obj = (int32_t)obj;
                    found = true;
                }
            }
        }
        return found;
    }
    for (uint32_t obj = 0; obj < s.objects.size(); obj++) {
        if (!precondition(s.objects[obj].flags, kind)) continue;
        Isect is;
        if (s.find_object(obj, r, shadow, is) && is.depth < best.depth) {
            best = is;
            best_obj = (int32_t)obj;
            found = true;
        }
    }
    return found;
}

bool scene_find_intersection(const Scene &s, const Ray &r, double max_depth, uint32_t kind,
                             Isect &best, int32_t &best_obj, int method) {
    return find_intersection(s, r, max_depth, kind, method, best, best_obj);
}

template <class F>
static void parallel_for(size_t n, int threads, F f) {
    if (threads <= 1 || n < 64) {
        for (size_t i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<size_t> next(0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&]() {
            for (;;) {
                size_t b = next.fetch_add(256);
                if (b >= n) break;
                size_t e = std::min(n, b + 256);
                for (size_t i = b; i < e; i++) f(i);
            }
        });
    for (auto &t : pool) t.join();
}

} // namespace orc

using namespace orc;

extern "C" {

void *orc_scene_create(const pvx_scene_desc *desc) {
    if (!desc) return nullptr;
    try {
        return scene_create(*desc);
    } catch (const std::exception &) { // e.g. a degenerate cone (the reference throws too)
        return nullptr;
    }
}

void orc_scene_destroy(void *s) { delete static_cast<Scene *>(s); }

// Box::Intersect on one ray: corners c1 <= c2, optional transform (16
// doubles matrix then 16 inverse, row-vector convention), ray origin +
// direction; out = depth1, depth2, side1, side2 (kSideHit numbering 1..6 =
// -x, +x, -y, +y, -z, +z, 0 = none).
int orc_box_intersect(const double *c1, const double *c2, const double *tr32, const double *ray6,
                      double *out4) {
    Transform t;
    if (tr32) {
        std::memcpy(t.m, tr32, sizeof t.m);
        std::memcpy(t.inv, tr32 + 16, sizeof t.inv);
    }
    Ray r{V3(ray6[0], ray6[1], ray6[2]), V3(ray6[3], ray6[4], ray6[5])};
    double d1 = 0, d2 = 0;
    int s1 = 0, s2 = 0;
    const bool hit = box_intersect(r, tr32 ? &t : nullptr, V3(c1[0], c1[1], c1[2]), V3(c2[0], c2[1], c2[2]), &d1,
                                   &d2, &s1, &s2);
    out4[0] = d1;
    out4[1] = d2;
    out4[2] = s1;
    out4[3] = s2;
    return hit ? 1 : 0;
}

// The object's FP32 bounding box (lowerLeft, size) and infinite flag.
int orc_object_bbox(void *sp, uint32_t obj, float *ll_size6) {
    const Scene &s = *static_cast<Scene *>(sp);
    const Object &o = s.objects[obj];
    for (int k = 0; k < 3; k++) {
        ll_size6[k] = o.bbox.ll[k];
        ll_size6[3 + k] = o.bbox.size[k];
    }
    return o.infinite ? 1 : 0;
}

int orc_find_intersection(void *sp, const pvx_ray *rays, const double *max_depth,
                          uint32_t ray_kind, size_t n, int method, pvx_hit *hits, int threads) {
    const Scene &s = *static_cast<Scene *>(sp);
    parallel_for(n, threads, [&](size_t i) {
        Ray r{V3(rays[i].o[0], rays[i].o[1], rays[i].o[2]),
              V3(rays[i].d[0], rays[i].d[1], rays[i].d[2])};
        Isect best;
        int32_t obj;
        double md = max_depth ? max_depth[i] : 0.0;
        find_intersection(s, r, md, ray_kind, method, best, obj);
        hits[i].t = best.depth;
        hits[i].obj = obj;
        hits[i].prim = obj >= 0 ? best.prim : 0;
    });
    return 0;
}

int orc_all_intersections(void *sp, const pvx_ray *rays, const uint32_t *obj, size_t n,
                          uint32_t max_hits, double *depths, int32_t *prims, uint32_t *counts) {
    const Scene &s = *static_cast<Scene *>(sp);
    for (size_t i = 0; i < n; i++) {
        Ray r{V3(rays[i].o[0], rays[i].o[1], rays[i].o[2]),
              V3(rays[i].d[0], rays[i].d[1], rays[i].d[2])};
        const auto &o = s.objects[obj[i]];
        std::vector<Isect> st;
        s.all_intersections(o.type, o.index, o.flags, r, st, 0);
        std::stable_sort(st.begin(), st.end(),
                         [](const Isect &a, const Isect &b) { return a.depth < b.depth; });
        counts[i] = (uint32_t)st.size();
        for (uint32_t k = 0; k < max_hits; k++) {
            depths[i * max_hits + k] = k < st.size() ? st[k].depth : HUGE_VAL;
            prims[i * max_hits + k] = k < st.size() ? st[k].prim : -1;
        }
    }
    return 0;
}

int orc_inside(void *sp, const double *pts, const uint32_t *obj, size_t n, uint8_t *out) {
    const Scene &s = *static_cast<Scene *>(sp);
    for (size_t i = 0; i < n; i++) {
        const auto &o = s.objects[obj[i]];
        out[i] = s.inside(o.type, o.index, o.flags, V3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]))
                     ? 1 : 0;
    }
    return 0;
}

int orc_tree_info(void *sp, int mesh, int64_t *n_nodes, int32_t *depth) {
    const Scene &s = *static_cast<Scene *>(sp);
    const BBoxTree &t = mesh < 0 ? s.tree : s.meshes[mesh].tree;
    *n_nodes = (int64_t)t.nodes.size();
    *depth = t.depth();
    return 0;
}

// Per-triangle data after Compute_Mesh_Triangle: normal(3) distance p1 p2 p3
// dominant smooth -> out[8] floats/ints packed as doubles.
int orc_mesh_triangle(void *sp, int mesh, int tri, double *out) {
    const Scene &s = *static_cast<Scene *>(sp);
    const Mesh &m = s.meshes[mesh];
    const MeshTri &t = m.tris[tri];
    const float *nf = &m.normals[3 * (size_t)t.normal_ind];
    out[0] = nf[0];
    out[1] = nf[1];
    out[2] = nf[2];
    out[3] = t.distance;
    out[4] = t.p1;
    out[5] = t.p2;
    out[6] = t.p3;
    out[7] = t.dominant_axis;
    out[8] = t.smooth;
    out[9] = m.valid[tri];
    return 0;
}

int orc_mesh_normal(void *sp, int mesh, const int32_t *tri, const double *pts, size_t n,
                    double *out) {
    const Scene &s = *static_cast<Scene *>(sp);
    for (size_t i = 0; i < n; i++) {
        V3 r = mesh_normal(s.meshes[mesh], tri[i], V3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
        out[3 * i] = r.x;
        out[3 * i + 1] = r.y;
        out[3 * i + 2] = r.z;
    }
    return 0;
}

// Triangle::Normal / SmoothTriangle::Normal of stand-alone triangle `tri`
// at n points.
int orc_triangle_normal(void *sp, int tri, const double *pts, size_t n, double *out) {
    const Scene &s = *static_cast<Scene *>(sp);
    for (size_t i = 0; i < n; i++) {
        V3 r = tri_normal(s.tris[tri], V3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
        out[3 * i] = r.x;
        out[3 * i + 1] = r.y;
        out[3 * i + 2] = r.z;
    }
    return 0;
}

int orc_solve_polynomial(int order, const double *coeffs, int sturm, double epsilon,
                         double *roots) {
    return solve_polynomial(order, coeffs, roots, sturm != 0, epsilon);
}

} // extern "C"
