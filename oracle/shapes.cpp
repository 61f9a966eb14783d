// FP64 CPU ORACLE — TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Sphere and mesh2 restatements (core/shape/sphere.cpp, core/shape/mesh.cpp,
// parser/parser.cpp:4512-4565).
#include <algorithm>

#include "oracle.hpp"

namespace orc {

static const double kSphereDepthTol = 1.0e-6; // sphere.cpp:63
static const double kMeshDepthTol = 1.0e-6;   // mesh.cpp:92

// Sphere::Intersect (sphere.cpp:214-241)
bool sphere_intersect(const Ray &r, V3 centre, double radius2, double *d1, double *d2) {
    V3 oc = centre - r.o;
    double oc2 = length_sqr(oc);
    double tca = dot(oc, r.d);
    if (oc2 >= radius2 && tca < kEpsilon) return false;
    double thc2 = radius2 - oc2 + tca * tca;
    if (thc2 > kEpsilon) {
        double hc = std::sqrt(thc2);
        *d1 = tca - hc;
        *d2 = tca + hc;
        return true;
    }
    return false;
}

// max3_coordinate (mesh.cpp:94)
static inline int max3_coordinate(double x, double y, double z) {
    return (x > y) ? ((x > z) ? 0 : 2) : ((y > z) ? 1 : 2);
}

// Mesh::compute_smooth_triangle (mesh.cpp:983-1008)
static void compute_smooth_triangle(MeshTri &t, V3 p1, V3 p2, V3 p3) {
    V3 p3mp2 = p3 - p2;
    t.v_axis = (uint8_t)max3_coordinate(std::fabs(p3mp2.x), std::fabs(p3mp2.y), std::fabs(p3mp2.z));
    V3 vt1 = normalized(p2 - p3);
    V3 vt2 = p1 - p3;
    double proj = dot(vt2, vt1);
    vt1 = vt1 * proj;
    V3 perp = normalized(vt1 - vt2);
    float pf[3] = {(float)perp.x, (float)perp.y, (float)perp.z};
    double uden = -dot(vt2, V3(pf[0], pf[1], pf[2]));
    float udf = (float)uden;
    for (int k = 0; k < 3; k++) t.perp[k] = pf[k] / udf;
}

// Mesh::Compute_Mesh_Triangle (mesh.cpp:838-945); returns false for a
// degenerate triangle (S_Normal is then the zero vector).
static bool compute_mesh_triangle(MeshTri &t, bool smooth, V3 p1, V3 p2, V3 p3, V3 &s_normal) {
    V3 v1 = p2 - p1, v2 = p3 - p1;
    s_normal = cross(v2, v1);
    double len = length(s_normal);
    if (len == 0.0) return false;
    s_normal = s_normal / len;
    t.distance = (float)dot(s_normal, p1);
    t.distance = (float)(t.distance * -1.0);
    double x = std::fabs(s_normal.x), y = std::fabs(s_normal.y), z = std::fabs(s_normal.z);
    t.dominant_axis = (uint8_t)max3_coordinate(x, y, z);
    bool swap = false;
    switch (t.dominant_axis) {
    case 0:
        swap = (p2.y - p3.y) * (p2.z - p1.z) < (p2.z - p3.z) * (p2.y - p1.y);
        break;
    case 1:
        swap = (p2.x - p3.x) * (p2.z - p1.z) < (p2.z - p3.z) * (p2.x - p1.x);
        break;
    default:
        swap = (p2.x - p3.x) * (p2.y - p1.y) < (p2.y - p3.y) * (p2.x - p1.x);
        break;
    }
    V3 q1 = p1, q2 = p2;
    if (swap) {
        std::swap(t.p1, t.p2);
        q1 = p2;
        q2 = p1;
        if (smooth) std::swap(t.n1, t.n2);
    }
    if (smooth) t.smooth = 1;
    compute_smooth_triangle(t, q1, q2, p3);
    return true;
}

// Parse_Mesh2 tail (parser.cpp:4494-4565) + Mesh::Compute_BBox (mesh.cpp) +
// Build_Mesh_BBox_Tree (mesh.cpp:1376-1409).
void mesh_prepare(Mesh &m, const pvx_mesh_desc &d) {
    uint32_t nv = d.n_vertices, nt = d.n_triangles;
    uint32_t nn = (d.normals && d.n_normals == nv) ? nv : 0;
    m.vertices.assign(d.vertices, d.vertices + 3 * (size_t)nv);
    m.normals.assign(3 * ((size_t)nn + nt), 0.0f);
    if (nn) std::copy(d.normals, d.normals + 3 * (size_t)nn, m.normals.begin());
    m.tris.assign(nt, MeshTri{});
    m.valid.assign(nt, 0);
    for (uint32_t i = 0; i < nt; i++) {
        MeshTri &t = m.tris[i];
        t.p1 = d.indices[3 * i];
        t.p2 = d.indices[3 * i + 1];
        t.p3 = d.indices[3 * i + 2];
        if (nn) {
            t.n1 = t.p1;
            t.n2 = t.p2;
            t.n3 = t.p3;
        }
        V3 p1 = m.vert(t.p1), p2 = m.vert(t.p2), p3 = m.vert(t.p3);
        t.smooth = 0;
        V3 n;
        bool ok;
        if (i < (nn ? nt : 0)) {
            auto nrm = [&](int k) {
                return V3(m.normals[3 * k], m.normals[3 * k + 1], m.normals[3 * k + 2]);
            };
            V3 dd1 = nrm(t.n1) - nrm(t.n2), dd2 = nrm(t.n1) - nrm(t.n3);
            double l1 = length_sqr(dd1), l2 = length_sqr(dd2);
            if (std::fabs(l1) > kEpsilon || std::fabs(l2) > kEpsilon) {
                ok = compute_mesh_triangle(t, true, p1, p2, p3, n);
                t.smooth = 1;
            } else {
                ok = compute_mesh_triangle(t, false, p1, p2, p3, n);
            }
        } else {
            ok = compute_mesh_triangle(t, false, p1, p2, p3, n);
        }
        m.valid[i] = ok ? 1 : 0;
        t.normal_ind = (int32_t)(i + nn);
        m.normals[3 * (nn + i)] = (float)n.x;
        m.normals[3 * (nn + i) + 1] = (float)n.y;
        m.normals[3 * (nn + i) + 2] = (float)n.z;
    }
    // Mesh::Compute_BBox
    V3 mins(kBoundHuge, kBoundHuge, kBoundHuge), maxs(-kBoundHuge, -kBoundHuge, -kBoundHuge);
    std::vector<FBox> boxes(nt);
    std::vector<int32_t> ids(nt);
    for (uint32_t i = 0; i < nt; i++) {
        const MeshTri &t = m.tris[i];
        V3 p[3] = {m.vert(t.p1), m.vert(t.p2), m.vert(t.p3)};
        V3 lo = p[0], hi = p[0];
        for (int a = 1; a < 3; a++)
            for (int k = 0; k < 3; k++) {
                lo[k] = std::min(lo[k], p[a][k]);
                hi[k] = std::max(hi[k], p[a][k]);
            }
        for (int k = 0; k < 3; k++) {
            mins[k] = std::min(mins[k], lo[k]);
            maxs[k] = std::max(maxs[k], hi[k]);
        }
        boxes[i] = make_bbox_from_min_max(lo, hi); // get_triangle_bbox (mesh.cpp:1335-1345)
        ids[i] = (int32_t)i;
    }
    m.bbox = make_bbox_from_min_max(mins, maxs);
    m.tree.build(boxes, ids, {}, {});
}

// Mesh::intersect_mesh_triangle (mesh.cpp:1040-1136), ray in mesh space.
bool mesh_triangle_intersect(const Mesh &m, const MeshTri &t, const Ray &r, double *depth) {
    const float *nf = &m.normals[3 * (size_t)t.normal_ind];
    V3 n(nf[0], nf[1], nf[2]);
    double ndd = dot(n, r.d);
    if (std::fabs(ndd) < kEpsilon) return false;
    double ndo = dot(n, r.o);
    *depth = -((double)t.distance + ndo) / ndd;
    if (*depth < kMeshDepthTol || *depth > kMaxDistance) return false;
    V3 p1 = m.vert(t.p1), p2 = m.vert(t.p2), p3 = m.vert(t.p3);
    int a, b; // projection axes
    switch (t.dominant_axis) {
    case 0: a = 1; b = 2; break;
    case 1: a = 0; b = 2; break;
    default: a = 0; b = 1; break;
    }
    double s = r.o[a] + *depth * r.d[a];
    double u = r.o[b] + *depth * r.d[b];
    if ((p2[a] - s) * (p2[b] - p1[b]) < (p2[b] - u) * (p2[a] - p1[a])) return false;
    if ((p3[a] - s) * (p3[b] - p2[b]) < (p3[b] - u) * (p3[a] - p2[a])) return false;
    if ((p1[a] - s) * (p1[b] - p3[b]) < (p1[b] - u) * (p1[a] - p3[a])) return false;
    return true;
}

// Mesh::Intersect (mesh.cpp:179-230) for an untransformed mesh (len = 1):
// brute force when use_tree is false, else intersect_bbox_tree
// (mesh.cpp:1452-1560) with its best-first early-out.
void mesh_all_intersections(const Mesh &m, const Transform *tr, const Ray &wr, std::vector<Isect> &stack,
                            bool use_tree) {
    double depth;
    Ray r = wr;
    double len = 1.0;
    if (tr) { // MInvTransRay, then the direction normalised
        r.o = mtrans_point(tr->inv, wr.o);
        r.d = mtrans_direction(tr->inv, wr.d);
        len = length(r.d);
        r.d = r.d / len;
    }
    if (!use_tree || m.tree.root < 0) {
        for (size_t i = 0; i < m.tris.size(); i++)
            if (mesh_triangle_intersect(m, m.tris[i], r, &depth))
                stack.push_back(Isect{depth / len, (int32_t)i, 0, -1});
        return;
    }
    RayInfo ri(r);
    PQueue q;
    double best = kBoundHuge, d;
    int32_t node;
    check_and_enqueue(q, m.tree.nodes[m.tree.root], m.tree.root, ri);
    while (!q.empty()) {
        q.remove_min(d, node);
        if (d > best) break;
        const BNode &n = m.tree.nodes[node];
        if (!n.kids.empty()) {
            for (int32_t k : n.kids) check_and_enqueue(q, m.tree.nodes[k], k, ri);
        } else if (mesh_triangle_intersect(m, m.tris[n.element], r, &depth)) {
            stack.push_back(Isect{depth / len, n.element, 0, -1});
            best = depth;
        }
    }
}

// Mesh::Normal (mesh.cpp:342-380): a placed mesh interpolates at the hit
// point in its own space and sends the normal out through MTransNormal.
static V3 mesh_normal_raw(const Mesh &m, int tri, V3 ip);

V3 mesh_normal_world(const Mesh &m, const Transform *tr, int tri, V3 ip) {
    if (!tr) return mesh_normal(m, tri, ip);
    V3 n = mesh_normal_raw(m, tri, mtrans_point(tr->inv, ip));
    return normalized(minv_trans_normal(tr->inv, n));
}

// Mesh::Normal without a placement: the face normal, or Smooth_Mesh_Normal
// normalised.
V3 mesh_normal(const Mesh &m, int tri, V3 ip) {
    V3 n = mesh_normal_raw(m, tri, ip);
    return m.tris[tri].smooth ? normalized(n) : n;
}

// Smooth_Mesh_Normal (mesh.cpp:410-439) or the face normal, in mesh space.
static V3 mesh_normal_raw(const Mesh &m, int tri, V3 ip) {
    const MeshTri &t = m.tris[tri];
    auto nrm = [&](int k) {
        return V3(m.normals[3 * k], m.normals[3 * k + 1], m.normals[3 * k + 2]);
    };
    if (!t.smooth) return nrm(t.normal_ind);
    V3 n1 = nrm(t.n1), n2 = nrm(t.n2), n3 = nrm(t.n3);
    V3 pip1 = ip - m.vert(t.p1);
    double u = dot(pip1, V3(t.perp[0], t.perp[1], t.perp[2]));
    V3 res;
    if (u < kEpsilon) {
        res = n1;
    } else {
        int axis = t.v_axis;
        double k1 = m.vertices[3 * t.p1 + axis];
        double k2 = m.vertices[3 * t.p2 + axis];
        double k3 = m.vertices[3 * t.p3 + axis];
        double v = (pip1[axis] / u + k1 - k2) / (k3 - k2);
        res = n1 + u * (n2 - n1 + v * (n3 - n2));
    }
    return res;
}

bool mesh_inside(const Mesh &, V3) {
    // Meshes without inside_vector are PATCH objects: Inside is never true
    // for them in CSG use (mesh.cpp Inside returns false without one).
    return false;
}

} // namespace orc
