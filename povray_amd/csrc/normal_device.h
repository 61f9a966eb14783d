// Device-side surface normals (Object::Normal for the shapes on the path).
#ifndef PVX_NORMAL_DEVICE_H
#define PVX_NORMAL_DEVICE_H

#include "shapes_device.h"

namespace pvx {

struct NormalParams {
    SceneView sv;
    const uint32_t *obj_type;
    const uint32_t *obj_index;
};

__device__ inline D3 mtrans_normal(const double *inv, D3 n) { // MTransNormal via inverse
    return D3{n.x * inv[0] + n.y * inv[1] + n.z * inv[2], n.x * inv[4] + n.y * inv[5] + n.z * inv[6],
              n.x * inv[8] + n.y * inv[9] + n.z * inv[10]};
}
__device__ inline D3 normalize3(D3 a) {
    double l = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return l != 0.0 ? D3{a.x / l, a.y / l, a.z / l} : a;
}

// Object::Normal of the surface that was hit (raw, not yet flipped).
// kRev false: the scene has no lathe / sor (their case, which retraces
// the hit list, is left out, and with it a large per-lane stack).
template <bool kRev = true>
__device__ inline D3 leaf_normal(const NormalParams &p, uint32_t type, uint32_t index, double t, D3 o, D3 d,
                          D3 ip, bool child = false) {
    const SceneView &sv = p.sv;
    switch (type) {
    case PVX_SPHERE: { // Sphere::Normal (sphere.cpp), spherical mode
        const double *s = sv.sph + 4 * (size_t)index;
        return D3{(ip.x - s[0]) / s[3], (ip.y - s[1]) / s[3], (ip.z - s[2]) / s[3]};
    }
    case PVX_PLANE: {
        const PlaneRec &pl = sv.planes[index];
        return D3{pl.n[0], pl.n[1], pl.n[2]};
    }
    case PVX_BOX: { // Box::Normal (box.cpp:600-625): side of the hit depth
        const BoxRec &b = sv.boxes[index];
        double hs[2];
        int side[2] = {0, 0};
        int nh = box_hits(b, sv.xforms, o, d, hs, side);
        int sd = 0;
        for (int k = 0; k < nh; k++)
            if (hs[k] == t) sd = side[k];
        D3 n{0, 0, 0};
        if (sd == 1) n.x = -1; else if (sd == 2) n.x = 1;
        else if (sd == 3) n.y = -1; else if (sd == 4) n.y = 1;
        else if (sd == 5) n.z = -1; else if (sd == 6) n.z = 1;
        if (b.xf >= 0) n = normalize3(mtrans_normal(sv.xforms[b.xf].inv, n));
        return n;
    }
    case PVX_QUADRIC: { // Quadric::Normal (quadric.cpp): the gradient, (1,0,0) where it vanishes
        const QuadricRec &q = sv.quadrics[index];
        D3 n{2.0 * q.qa * ip.x + q.qb * ip.y + q.qc * ip.z + q.qd, q.qb * ip.x + 2.0 * q.qe * ip.y + q.qf * ip.z + q.qg,
             q.qc * ip.x + q.qf * ip.y + 2.0 * q.qh * ip.z + q.qi};
        const double len = sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
        return len == 0.0 ? D3{1.0, 0.0, 0.0} : D3{n.x / len, n.y / len, n.z / len};
    }
    case PVX_POLYGON: { // Polygon::Normal (polygon.cpp:308): S_Normal
        const PolygonRec &pg = sv.polygons[index];
        return D3{pg.n[0], pg.n[1], pg.n[2]};
    }
    case PVX_DISC: { // Disc::Normal (disc.cpp): the stored normal
        const DiscRec &c = sv.discs[index];
        return D3{c.n[0], c.n[1], c.n[2]};
    }
    case PVX_CONE: { // Cone::Normal (cone.cpp:440-480): the part of the hit at depth t
        const ConeRec &c = sv.cones[index];
        double hs[4];
        int part[4] = {0, 0, 0, 0};
        const int nh = cone_hits(c, sv.xforms, o, d, hs, part);
        int pt = 0;
        for (int k = 0; k < nh; k++)
            if (hs[k] == t) pt = part[k]; // the last one pushed, as the IStack scan from the top
        const Xform &x = sv.xforms[c.xf];
        D3 n = xf_point(x.inv, ip);
        if (pt == 3) n.z = (c.flags & kConeCylinder) ? 0.0 : -n.z;
        else if (pt == 1) n = D3{0.0, 0.0, -1.0};
        else if (pt == 2) n = D3{0.0, 0.0, 1.0};
        return normalize3(mtrans_normal(x.inv, n));
    }
    case PVX_TORUS: { // Torus::Normal (torus.cpp:458-490)
        const TorusRec &tr = sv.tori[index];
        const Xform &x = sv.xforms[tr.xf];
        D3 P = xf_point(x.inv, ip);
        double dist = sqrt(P.x * P.x + P.z * P.z);
        D3 M{0, 0, 0};
        if (dist > kEpsilon) M = D3{tr.R * P.x / dist, 0.0, tr.R * P.z / dist};
        return normalize3(mtrans_normal(x.inv, sub3(P, M)));
    }
    case PVX_LATHE:
    case PVX_SOR: { // Lathe::Normal (lathe.cpp:466-505), Sor::Normal (sor.cpp:540-585)
        if constexpr (!kRev) {
            return D3{0.0, 1.0, 0.0}; // not in such scenes
        } else {
            // the hit's segment and spline parameter: the hit list is computed
            // again and the last entry at the hit's depth taken (Find_Intersection
            // keeps the last of equal depths)
            int side = 0;
            double d1 = 0.0;
            rev_visit(sv, index, o, d, child, [&](double dep, int sd, double w) {
                if (dep == t) {
                    side = sd;
                    d1 = w;
                }
            });
            const RevRec &s = sv.revs[index];
            const Xform &x = sv.xforms[s.xf];
            const D3 P = xf_point(x.inv, ip);
            D3 n;
            if ((s.flags & 7u) == PVX_SOR_CUBIC) {
                const int part = side & 3;
                if (part == kSorCurve) {
                    const RevSeg &g = sv.rev_segs[s.seg0 + (side >> 2)];
                    const double k = 0.5 * (P.y * (3.0 * g.a[0] * P.y + 2.0 * g.b[0]) + g.c[0]);
                    n = D3{P.x, -k, P.z};
                } else {
                    n = D3{0.0, part == kSorBase ? -1.0 : 1.0, 0.0};
                }
            } else {
                const RevSeg &g = sv.rev_segs[s.seg0 + side];
                double r = P.x * P.x + P.z * P.z;
                if (r > kEpsilon) {
                    r = sqrt(r);
                    const double dx = d1 * (3.0 * g.a[0] * d1 + 2.0 * g.b[0]) + g.c[0];
                    const double dy = d1 * (3.0 * g.a[1] * d1 + 2.0 * g.b[1]) + g.c[1];
                    n = D3{dy * P.x, -dx * r, dy * P.z};
                } else {
                    n = D3{0.0, 1.0, 0.0};
                }
            }
            return normalize3(mtrans_normal(x.inv, n));
        }
    }
    case PVX_TRIANGLE: { // Triangle::Normal (triangle.cpp:641-644)
        const TriangleRec &t = sv.triangles[index];
        return D3{t.n[0], t.n[1], t.n[2]};
    }
    case PVX_SMOOTH_TRIANGLE: { // SmoothTriangle::Normal (triangle.cpp:712-738)
        const TriangleRec &t = sv.triangles[index];
        const D3 pi{ip.x - t.p1[0], ip.y - t.p1[1], ip.z - t.p1[2]};
        const double u = ((pi.x * t.perp[0]) + (pi.y * t.perp[1])) + (pi.z * t.perp[2]);
        if (u < kEpsilon) return D3{t.n1[0], t.n1[1], t.n1[2]};
        const int a = t.vaxis;
        const double v = (comp(pi, a) / u + t.p1[a] - t.p2[a]) / (t.p3[a] - t.p2[a]);
        D3 r;
        r.x = t.n1[0] + u * (t.n2[0] - t.n1[0] + v * (t.n3[0] - t.n2[0]));
        r.y = t.n1[1] + u * (t.n2[1] - t.n1[1] + v * (t.n3[1] - t.n2[1]));
        r.z = t.n1[2] + u * (t.n2[2] - t.n1[2] + v * (t.n3[2] - t.n2[2]));
        return normalize3(r);
    }
    default:
        return D3{0, 1, 0};
    }
}

template <bool kRev = true>
__device__ inline D3 hit_normal(const NormalParams &p, const pvx_hit &h, D3 o, D3 d, D3 ip) {
    const SceneView &sv = p.sv;
    uint32_t type = p.obj_type[h.obj], index = p.obj_index[h.obj];
    if (type == PVX_MESH) { // Mesh::Normal + Smooth_Mesh_Normal (mesh.cpp:342-380, 2366-2394)
        const ShadeTri &st = sv.shade_tris[sv.obj_tri_base[h.obj] + (uint32_t)h.prim];
        const int32_t xf = sv.obj_mesh_xf[h.obj];
        if (!(st.vaxis_smooth >> 4)) {
            const float *fn = sv.mesh_nrm + 3 * (size_t)st.fn;
            const D3 n{fn[0], fn[1], fn[2]};
            return xf < 0 ? n : normalize3(mtrans_normal(sv.xforms[xf].inv, n));
        }
        // a placed mesh interpolates at the hit point in its own space
        const D3 q = xf < 0 ? ip : xf_point(sv.xforms[xf].inv, ip);
        const float *n1 = sv.mesh_nrm + 3 * (size_t)st.n1, *n2 = sv.mesh_nrm + 3 * (size_t)st.n2,
                    *n3 = sv.mesh_nrm + 3 * (size_t)st.n3;
        const float *p1 = sv.mesh_vtx + 3 * (size_t)st.p1, *p2 = sv.mesh_vtx + 3 * (size_t)st.p2,
                    *p3 = sv.mesh_vtx + 3 * (size_t)st.p3;
        D3 pip1{q.x - (double)p1[0], q.y - (double)p1[1], q.z - (double)p1[2]};
        double u = pip1.x * (double)st.perp[0] + pip1.y * (double)st.perp[1] + pip1.z * (double)st.perp[2];
        D3 N1{n1[0], n1[1], n1[2]};
        D3 r;
        if (u < kEpsilon) {
            r = N1;
        } else {
            int axis = (int)(st.vaxis_smooth & 3u);
            double k1 = p1[axis], k2 = p2[axis], k3 = p3[axis];
            double v = (comp(pip1, axis) / u + k1 - k2) / (k3 - k2);
            D3 N2{n2[0], n2[1], n2[2]}, N3{n3[0], n3[1], n3[2]};
            D3 t1{N2.x - N1.x + v * (N3.x - N2.x), N2.y - N1.y + v * (N3.y - N2.y),
                  N2.z - N1.z + v * (N3.z - N2.z)};
            r = D3{N1.x + u * t1.x, N1.y + u * t1.y, N1.z + u * t1.z};
        }
        if (xf >= 0) r = mtrans_normal(sv.xforms[xf].inv, r);
        return normalize3(r);
    }
    if (type == PVX_CSG) {
        const CsgChild &c = sv.csg_children[(uint32_t)h.prim];
        return leaf_normal<kRev>(p, c.type, c.index, h.t, o, d, ip, true);
    }
    return leaf_normal<kRev>(p, type, index, h.t, o, d, ip);
}


} // namespace pvx

#endif
