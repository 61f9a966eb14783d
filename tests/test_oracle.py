"""Pins for the FP64 CPU oracle (the parity checker of the GPU engine).

The reference ships no golden vectors for this path and cannot be built here
(its build needs the autoconf-generated config.h), so every oracle function is
pinned by what the mathematics fixes, with independent NumPy code:
closed-form roots, polynomials built from known roots, point-membership of CSG
along the ray, and the reference's own guarantee that brute force (bounding
method 0) and the bounding-slab tree (method 1) find the same closest hit.
"""
import numpy as np
import pytest

from oracle.oracle import OracleScene, solve_polynomial
from povray_amd import _abi as A
from povray_amd import scenes


def _rays_random(rng, n, centre=(0, 0, 0), spread=8.0, dist=30.0):
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1)[:, None] * dist + np.asarray(centre)
    tgt = np.asarray(centre) + rng.uniform(-spread, spread, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.concatenate([o, d], 1)


# ---------------------------------------------------------------------------
# Solve_Polynomial
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("sturm", [True, False])
def test_quartic_known_roots(sturm):
    rng = np.random.default_rng(1)
    for _ in range(200):
        roots = np.sort(rng.uniform(0.01, 50.0, 4))
        if np.min(np.diff(roots)) < 1e-2:
            continue
        c = np.poly(roots)
        got = solve_polynomial(c, sturm=sturm, epsilon=0.0)
        assert got.shape == (4,), (roots, got)
        np.testing.assert_allclose(got, roots, rtol=1e-7, atol=1e-9)


def test_sturm_only_counts_positive_roots():
    # polysolve brackets roots in [0, MAX_DISTANCE]: negative roots are not reported
    c = np.poly([-3.0, -1.0, 2.0, 5.0])
    np.testing.assert_allclose(solve_polynomial(c, sturm=True, epsilon=0.0), [2.0, 5.0], rtol=1e-9)
    # Ferrari returns all real roots
    np.testing.assert_allclose(solve_polynomial(c, sturm=False, epsilon=0.0),
                               [-3.0, -1.0, 2.0, 5.0], rtol=1e-9)


def test_root_elimination_drops_to_cubic():
    # |c4/c3| < epsilon: the constant term is treated as zero (a root at 0 removed)
    c = np.poly([1e-9, 1.0, 2.0, 3.0])
    got = solve_polynomial(c, sturm=True, epsilon=1e-4)
    # exactly the roots of the cubic c[0..3] (c4 dropped), not of the quartic
    np.testing.assert_allclose(got, np.sort(np.roots(c[:4]).real), rtol=1e-11)
    assert got.shape == (3,)


def test_quadratic_and_cubic_orders():
    # leading coefficients below SMALL_ENOUGH reduce the order
    c = np.concatenate([[0.0, 0.0], np.poly([1.5, 4.0])])
    np.testing.assert_allclose(solve_polynomial(c, sturm=False, epsilon=0.0), [1.5, 4.0], rtol=1e-12)
    c = np.concatenate([[0.0], np.poly([0.5, 2.0, 7.0])])
    np.testing.assert_allclose(solve_polynomial(c, sturm=False, epsilon=0.0), [0.5, 2.0, 7.0],
                               rtol=1e-9)
    np.testing.assert_allclose(solve_polynomial(c, sturm=True, epsilon=0.0), [0.5, 2.0, 7.0],
                               rtol=1e-9)


def test_no_real_roots():
    c = np.poly([1 + 1j, 1 - 1j, 3 + 2j, 3 - 2j]).real
    assert solve_polynomial(c, sturm=True, epsilon=0.0).size == 0
    assert solve_polynomial(c, sturm=False, epsilon=0.0).size == 0


# ---------------------------------------------------------------------------
# spheres and planes: closed forms
# ---------------------------------------------------------------------------
def test_sphere_closest_hit_closed_form():
    rng = np.random.default_rng(2)
    s = scenes.sphere_scene(n=1)
    s.sphere_center = np.array([[0.3, -0.2, 0.5]])
    s.sphere_radius = np.array([1.7])
    o = OracleScene(s)
    rays = _rays_random(rng, 4000, centre=(0.3, -0.2, 0.5), spread=2.5, dist=10.0)
    hits = o.find_intersection(rays, method=0)
    oc = rays[:, :3] - s.sphere_center[0]
    b = np.sum(oc * rays[:, 3:], 1)
    cc = np.sum(oc * oc, 1) - 1.7 ** 2
    disc = b * b - cc
    hit = disc > 1e-9
    t = -b - np.sqrt(np.where(hit, disc, 0))
    assert np.array_equal(hits["obj"] >= 0, hit)
    np.testing.assert_allclose(hits["t"][hit], t[hit], rtol=1e-10)
    assert np.all(hits["t"][~hit] == 2.0e10)  # Intersection::Depth default BOUND_HUGE


def test_sphere_inside_and_origin_inside():
    s = scenes.sphere_scene(n=1)
    s.sphere_center = np.zeros((1, 3))
    s.sphere_radius = np.array([2.0])
    o = OracleScene(s)
    pts = np.array([[0, 0, 0], [1.9, 0, 0], [2.1, 0, 0], [0, 0, -3]], float)
    assert list(o.inside(pts, [0, 0, 0, 0])) == [True, True, False, False]
    # ray from inside: only the far root counts
    h = o.find_intersection(np.array([[0, 0, 0, 1, 0, 0.0]]), method=0)
    assert h["obj"][0] == 0 and abs(h["t"][0] - 2.0) < 1e-12


def test_plane_closed_form():
    s = scenes.mesh_scene(level=1)
    s.objects = s.objects[1:]  # plane only
    o = OracleScene(s)
    rng = np.random.default_rng(3)
    rays = _rays_random(rng, 2000, dist=5.0)
    h = o.find_intersection(rays, method=0)
    # y = -1.25 plane
    t = (-1.25 - rays[:, 1]) / rays[:, 4]
    ok = (t >= 1e-4) & (np.abs(rays[:, 4]) >= 1e-10)
    assert np.array_equal(h["obj"] >= 0, ok)
    np.testing.assert_allclose(h["t"][ok], t[ok], rtol=1e-12)


# ---------------------------------------------------------------------------
# mesh2 triangles
# ---------------------------------------------------------------------------
def _moller_trumbore(rays, v0, v1, v2):
    """Independent FP64 ray/triangle test: returns (t, u, v) per ray x tri."""
    e1, e2 = v1 - v0, v2 - v0
    d = rays[:, None, 3:]
    o = rays[:, None, :3]
    p = np.cross(d, e2[None])
    det = np.sum(e1[None] * p, -1)
    inv = 1.0 / np.where(np.abs(det) < 1e-300, 1e-300, det)
    tv = o - v0[None]
    u = np.sum(tv * p, -1) * inv
    q = np.cross(tv, e1[None])
    v = np.sum(d * q, -1) * inv
    t = np.sum(e2[None] * q, -1) * inv
    return t, u, v


def test_mesh_triangles_against_moller_trumbore():
    s = scenes.mesh_scene(level=2, width=64, height=36)
    s.objects = s.objects[:1]  # mesh only
    verts, faces, _ = s.meshes[0]
    o = OracleScene(s)
    rng = np.random.default_rng(4)
    rays = _rays_random(rng, 3000, spread=1.2, dist=4.0)
    h = o.find_intersection(rays, method=0)
    v = verts.astype(np.float64)
    t, u, w = _moller_trumbore(rays, v[faces[:, 0]], v[faces[:, 1]], v[faces[:, 2]])
    inside = (u > 1e-7) & (w > 1e-7) & (u + w < 1 - 1e-7) & (t > 1e-4)
    amb = ((u > -1e-7) & (w > -1e-7) & (u + w < 1 + 1e-7) & (t > 1e-4)) & ~inside
    tt = np.where(inside, t, np.inf)
    best = tt.min(1)
    bi = tt.argmin(1)
    # skip rays where an edge-ambiguous triangle could be closer
    amb_t = np.where(amb, t, np.inf).min(1)
    clear = amb_t > best * (1 + 1e-9)
    hit = np.isfinite(best) & clear
    assert hit.sum() > 500
    assert np.array_equal(h["prim"][hit], bi[hit])
    np.testing.assert_allclose(h["t"][hit], best[hit], rtol=1e-6)
    miss = ~np.isfinite(best) & ~np.isfinite(amb_t)
    assert np.all(h["obj"][miss] == -1)


def test_mesh_triangle_preparation_invariants():
    s = scenes.mesh_scene(level=2)
    o = OracleScene(s)
    verts, faces, _ = s.meshes[0]
    for tri in range(0, faces.shape[0], 7):
        info = o.mesh_triangle(0, tri)
        n = info[:3]
        dist = info[3]
        p = verts[[int(info[4]), int(info[5]), int(info[6])]].astype(np.float64)
        assert abs(np.linalg.norm(n) - 1) < 1e-6
        # plane N.P + D = 0 through the (FP32) vertices
        assert np.all(np.abs(p @ n + dist) < 1e-5)
        axis = int(info[7])
        assert axis == int(np.argmax(np.abs(n)))
        # same vertex set as the face, P3 never swapped
        assert sorted(map(int, info[4:7])) == sorted(faces[tri].tolist())
        assert int(info[6]) == faces[tri][2]
        # after the swap, the projected triangle winds so that the edge tests
        # (P2-s)x(P2-P1) >= 0 hold for its centroid
        a, b = [(1, 2), (0, 2), (0, 1)][axis]
        c = p.mean(0)
        for (pa, pb) in ((p[1], p[0]), (p[2], p[1]), (p[0], p[2])):
            assert (pa[a] - c[a]) * (pa[b] - pb[b]) >= (pa[b] - c[b]) * (pa[a] - pb[a])


# ---------------------------------------------------------------------------
# bounding-slab tree == brute force (trace.cpp:230-344 guarantee)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["spheres", "mesh", "csg"])
def test_tree_matches_brute_force(kind):
    if kind == "spheres":
        s = scenes.sphere_scene(n=400, width=120, height=90)
    elif kind == "mesh":
        s = scenes.mesh_scene(level=4, width=120, height=68)
    else:
        s = scenes.csg_scene(n_objects=240, width=120, height=68)
    o = OracleScene(s)
    rays = scenes.camera_rays_np(s.camera)
    h0 = o.find_intersection(rays, method=0)
    h1 = o.find_intersection(rays, method=1)
    assert (h0["obj"] >= 0).mean() > 0.05
    assert np.array_equal(h0["obj"], h1["obj"])
    assert np.array_equal(h0["prim"], h1["prim"])
    assert np.array_equal(h0["t"], h1["t"])


# ---------------------------------------------------------------------------
# boxes, tori and CSG: independent point membership along the ray
# ---------------------------------------------------------------------------
def _np_inside(s, type_, index, flags, p):
    """Independent NumPy membership (sphere / box / torus / CSG tree)."""
    inv = bool(flags & A.PVX_FLAG_INVERTED)
    if type_ == A.PVX_SPHERE:
        r = np.sum((p - s.sphere_center[index]) ** 2, -1) < s.sphere_radius[index] ** 2
    elif type_ == A.PVX_BOX:
        t = s.transforms[s.box_transform[index]]
        q = np.c_[p, np.ones(len(p))] @ t[16:].reshape(4, 4)
        r = np.all((q[:, :3] >= s.box_corner1[index]) & (q[:, :3] <= s.box_corner2[index]), -1)
    elif type_ == A.PVX_TORUS:
        t = s.transforms[s.torus_transform[index]]
        q = np.c_[p, np.ones(len(p))] @ t[16:].reshape(4, 4)
        R, rr = s.torus_major[index], s.torus_minor[index]
        rad = np.sqrt(q[:, 0] ** 2 + q[:, 2] ** 2)
        r = q[:, 1] ** 2 + (rad - R) ** 2 <= rr * rr
    else:
        op, first, cnt, _ = s.csg[index]
        ins = [_np_inside(s, *s.csg_children[first + k][:3], p) for k in range(cnt)]
        r = np.all(ins, 0) if op == A.PVX_CSG_INTERSECTION else np.any(ins, 0)
        return r
    return ~r if inv else r


def test_csg_hits_are_first_boundary_crossings():
    s = scenes.csg_scene(n_objects=80, width=200, height=112)
    o = OracleScene(s)
    rays = scenes.camera_rays_np(s.camera)
    h = o.find_intersection(rays, method=0)
    idx = np.nonzero(h["obj"] >= 0)[0]
    assert idx.size > 100
    rng = np.random.default_rng(6)
    idx = rng.choice(idx, min(300, idx.size), replace=False)
    for i in idx:
        root = s.objects[h["obj"][i]][1]
        t = h["t"][i]
        o_, d_ = rays[i, :3], rays[i, 3:]
        eps = 1e-6 * max(1.0, t)
        before = _np_inside(s, A.PVX_CSG, root, 0, (o_ + d_ * (t - eps))[None])[0]
        after = _np_inside(s, A.PVX_CSG, root, 0, (o_ + d_ * (t + eps))[None])[0]
        assert before != after, f"ray {i}: hit at t={t} is not a boundary crossing"
        # no earlier crossing of this object's boundary along the ray
        ts = np.linspace(1e-3, t - 10 * eps, 400)
        mem = _np_inside(s, A.PVX_CSG, root, 0, o_[None] + ts[:, None] * d_[None])
        assert np.all(mem == mem[0]) and mem[0] == before


def test_csg_misses_have_constant_membership():
    s = scenes.csg_scene(n_objects=40, width=96, height=54)
    o = OracleScene(s)
    rays = scenes.camera_rays_np(s.camera)
    h = o.find_intersection(rays, method=0)
    miss = np.nonzero(h["obj"] < 0)[0][::37]
    ts = np.linspace(1e-3, 120.0, 1500)
    for i in miss:
        pts = rays[i, :3][None] + ts[:, None] * rays[i, 3:][None]
        for obj in range(s.n_objects):
            mem = _np_inside(s, A.PVX_CSG, s.objects[obj][1], 0, pts)
            assert np.all(mem == mem[0]), f"ray {i} crosses object {obj} but the oracle misses"


def test_box_and_torus_inside_match_numpy():
    s = scenes.csg_scene(n_objects=80)
    o = OracleScene(s)
    rng = np.random.default_rng(8)
    pts = rng.uniform(-8, 8, (4000, 3)) + s.sphere_center.mean(0)
    for obj in range(s.n_objects):
        root = s.objects[obj][1]
        want = _np_inside(s, A.PVX_CSG, root, 0, pts)
        got = o.inside(pts, np.full(len(pts), obj))
        assert np.array_equal(got, want)
