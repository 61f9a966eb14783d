"""Cones and cylinders (core/shape/cone.cpp), discs (core/shape/disc.cpp)
and quadrics (core/shape/quadric.cpp).

CPU: the oracle's cone (the shape's own placement from its two points and
radii, the canonical-space intersection, Inside) is pinned against closed
forms computed here directly in world space: the lateral surface
|w - (w.a)a| = r0 + k (w.a) with w = p - base, a the unit axis, plus the two
end discs, and the point-membership test of the same solid.  These never use
the shape's transform, so a wrong placement, swap of ends, radius, cap or
open flag fails them.
GPU: closest hits, intersection lists, Inside and shaded frames against the
oracle."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from povray_amd import _abi as A
from povray_amd import scenes


def _single(apex, base, r1, r2, flags, transform=None):
    s = scenes.cone_scene(n=1, n_csg=0, n_discs=0, n_quadrics=0, n_polygons=0, width=8, height=8)
    s.cone_apex = np.array([apex], float)
    s.cone_base = np.array([base], float)
    s.cone_radius = np.array([[r1, r2]], float)
    s.cone_flags = np.array([flags], np.uint8)
    if transform is None:
        s.cone_transform = np.array([-1], np.int32)
        s.transforms = None
    else:
        s.cone_transform = np.array([0], np.int32)
        s.transforms = transform.reshape(1, 32)
    s.objects = s.objects[:1]
    s.plane_normal = s.plane_distance = None
    s.materials = {k: v[:1] for k, v in s.materials.items()}
    return s


def _closed_form(o, d, apex, base, r1, r2, closed):
    """World-space hits (sorted) of rays with the solid cone frustum."""
    apex, base = np.asarray(apex, float), np.asarray(base, float)
    L = np.linalg.norm(apex - base)
    a = (apex - base) / L
    k = (r1 - r2) / L
    out = [[] for _ in range(o.shape[0])]
    w0 = o - base
    wa, da = w0 @ a, d @ a
    # |w|^2 - (w.a)^2 - (r2 + k w.a)^2 = 0, w = w0 + t d
    A2 = (d * d).sum(1) - da * da - k * k * da * da
    B2 = 2 * ((w0 * d).sum(1) - wa * da - k * da * (r2 + k * wa))
    C2 = (w0 * w0).sum(1) - wa * wa - (r2 + k * wa) ** 2
    disc = B2 * B2 - 4 * A2 * C2
    for i in range(o.shape[0]):
        ts = []
        if abs(A2[i]) > 1e-12 and disc[i] >= 0:
            q = np.sqrt(disc[i])
            ts = [(-B2[i] - q) / (2 * A2[i]), (-B2[i] + q) / (2 * A2[i])]
        for t in ts:
            s_ = wa[i] + t * da[i]
            rr = r2 + k * s_
            if t > 1e-7 and 0 <= s_ <= L and rr >= 0:
                out[i].append(t)
        if closed and abs(da[i]) > 1e-12:
            for centre, rad in ((apex, r1), (base, r2)):
                t = ((centre - o[i]) @ a) / da[i]
                p = o[i] + t * d[i]
                if t > 1e-7 and np.linalg.norm(p - centre) <= rad:
                    out[i].append(t)
    return [sorted(x) for x in out]


def _rays_at(apex, base, n, seed):
    rng = np.random.default_rng(seed)
    mid = 0.5 * (np.asarray(apex) + np.asarray(base))
    o = mid + rng.normal(size=(n, 3)) * 4.0
    tgt = mid + rng.normal(size=(n, 3)) * 0.7
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.hstack([o, d])


CASES = [
    # apex, base, r1, r2, flags
    ((0.3, 1.5, -0.2), (0.1, -1.0, 0.4), 0.9, 0.3, 0),                 # cone, apex end bigger
    ((0.3, 1.5, -0.2), (0.1, -1.0, 0.4), 0.3, 0.9, 0),                 # smaller first: ends swapped
    ((-1.0, 0.2, 0.5), (1.2, 0.4, -0.3), 0.0, 0.7, 0),                 # pointed
    ((0.0, 0.0, 2.0), (0.0, 0.0, -1.0), 0.6, 0.6, 0),                  # equal radii: a cylinder
    ((0.5, -0.5, 1.0), (-0.5, 1.0, -1.0), 0.5, 0.0, A.PVX_CONE_CYLINDER),  # cylinder statement
    ((0.5, -0.5, 1.0), (-0.5, 1.0, -1.0), 0.5, 0.0, A.PVX_CONE_CYLINDER | A.PVX_CONE_OPEN),
    ((0.3, 1.5, -0.2), (0.1, -1.0, 0.4), 0.9, 0.3, A.PVX_CONE_OPEN),   # open cone
    ((0.0, 3.0, 0.0), (0.0, -3.0, 0.0), 1.0, 0.2, 0),                  # axis along y (rotation edge case)
    ((0.0, 0.0, -3.0), (0.0, 0.0, 3.0), 1.0, 0.5, 0),                  # axis along -z
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_oracle_cone_closed_form(case):
    apex, base, r1, r2, fl = CASES[case]
    if fl & A.PVX_CONE_CYLINDER:
        r2 = r1
    s = _single(apex, base, r1, r2 if not (fl & A.PVX_CONE_CYLINDER) else 0.0, fl)
    rays = _rays_at(apex, base, 3000, case)
    d, _, c = OracleScene(s).all_intersections(rays, np.zeros(rays.shape[0], np.uint32), 8)
    ref = _closed_form(rays[:, :3], rays[:, 3:], apex, base, r1, r2, not (fl & A.PVX_CONE_OPEN))
    n_hit = 0
    for i in range(rays.shape[0]):
        got = list(d[i, :c[i]])
        exp = ref[i]
        if len(got) != len(exp):
            # only grazing rays (a hit at a rim or a tangent) may differ in count
            near_rim = any(abs(x - y) < 1e-6 * max(1.0, abs(x)) for x in exp for y in exp if x is not y)
            assert near_rim or len(exp) == 0 and len(got) == 0, f"ray {i}: {got} vs {exp}"
            continue
        n_hit += len(got) > 0
        assert np.allclose(got, exp, rtol=1e-9, atol=1e-9), f"ray {i}: {got} vs {exp}"
    assert n_hit > 300


def test_oracle_cone_inside_and_transform():
    apex, base, r1, r2 = (0.3, 1.5, -0.2), (0.1, -1.0, 0.4), 0.9, 0.3
    rng = np.random.default_rng(3)
    pts = rng.uniform(-2.0, 2.0, (20000, 3))
    for fl in (0, A.PVX_CONE_OPEN):
        s = _single(apex, base, r1, r2, fl)
        got = OracleScene(s).inside(pts, np.zeros(pts.shape[0], np.uint32))
        a = np.subtract(apex, base)
        L = np.linalg.norm(a)
        a = a / L
        w = pts - base
        sa = w @ a
        radial = np.linalg.norm(w - np.outer(sa, a), axis=1)
        rr = r2 + (r1 - r2) / L * sa
        margin = np.minimum.reduce([rr - radial, sa, L - sa])
        exp = margin > 0
        far = np.abs(margin) > 1e-6
        assert got[far].tolist() == exp[far].tolist()
        assert exp.sum() > 500
    # a rotation composed after the shape's own placement: hits of the
    # rotated cone are those of the cone whose points are rotated
    rot = scenes.rotation_transform((30.0, 50.0, 70.0))
    xf = scenes._xf_row(rot)
    m = np.asarray(xf[:16]).reshape(4, 4)
    ra = np.asarray(apex) @ m[:3, :3] + m[3, :3]
    rb = np.asarray(base) @ m[:3, :3] + m[3, :3]
    s1 = _single(apex, base, r1, r2, 0, transform=np.asarray(xf))
    s2 = _single(tuple(ra), tuple(rb), r1, r2, 0)
    rays = _rays_at(tuple(ra), tuple(rb), 2000, 9)
    h1 = OracleScene(s1).find_intersection(rays, method=0)
    h2 = OracleScene(s2).find_intersection(rays, method=0)
    assert np.array_equal(h1["obj"], h2["obj"])
    ok = h1["obj"] >= 0
    assert ok.sum() > 300
    assert np.allclose(h1["t"][ok], h2["t"][ok], rtol=1e-9)


def test_cone_scene_generator_and_tree():
    s = scenes.cone_scene()
    assert s.n_objects == 300 + 40 + 60 + 8 + 2 + 10 + 1
    o = OracleScene(s)
    rays = scenes.camera_rays_np(s.camera)
    h1, h0 = o.find_intersection(rays, method=1), o.find_intersection(rays, method=0)
    assert np.array_equal(h1["obj"], h0["obj"]) and np.array_equal(h1["t"], h0["t"])
    assert np.mean((h1["obj"] >= 0) & (h1["obj"] < 300)) > 0.1
    assert np.mean((h1["obj"] >= 340) & (h1["obj"] < 400)) > 0.01  # discs are seen
    assert np.mean((h1["obj"] >= 400) & (h1["obj"] < 410)) > 0.003  # quadrics are seen
    assert np.mean((h1["obj"] >= 410) & (h1["obj"] < 420)) > 0.002  # polygons are seen


def _single_disc(center, normal, radius, hole, transform=None):
    s = scenes.cone_scene(n=1, n_csg=0, n_discs=1, n_quadrics=0, n_polygons=0, width=8, height=8)
    s.disc_center = np.array([center], float)
    s.disc_normal = np.array([normal], float)
    s.disc_radius = np.array([[radius, hole]], float)
    if transform is None:
        s.disc_transform = np.array([-1], np.int32)
        s.transforms = None
        s.cone_transform = np.array([-1], np.int32)
    else:
        s.disc_transform = np.array([0], np.int32)
        s.cone_transform = np.array([-1], np.int32)
        s.transforms = np.asarray(transform).reshape(1, 32)
    s.objects = s.objects[1:2]  # the disc
    s.plane_normal = s.plane_distance = None
    s.materials = {k: v[1:2] for k, v in s.materials.items()}
    return s


DISCS = [((0.2, -0.3, 0.5), (0.3, 0.8, -0.2), 1.1, 0.0),
         ((0.2, -0.3, 0.5), (0.3, 0.8, -0.2), 1.1, 0.6),   # with a hole
         ((0.0, 0.0, 0.0), (0.0, 0.0, -2.0), 0.9, 0.0),    # normal along -z
         ((1.0, 2.0, 3.0), (1e-12, 0.0, 1.0), 0.7, 0.2)]   # within EPSILON of +z: snapped


@pytest.mark.parametrize("case", range(len(DISCS)))
def test_oracle_disc_closed_form(case):
    c, nrm, rad, hole = DISCS[case]
    s = _single_disc(c, nrm, rad, hole)
    rays = _rays_at(c, np.add(c, 0.1), 3000, 20 + case)
    d, _, cnt = OracleScene(s).all_intersections(rays, np.zeros(rays.shape[0], np.uint32), 4)
    n = np.asarray(nrm) / np.linalg.norm(nrm)
    o, dd = rays[:, :3], rays[:, 3:]
    t = ((np.asarray(c) - o) @ n) / (dd @ n)
    p = o + t[:, None] * dd
    r2 = ((p - np.asarray(c)) ** 2).sum(1)
    exp = (t > 1e-6) & (r2 >= hole * hole) & (r2 <= rad * rad)
    edge = (np.abs(r2 - rad * rad) < 1e-7) | (np.abs(r2 - hole * hole) < 1e-7)
    got = cnt == 1
    assert (got == exp)[~edge].all()
    assert exp.sum() > 200
    assert np.allclose(d[got, 0], t[got], rtol=1e-9)
    # Inside: the half-space behind the normal
    pts = np.asarray(c) + np.random.default_rng(case).uniform(-2, 2, (5000, 3))
    side = (pts - np.asarray(c)) @ n
    ins = OracleScene(s).inside(pts, np.zeros(pts.shape[0], np.uint32))
    far = np.abs(side) > 1e-9
    assert (ins[far] == (side[far] < 0)).all()


def test_oracle_disc_transform():
    rot = scenes.rotation_transform((20.0, 40.0, 60.0))
    xf = np.asarray(scenes._xf_row(rot))
    m = xf[:16].reshape(4, 4)
    c, nrm = np.array([0.3, 0.1, -0.4]), np.array([0.2, 1.0, 0.3])
    rc = c @ m[:3, :3] + m[3, :3]
    rn = nrm @ m[:3, :3]  # a rotation carries normals like points' directions
    s1 = _single_disc(tuple(c), tuple(nrm), 1.0, 0.3, transform=xf)
    s2 = _single_disc(tuple(rc), tuple(rn), 1.0, 0.3)
    rays = _rays_at(tuple(rc), tuple(rc + 0.1), 2000, 31)
    h1 = OracleScene(s1).find_intersection(rays, method=0)
    h2 = OracleScene(s2).find_intersection(rays, method=0)
    assert np.array_equal(h1["obj"], h2["obj"])
    ok = h1["obj"] >= 0
    assert ok.sum() > 200
    assert np.allclose(h1["t"][ok], h2["t"][ok], rtol=1e-9)


def test_oracle_quadric_ellipsoid_closed_form():
    """An ellipsoid quadric moved by a rotation + translation: the oracle's
    transformed coefficients give the hits of the ray taken into the
    ellipsoid's own frame and solved there."""
    ax = np.array([0.7, 1.3, 0.9])
    rot = scenes.rotation_transform((25.0, 70.0, 140.0))
    t = scenes.compose(rot, scenes.translation_transform(np.array([0.4, -0.2, 0.3])))
    xf = np.asarray(scenes._xf_row(t))
    s = scenes.cone_scene(n=1, n_csg=0, n_discs=0, n_quadrics=0, n_polygons=0, width=8, height=8)
    s.quadric_coef = np.array([[1 / ax[0] ** 2, 1 / ax[1] ** 2, 1 / ax[2] ** 2, 0, 0, 0, 0, 0, 0, -1.0]])
    s.quadric_transform = np.array([0], np.int32)
    s.transforms = xf.reshape(1, 32)
    s.cone_transform = np.array([-1], np.int32)
    s.objects = np.array([[A.PVX_QUADRIC, 0, 0, 0]], np.uint32)
    s.plane_normal = s.plane_distance = None
    s.materials = {k: v[:1] for k, v in s.materials.items()}
    inv = xf[16:].reshape(4, 4)
    rays = _rays_at((0.4, -0.2, 0.3), (0.4, -0.2, 0.3), 3000, 41)
    d, _, cnt = OracleScene(s).all_intersections(rays, np.zeros(rays.shape[0], np.uint32), 4)
    o = rays[:, :3] @ inv[:3, :3] + inv[3, :3]
    dd = rays[:, 3:] @ inv[:3, :3]
    A2 = ((dd / ax) ** 2).sum(1)
    B2 = 2 * ((o * dd) / ax ** 2).sum(1)
    C2 = ((o / ax) ** 2).sum(1) - 1
    disc = B2 * B2 - 4 * A2 * C2
    hit = disc > 1e-9
    n_hit = 0
    for i in np.nonzero(hit)[0]:
        q = np.sqrt(disc[i])
        ts = sorted(t for t in ((-B2[i] - q) / (2 * A2[i]), (-B2[i] + q) / (2 * A2[i])) if t > 1e-6)
        assert cnt[i] == len(ts) and np.allclose(d[i, :cnt[i]], ts, rtol=1e-9), (i, d[i, :cnt[i]], ts)
        n_hit += 1
    assert (cnt[disc < -1e-9] == 0).all() and n_hit > 300
    pts = np.random.default_rng(5).uniform(-2, 2, (5000, 3))
    lp = pts @ inv[:3, :3] + inv[3, :3]
    val = ((lp / ax) ** 2).sum(1) - 1
    ins = OracleScene(s).inside(pts, np.zeros(pts.shape[0], np.uint32))
    far = np.abs(val) > 1e-9
    assert (ins[far] == (val[far] < 0)).all()


def test_oracle_polygon_closed_form():
    """A square with a square hole, given as two closed sub-polygons (the
    outer one written without its closing point, which the reader adds):
    hits are where the ray meets the plane inside the outer square and
    outside the hole."""
    s = scenes.cone_scene(n=1, n_csg=0, n_discs=0, n_quadrics=0, n_polygons=0, width=8, height=8)
    c = np.array([0.3, -0.2, 0.5])
    u = np.array([0.6, 0.8, 0.0])
    v = np.array([0.0, 0.0, 1.0])
    sq = lambda h: [c + h * (a * u + b * v) for a, b in ((-1, -1), (1, -1), (1, 1), (-1, 1))]
    hole = sq(0.4)
    pts = sq(1.0) + [sq(1.0)[0]] + hole + [hole[0]]
    s.polygon_first = np.array([0, len(pts)], np.uint32)
    s.polygon_points = np.array(pts)
    s.polygon_transform = None
    s.objects = np.array([[A.PVX_POLYGON, 0, 0, 0]], np.uint32)
    s.plane_normal = s.plane_distance = None
    s.cone_transform = np.array([-1], np.int32)
    s.transforms = None
    s.materials = {k: v_[:1] for k, v_ in s.materials.items()}
    rays = _rays_at(tuple(c), tuple(c), 4000, 51)
    d, _, cnt = OracleScene(s).all_intersections(rays, np.zeros(rays.shape[0], np.uint32), 2)
    n = np.cross(u, v)
    o, dd = rays[:, :3], rays[:, 3:]
    t = ((c - o) @ n) / (dd @ n)
    p = o + t[:, None] * dd - c
    a, b = np.abs(p @ u), np.abs(p @ v)
    exp = (t > 1e-7) & (np.maximum(a, b) <= 1.0) & (np.maximum(a, b) >= 0.4)
    edge = (np.abs(np.maximum(a, b) - 1.0) < 1e-9) | (np.abs(np.maximum(a, b) - 0.4) < 1e-9)
    got = cnt == 1
    assert (got == exp)[~edge].all()
    assert exp.sum() > 300 and (~exp).sum() > 300
    assert np.allclose(d[got, 0], t[got], rtol=1e-9)


@pytest.mark.gpu
def test_gpu_cones_match_oracle(gpu):
    s = scenes.cone_scene()
    gs = gpu.GpuScene(s)
    o = OracleScene(s)
    rays = scenes.camera_rays_np(s.camera)
    for kind, md in ((A.PVX_RAY_IMAGE, None),
                     (A.PVX_RAY_SHADOW, np.random.default_rng(5).uniform(2.0, 30.0, rays.shape[0]))):
        g, c = gs.find_intersection(rays, md, kind), o.find_intersection(rays, md, kind, method=1)
        bad = (g["obj"] != c["obj"]) | (g["prim"] != c["prim"])
        assert not bad.any(), f"kind {kind}: {int(bad.sum())} id mismatches"
        assert np.array_equal(g["t"], c["t"])
    # intersection lists and Inside, per object
    rng = np.random.default_rng(6)
    obj = rng.integers(0, s.n_objects - 1, 4000).astype(np.uint32)
    rr = rays[rng.integers(0, rays.shape[0], 4000)]
    for a, b in zip(gs.all_intersections(rr, obj, 8), o.all_intersections(rr, obj, 8)):
        assert np.array_equal(a, b)
    pts = rng.uniform(-7.0, 7.0, (20000, 3))
    po = rng.integers(0, s.n_objects - 1, 20000).astype(np.uint32)
    assert np.array_equal(gs.inside(pts, po), o.inside(pts, po))
    # the GPU-built tree gives the same hits
    gb = gpu.GpuScene(s, builder=A.PVX_BUILDER_GPU).find_intersection(rays)
    assert np.array_equal(gb["obj"], gs.find_intersection(rays)["obj"])


@pytest.mark.gpu
def test_gpu_render_cones_matches_oracle(gpu):
    s = scenes.cone_scene(width=320, height=240)
    gs = gpu.GpuScene(s)
    gs.set_shading()
    img, st = gs.render()
    ref, (n_closest, n_shadow) = OracleScene(s).render(s)
    diff = np.abs(img - ref).max(axis=2)
    ok = np.mean(diff <= 1.0 / 255.0)
    assert ok >= 0.999, f"only {ok * 100:.3f}% of pixels within 1/255"
    assert abs(st["closest_rays"] - n_closest) <= 1e-3 * n_closest
