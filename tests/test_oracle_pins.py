"""Pins of the oracle's shading, lighting, radiosity and solver restatements
against what the mathematics fixes, computed here by independent routes
(analytic ray-sphere / ray-plane geometry, Snell's law in angle form, the
textbook radical inverse, the differential form factor of a rectangle,
segment-box occlusion), so that a dropped term, a wrong sign or index or a
transposed operand in the oracle fails one of them.

The reference ships no fixtures for this path; these checks stand in for
them.  Everything here runs on the CPU.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from oracle.oracle import OracleScene
from povray_amd import _abi as A
from povray_amd import scenes


def _scene(objects, materials, lights=(), settings=None, **shapes):
    objs = np.zeros((len(objects), 4), np.uint32)
    for i, (t, ix) in enumerate(objects):
        objs[i, 0], objs[i, 1] = t, ix
    n = len(objects)
    mats = dict(pigment=np.ones((n, 3)), ambient=np.zeros(n), diffuse=np.zeros(n), phong=np.zeros(n),
                phong_size=np.full(n, 40.0), reflection=np.zeros(n), transmit=np.zeros(n), ior=np.ones(n))
    for k, v in materials.items():
        mats[k] = np.asarray(v, np.float64)
    st = dict(max_trace_level=10, adc_bailout=1.0 / 255.0, ambient_light=(1.0, 1.0, 1.0),
              background=(0.0, 0.0, 0.0))
    st.update(settings or {})
    return scenes.Scene(name="pin", objects=objs, lights=list(lights), materials=mats, settings=st,
                        camera=scenes.look_at_camera((0.0, 0.0, -5.0), (0.0, 0.0, 0.0), 8, 8), **shapes)


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _rays(o, d):
    o = np.broadcast_to(np.asarray(o, np.float64), np.shape(d))
    return np.concatenate([o, _unit(d)], axis=-1)


def _sphere_hit(o, d, c, r):
    """Nearest positive root of |o + t d - c| = r (d unit), or inf."""
    oc = o - c
    b = np.sum(oc * d, axis=-1)
    disc = b * b - (np.sum(oc * oc, axis=-1) - r * r)
    t = np.where(disc > 0, -b - np.sqrt(np.maximum(disc, 0)), np.inf)
    t2 = np.where(disc > 0, -b + np.sqrt(np.maximum(disc, 0)), np.inf)
    return np.where(t > 1e-6, t, np.where(t2 > 1e-6, t2, np.inf)), disc


# ---------------------------------------------------------------------------
# ComputeDiffuseColour / ComputePhongColour (trace.cpp:2441-2558)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("brilliance", [1.0, 2.5])
def test_lambert_phong_sphere_closed_form(brilliance):
    """One sphere, one point light, no other object: each pixel is
    ambient*pigment + diffuse*pigment*cos(theta)^brilliance +
    phong*(R.L)^size, with the hit point, normal and mirror direction R of
    the eye ray solved here analytically."""
    pig = np.array([0.8, 0.5, 0.2])
    L = np.array([3.0, 4.0, -6.0])
    s = _scene([(A.PVX_SPHERE, 0)], dict(pigment=[pig], ambient=[0.1], diffuse=[0.6], phong=[0.5],
                                          phong_size=[7.0], brilliance=[brilliance]),
               lights=[dict(kind="point", position=L, colour=np.ones(3))],
               sphere_center=np.zeros((1, 3)), sphere_radius=np.array([1.0]))
    rng = np.random.default_rng(5)
    o = np.array([0.0, 0.0, -5.0])
    rad, ang = 0.2 * np.sqrt(rng.uniform(0, 0.97, 400)), rng.uniform(0, 2 * np.pi, 400)
    d = _unit(np.column_stack([rad * np.cos(ang), rad * np.sin(ang), np.ones(400)]))
    got = OracleScene(s).trace_rays(s, _rays(o, d))
    t, disc = _sphere_hit(o, d, np.zeros(3), 1.0)
    assert np.all(np.isfinite(t))
    p = o + t[:, None] * d
    n = p  # unit sphere at the origin
    ldir = _unit(L - p)
    cos = np.sum(n * ldir, axis=1)
    lit = cos > 1e-10  # the sphere shadows the points that face away
    refl = d - 2.0 * np.sum(d * n, axis=1)[:, None] * n
    ca = np.sum(refl * ldir, axis=1)
    want = np.tile(0.1 * pig, (400, 1))
    diff = 0.6 * np.abs(cos) ** brilliance
    want[lit] += diff[lit, None] * pig
    spec = np.where(lit & (ca > 0), 0.5 * np.maximum(ca, 0) ** 7.0, 0.0)
    want += spec[:, None]
    assert lit.sum() > 100 and (~lit).sum() > 20
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-7)


def test_centre_pixel_of_lit_sphere():
    """Eye ray along the normal at the sphere's front pole, light at angle
    theta from that normal: cos = R.L = cos(theta) exactly."""
    th = math.radians(35.0)
    hp = np.array([0.0, 0.0, -1.0])
    L = hp + 20.0 * np.array([math.sin(th), 0.0, -math.cos(th)])
    s = _scene([(A.PVX_SPHERE, 0)], dict(pigment=[[1.0, 0.5, 0.25]], ambient=[0.2], diffuse=[0.7],
                                          phong=[0.9], phong_size=[3.0]),
               lights=[dict(kind="point", position=L, colour=np.array([1.0, 0.8, 0.6]))],
               sphere_center=np.zeros((1, 3)), sphere_radius=np.array([1.0]))
    got = OracleScene(s).trace_rays(s, _rays([0.0, 0.0, -5.0], [[0.0, 0.0, 1.0]]))[0]
    lc = np.array([1.0, 0.8, 0.6])
    pig = np.array([1.0, 0.5, 0.25])
    want = 0.2 * pig + 0.7 * math.cos(th) * pig * lc + 0.9 * math.cos(th) ** 3 * lc
    np.testing.assert_allclose(got, want, rtol=2e-6)


# ---------------------------------------------------------------------------
# ComputeReflection (trace.cpp:1264-1321)
# ---------------------------------------------------------------------------
def test_mirror_plane_image_of_emitter():
    """A full mirror y = 0 under a self-lit sphere: a pixel shows the sphere
    (colour 1) iff the eye ray, or its mirror image (y -> -y) after the
    plane, meets the sphere; otherwise the black background."""
    c = np.array([0.3, 1.2, 4.0])
    s = _scene([(A.PVX_PLANE, 0), (A.PVX_SPHERE, 0)],
               dict(pigment=[[0, 0, 0], [1, 1, 1]], ambient=[0.0, 1.0], reflection=[1.0, 0.0]),
               plane_normal=np.array([[0.0, 1.0, 0.0]]), plane_distance=np.array([0.0]),
               sphere_center=c[None], sphere_radius=np.array([0.6]))
    rng = np.random.default_rng(7)
    o = np.array([0.0, 1.0, -5.0])
    d = _unit(np.column_stack([rng.uniform(-0.2, 0.2, 3000), rng.uniform(-0.35, 0.1, 3000), np.ones(3000)]))
    got = OracleScene(s).trace_rays(s, _rays(o, d))
    t_direct, disc1 = _sphere_hit(o, d, c, 0.6)
    t_plane = np.where(d[:, 1] < 0, -o[1] / d[:, 1], np.inf)
    see = np.isfinite(t_direct) & (t_direct < t_plane)
    p = o + np.where(np.isfinite(t_plane), t_plane, 0.0)[:, None] * d
    r = d * np.array([1.0, -1.0, 1.0])
    t_mirror, disc2 = _sphere_hit(p, r, c, 0.6)
    see |= ~see & np.isfinite(t_plane) & ~(np.isfinite(t_direct) & (t_direct < t_plane)) & np.isfinite(t_mirror)
    clear = (np.abs(disc1) > 1e-6) & (np.abs(disc2) > 1e-6)  # not grazing the silhouette
    assert see[clear].sum() > 200 and (~see[clear]).sum() > 200
    np.testing.assert_array_equal(got[clear, 1], np.where(see[clear], 1.0, 0.0).astype(np.float32))


# ---------------------------------------------------------------------------
# ComputeRefraction / TraceRefractionRay (trace.cpp:1323-1483)
# ---------------------------------------------------------------------------
def _glass(n_objects, glass_index, ior=1.5):
    tr = np.zeros(n_objects)
    tr[glass_index] = 1.0
    return tr, np.where(tr > 0, ior, 1.0)


def test_slab_refraction_lateral_shift():
    """Glass slab 0 <= z <= T (ior 1.5) in front of a self-lit sphere.  In
    the plane y = 0 a ray at incidence theta crosses the slab at
    theta' = asin(sin(theta)/1.5) (Snell in angle form) and leaves it
    parallel to itself at x_in + T tan(theta'); the pixel is 1 iff that ray
    meets the sphere."""
    T = 3.0
    c = np.array([4.0, 0.0, 8.0])
    transmit, ior = _glass(2, 0)
    s = _scene([(A.PVX_BOX, 0), (A.PVX_SPHERE, 0)],
               dict(pigment=[[1, 1, 1], [1, 1, 1]], ambient=[0.0, 1.0], transmit=transmit, ior=ior),
               box_corner1=np.array([[-50.0, -50.0, 0.0]]), box_corner2=np.array([[50.0, 50.0, T]]),
               box_transform=np.array([-1], np.int32),
               sphere_center=c[None], sphere_radius=np.array([0.5]))
    theta = np.linspace(-0.6, 0.6, 2401)
    o = np.array([0.0, 0.0, -3.0])
    d = np.column_stack([np.sin(theta), np.zeros_like(theta), np.cos(theta)])
    got = OracleScene(s).trace_rays(s, _rays(o, d))[:, 0]
    x_in = o[0] + 3.0 * np.tan(theta)
    th_in = np.arcsin(np.sin(theta) / 1.5)
    x_out = x_in + T * np.tan(th_in)
    exit_pt = np.column_stack([x_out, np.zeros_like(x_out), np.full_like(x_out, T)])
    t, disc = _sphere_hit(exit_pt, d, c, 0.5)
    want = np.isfinite(t)
    clear = np.abs(disc) > 1e-4
    assert want[clear].sum() > 50 and (~want[clear]).sum() > 1000
    np.testing.assert_array_equal(got[clear], want[clear].astype(np.float32))
    # an unrefracted ray would light a visibly different set of pixels
    straight = np.column_stack([o[0] + (3.0 + T) * np.tan(theta), 0 * theta, np.full_like(theta, T)])
    assert (np.isfinite(_sphere_hit(straight, d, c, 0.5)[0])[clear] != want[clear]).sum() > 30


def test_total_internal_reflection_in_glass_cube():
    """A ray entering the face z = 0 of a glass cube [0,1]^3 at incidence
    theta meets the side x = 1 at 90 deg - theta' >= 48.2 deg, above the
    critical angle asin(1/1.5) = 41.8 deg: it is totally reflected, leaves
    through z = 1 at x = 1 - (1 - z1) tan(theta') and goes on with direction
    (-sin(theta), 0, cos(theta)).  A self-lit sphere placed on that exit ray
    is seen; one placed where a transmitted ray would go is not."""
    theta = math.radians(40.0)
    thp = math.asin(math.sin(theta) / 1.5)
    x0 = 0.55
    z1 = (1.0 - x0) / math.tan(thp)
    assert z1 < 1.0
    x_exit = 1.0 - (1.0 - z1) * math.tan(thp)
    dir_out = np.array([-math.sin(theta), 0.0, math.cos(theta)])
    seen_at = np.array([x_exit, 0.5, 1.0]) + 4.0 * dir_out
    # where the ray would go had it been transmitted at x = 1
    t_side = np.array([1.0, 0.5, z1]) + 4.0 * np.array([math.sin(theta), 0.0, math.cos(theta)])
    transmit, ior = _glass(3, 0)
    s = _scene([(A.PVX_BOX, 0), (A.PVX_SPHERE, 0), (A.PVX_SPHERE, 1)],
               dict(pigment=[[1, 1, 1], [1, 0, 0], [0, 1, 0]], ambient=[0.0, 1.0, 1.0], transmit=transmit,
                    ior=ior),
               box_corner1=np.array([[0.0, 0.0, 0.0]]), box_corner2=np.array([[1.0, 1.0, 1.0]]),
               box_transform=np.array([-1], np.int32),
               sphere_center=np.stack([seen_at, t_side]), sphere_radius=np.array([0.05, 0.05]))
    o = np.array([x0 - 2.0 * math.tan(theta), 0.5, -2.0])
    got = OracleScene(s).trace_rays(s, _rays(o, [[math.sin(theta), 0.0, math.cos(theta)]]))[0]
    np.testing.assert_array_equal(got, np.array([1.0, 0.0, 0.0], np.float32))


# ---------------------------------------------------------------------------
# Area lights (TraceAreaLightShadowRay / TraceAreaLightSubsetShadowRay,
# trace.cpp:2131-2272)
# ---------------------------------------------------------------------------
def _segment_hits_box(p, q, lo, hi):
    d = q - p
    t0, t1 = 0.0, 1.0
    for a in range(3):
        if abs(d[a]) < 1e-15:
            if p[a] < lo[a] or p[a] > hi[a]:
                return False, 1.0
            continue
        ta, tb = (lo[a] - p[a]) / d[a], (hi[a] - p[a]) / d[a]
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    margin = min(abs(t1 - t0), 1.0)
    return t0 <= t1, margin


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_area_light_visible_fraction(seed):
    """A 4x4 area light over a lit floor point: with no occluder the pixel
    equals the same scene lit by a point light at the light's centre; with
    a box occluder it is that value times the fraction of the 16 grid
    points (axis * (i/3 - 1/2)) whose segment to the point misses the box."""
    rng = np.random.default_rng(seed)
    P = np.array([rng.uniform(-0.5, 0.5), 0.0, rng.uniform(-0.5, 0.5)])
    C = np.array([0.0, 5.0, 0.0])
    ax1, ax2 = np.array([2.0, 0.0, 0.0]), np.array([0.0, 0.0, 2.0])
    lo = np.array([rng.uniform(-3.0, -0.5), 2.0, rng.uniform(-3.0, -0.5)])
    hi = np.array([rng.uniform(-0.2, 0.9), 2.2, rng.uniform(-0.2, 0.9)])
    grid = [C + ax1 * (i / 3 - 0.5) + ax2 * (j / 3 - 0.5) for i in range(4) for j in range(4)]
    vis = []
    for g in grid:
        hit, margin = _segment_hits_box(P, g, lo, hi)
        vis.append(not hit)
    frac = sum(vis) / 16.0

    def render(light, occluded):
        objs = [(A.PVX_PLANE, 0)] + ([(A.PVX_BOX, 0)] if occluded else [])
        n = len(objs)
        sc = _scene(objs, dict(pigment=np.ones((n, 3)), diffuse=np.r_[1.0, np.zeros(n - 1)]), lights=[light],
                    plane_normal=np.array([[0.0, 1.0, 0.0]]), plane_distance=np.array([0.0]),
                    box_corner1=lo[None] if occluded else None, box_corner2=hi[None] if occluded else None,
                    box_transform=np.array([-1], np.int32) if occluded else None)
        o = P + np.array([0.0, 1.0, -1.0])
        return OracleScene(sc).trace_rays(sc, _rays(o, [P - o]))[0]

    area = dict(kind="area", position=C, axis1=ax1, axis2=ax2, size=(4, 4), colour=np.ones(3))
    point = dict(kind="point", position=C, colour=np.ones(3))
    ref = render(point, False)
    assert ref[0] > 0.1
    np.testing.assert_array_equal(render(area, False), ref)
    got = render(area, True)
    np.testing.assert_allclose(got, ref * frac, rtol=1e-6)
    assert 0.0 < frac < 1.0 or seed != 1


# ---------------------------------------------------------------------------
# Radiosity: direction pool (randomsequence.cpp:730-747, 907-914)
# ---------------------------------------------------------------------------
def _radical_inverse(i, base):
    h, q = 0.0, 1.0 / base
    while i > 0:
        h += (i % base) * q
        i //= base
        q /= base
    return h


def test_halton_pool_is_cosine_weighted_radical_inverse():
    """Entry i of the pool is the cosine-weighted hemisphere point of the
    2-D Halton point (phi_2(i+1), phi_3(i+1)): r^2 = x^2 + z^2 = phi_2 and
    atan2(z, x) = 2 pi phi_3, y = sqrt(1 - r^2) up.  Over the pool the
    cosine-weighted moments hold: E[y] = 2/3, E[y^2] = 1/2, E[x] = E[z] = 0."""
    pool = O.halton_pool()
    assert pool.shape == (1600, 3)
    i = np.arange(1600)
    u = np.array([_radical_inverse(k + 1, 2) for k in i])
    v = np.array([_radical_inverse(k + 1, 3) for k in i])
    np.testing.assert_allclose(pool[:, 0] ** 2 + pool[:, 2] ** 2, u, atol=1e-12)
    ang = np.mod(np.arctan2(pool[:, 2], pool[:, 0]), 2 * np.pi) / (2 * np.pi)
    wrap = np.minimum(np.abs(ang - v), 1 - np.abs(ang - v))
    assert wrap.max() < 1e-9
    np.testing.assert_allclose(np.linalg.norm(pool, axis=1), 1.0, atol=1e-12)
    assert np.all(pool[:, 1] >= 0)
    assert abs(pool[:, 1].mean() - 2 / 3) < 2e-3
    assert abs((pool[:, 1] ** 2).mean() - 0.5) < 2e-3
    assert abs(pool[:, 0].mean()) < 5e-3 and abs(pool[:, 2].mean()) < 5e-3
    # the first points of the two sequences are the textbook ones
    np.testing.assert_allclose(u[:4], [0.5, 0.25, 0.75, 0.125])
    np.testing.assert_allclose(v[:4], [1 / 3, 2 / 3, 1 / 9, 4 / 9])


# ---------------------------------------------------------------------------
# Radiosity: AverageNearBlock (radiosity.cpp:1575-1720)
# ---------------------------------------------------------------------------
def _rec(p, n, E, r=1.0, nearest=None, to_nearest=(0.0, 1.0, 0.0), depth=0):
    return np.r_[p, n, E, r, r if nearest is None else nearest, to_nearest, depth]


def test_average_near_block_exact_reuse_and_symmetry():
    up = np.array([0.0, 1.0, 0.0])
    # a record at the query point with the query's normal: full weight
    E, W = O.rad_average(_rec([1, 0, 1], up, [0.2, 0.4, 0.6]), 0.5, [1, 0, 1], up, 0.5)
    np.testing.assert_allclose(E, [0.2, 0.4, 0.6])
    assert W == pytest.approx(1.0)
    # two records at equal distance on either side: the mean of the two
    recs = np.stack([_rec([0.9, 0, 1], up, [1, 0, 0]), _rec([1.1, 0, 1], up, [0, 0, 1])])
    E, W = O.rad_average(recs, 0.5, [1, 0, 1], up, 0.5)
    np.testing.assert_allclose(E, [0.5, 0.0, 0.5], rtol=1e-12)


def test_average_near_block_ward_weights():
    """Ward's weight with POV-Ray's error: err = d/R + 2 sqrt(1 - n.n');
    w = 1 - err/bound while err < bound (plain passes), the pseudo-sigmoid
    1 - (2x)^2/2 below x = err/bound = 1/2 and (2(1 - x))^2/2 above it on the
    final pass; zero from err = bound on."""
    up = np.array([0.0, 1.0, 0.0])
    bound, R = 0.5, 2.0
    for d in (0.0, 0.1, 0.3, 0.6, 0.9, 0.99, 1.01, 1.5):
        rec = _rec([d, 0, 0], up, [1, 1, 1], r=R, to_nearest=(0.0, 1.0, 0.0))
        _, W = O.rad_average(rec, bound, [0, 0, 0], up, bound)
        x = (d / R) / bound
        assert W == pytest.approx(max(0.0, 1.0 - x) if 1.0 - x > 0.001 else 0.0, abs=1e-12), d
        _, Wf = O.rad_average(rec, bound, [0, 0, 0], up, bound, final=True)
        want = (1 - (2 * x) ** 2 / 2 if x < 0.5 else (2 * (1 - x)) ** 2 / 2) if x < 1 else 0.0
        assert Wf == pytest.approx(want if want > 0.001 else 0.0, abs=1e-12), d
    # normal deviation at the same point
    for deg in (2.0, 5.0, 8.0):
        a = math.radians(deg)
        n2 = np.array([math.sin(a), math.cos(a), 0.0])
        _, W = O.rad_average(_rec([0, 0, 0], n2, [1, 1, 1], r=R), bound, [0, 0, 0], up, bound)
        err = 2.0 * math.sqrt(1.0 - math.cos(a))
        assert W == pytest.approx(max(0.0, 1.0 - err / bound), abs=1e-12), deg


def test_average_near_block_in_front_and_nearest_surface():
    up = np.array([0.0, 1.0, 0.0])
    bound = 0.5
    # a query behind the record's tangent plane (the record lies above it)
    # does not use the record
    rec = _rec([0.0, 0.3, 0.3], up, [1, 1, 1], r=2.0)
    _, W = O.rad_average(rec, bound, [0, 0, 0], up, bound)
    assert W == 0.0
    # the mirrored offset, query in front of the record, does
    rec = _rec([0.0, -0.3, 0.3], up, [1, 1, 1], r=2.0)
    _, W = O.rad_average(rec, bound, [0, 0, 0], up, bound)
    assert W > 0.0
    # towards the record's nearest surface the validity radius shrinks to
    # c * nearest + (1 - c) * R, c = cos(angle to that direction)
    d = 0.4
    to_near = np.array([-1.0, 0.0, 0.0])  # from the record towards the query
    rec = _rec([d, 0, 0], up, [1, 1, 1], r=2.0, nearest=0.9, to_nearest=to_near)
    _, W = O.rad_average(rec, bound, [0, 0, 0], up, bound)
    ri = 1.0 * 0.9 + 0.0 * 2.0
    assert W == pytest.approx(1.0 - (d / ri) / bound, abs=1e-12)


# ---------------------------------------------------------------------------
# Radiosity: one bounce under a rectangular emitter vs the form factor
# ---------------------------------------------------------------------------
def _corner_form_factor(X, Y):
    """Differential area -> parallel rectangle with one corner above it,
    sides X, Y in units of the height (the textbook closed form)."""
    a, b = math.sqrt(1 + X * X), math.sqrt(1 + Y * Y)
    return (X / a * math.atan(Y / a) + Y / b * math.atan(X / b)) / (2 * math.pi)


def _rect_form_factor(px, pz, x0, x1, z0, z1, h):
    f = 0.0
    for xa, sx in ((x1 - px, 1), (x0 - px, -1)):
        for za, sz in ((z1 - pz, 1), (z0 - pz, -1)):
            X, Y = abs(xa) / h, abs(za) / h
            sgn = (1 if xa >= 0 else -1) * sx * (1 if za >= 0 else -1) * sz
            f += sgn * _corner_form_factor(X, Y)
    return f


def test_one_bounce_radiosity_matches_rectangle_form_factor():
    """White diffuse floor (diffuse 1) under a self-lit panel (emission Le)
    at height h, nothing else, radiosity count 1600 (the whole pool),
    recursion_limit 1, no reuse: a floor pixel is Le times the point's form
    factor to the panel, up to the quadrature error of 1600 Halton
    directions."""
    h, Le = 2.0, 3.0
    x0, x1, z0, z1 = -1.0, 1.5, -0.5, 2.0
    mats = dict(pigment=[[1, 1, 1], [1, 1, 1]], diffuse=[1.0, 0.0],
                emission=[[0, 0, 0], [Le, Le, Le]])
    s = _scene([(A.PVX_PLANE, 0), (A.PVX_BOX, 0)], mats,
               plane_normal=np.array([[0.0, 1.0, 0.0]]), plane_distance=np.array([0.0]),
               box_corner1=np.array([[x0, h, z0]]), box_corner2=np.array([[x1, h + 0.01, z1]]),
               box_transform=np.array([-1], np.int32))
    s.camera = scenes.look_at_camera((0.2, 1.0, -1.5), (0.2, 0.0, 0.8), 16, 12)
    img, cnt = OracleScene(s).radiosity_render(s, count=1600, recursion_limit=1, error_bound=1e-6,
                                               minimum_reuse=0.015, maximum_reuse=0.2, brightness=1.0,
                                               nearest_count=1, pretrace_start=0.08, pretrace_end=0.04)
    rays = scenes.camera_rays_np(s.camera)
    floor = rays[:, 4] < -0.05  # pixels that see the floor (the others look up)
    assert floor.sum() > 120
    rays = rays[floor]
    t = -rays[:, 1] / rays[:, 4]
    p = rays[:, :3] + t[:, None] * rays[:, 3:]
    ff = np.array([_rect_form_factor(q[0], q[2], x0, x1, z0, z1, h) for q in p])
    got = img.reshape(-1, 3)[floor, 0] / Le
    assert (ff > 0.05).sum() > 60 and ff.max() > 0.2
    np.testing.assert_allclose(got, ff, atol=0.01)
    assert np.abs(got - ff).mean() < 0.003


# ---------------------------------------------------------------------------
# Solve_Polynomial branches (polynomialsolver.cpp:1047-1087, 1298-1450, 1585-1732)
# ---------------------------------------------------------------------------
def _poly(roots, lead=1.0):
    return lead * np.poly(np.asarray(roots, np.float64))


def test_vieta_quartic_roots():
    rng = np.random.default_rng(11)
    for _ in range(200):
        r = np.sort(rng.uniform(0.1, 20.0, 4))
        if np.min(np.diff(r)) < 0.05:
            continue
        got = np.sort(O.solve_quartic_vieta(_poly(r, rng.uniform(0.5, 3.0))))
        np.testing.assert_allclose(got, r, rtol=1e-6)
    # two real roots and a complex pair
    c = np.polymul(np.poly([2.0, 5.0]), [1.0, 0.0, 4.0])
    np.testing.assert_allclose(np.sort(O.solve_quartic_vieta(c)), [2.0, 5.0], rtol=1e-9)
    # roots symmetric about their mean: the depressed quartic has no odd
    # term, and the resolvent's 2z - p vanishes (the d1 < SMALL_ENOUGH branch)
    c = _poly([1.0, 2.0, 4.0, 5.0])
    np.testing.assert_allclose(np.sort(O.solve_quartic_vieta(c)), [1.0, 2.0, 4.0, 5.0], rtol=1e-9)


def test_difficult_coeffs_branch_choice():
    """Solve_Polynomial solves an order-4 polynomial by Sturm sequences when
    a coefficient is more than 1e12 times smaller than `biggest`.  The
    reference's loop stores the SIGNED coefficient as `biggest`, so once a
    negative coefficient is kept every later coefficient replaces it: the
    switch then depends on the signs, which these cases pin."""
    spread = [1e13, -3.0, 1.0, 2.0, -1.0]
    assert O.difficult_coeffs(spread)                    # positive biggest first
    assert not O.difficult_coeffs([1.0, -2.0, 3.0, 1.0, -1.0])
    # negative biggest first: the next coefficient replaces it
    assert not O.difficult_coeffs([-1e13, 1.0, 2.0, 3.0, 4.0])
    # ... unless a later one is again 1e12 times larger than another
    assert O.difficult_coeffs([-1e13, 1e-13, 2.0, 3.0, 4.0])
    # the chosen branch is the one Solve_Polynomial runs (sturm off)
    for c in (spread, [-1e13, 1.0, 2.0, 3.0, 4.0], _poly([1.0, 2.0, 3.0, 7.0])):
        c = np.asarray(c, np.float64)
        want = O.polysolve(c) if O.difficult_coeffs(c) else O.solve_quartic_vieta(c)
        np.testing.assert_array_equal(O.solve_polynomial(c, sturm=False, epsilon=0.0), np.sort(want))


# ---------------------------------------------------------------------------
# Box::Intersect side choice (box.cpp:172-505)
# ---------------------------------------------------------------------------
def test_box_sides_on_faces_and_edges():
    c1, c2 = np.array([-1.0, -2.0, -0.5]), np.array([1.5, 1.0, 2.0])
    rng = np.random.default_rng(3)
    n_checked = 0
    for _ in range(500):
        o = rng.uniform(-6, 6, 3)
        if np.all((o > c1) & (o < c2)):
            continue
        target = rng.uniform(c1, c2)
        d = _unit(target - o)
        hit, d1, d2, s1, s2 = O.box_intersect(c1, c2, np.r_[o, d])
        assert hit
        # slab entry / exit
        with np.errstate(divide="ignore"):
            ta, tb = (c1 - o) / d, (c2 - o) / d
        tn, tf = np.minimum(ta, tb), np.maximum(ta, tb)
        t_in, t_out = tn.max(), tf.min()
        gap = np.sort(tn)[-1] - np.sort(tn)[-2]
        if gap < 1e-5:
            continue
        assert d1 == pytest.approx(t_in, rel=1e-12) and d2 == pytest.approx(t_out, rel=1e-12)
        p = o + d1 * d
        a = int(np.argmax(tn))
        face = 2 * a + (1 if p[a] > 0.5 * (c1[a] + c2[a]) else 0) + 1
        assert s1 == face
        n_checked += 1
    assert n_checked > 300
    # a ray aimed exactly at the edge x = c2.x, y = c2.y (entering at both
    # faces at once) is given the face of the larger direction component
    edge = np.array([c2[0], c2[1], 0.7])
    for dx, dy, want in ((-1.0, -0.4, 2), (-0.4, -1.0, 4), (-0.7, -0.71, 4)):
        d = _unit([dx, dy, 0.05])
        o = edge - 3.0 * d
        hit, d1, _, s1, _ = O.box_intersect(c1, c2, np.r_[o, d])
        assert hit and s1 == want, (dx, dy, s1)
        assert d1 == pytest.approx(3.0, rel=1e-9)


# ---------------------------------------------------------------------------
# FP32 object-box pretest (object.cpp:917-942, 1074-1108) and the shapes'
# Compute_BBox (quadric.cpp:764-1240, polygon.cpp:806-842)
# ---------------------------------------------------------------------------
def _f32_above(x):
    return float(np.float32(x)) > x


def test_quadric_boxes_closed_form():
    # ellipsoid x^2/a^2 + y^2/b^2 + z^2/c^2 = 1 moved to (1, -2, 3)
    a, b, c, m = 0.7, 1.3, 2.1, np.array([1.0, -2.0, 3.0])
    A_, E_, H_ = 1 / a ** 2, 1 / b ** 2, 1 / c ** 2
    coef = [A_, E_, H_, 0.0, 0.0, 0.0, -2 * A_ * m[0], -2 * E_ * m[1], -2 * H_ * m[2],
            A_ * m[0] ** 2 + E_ * m[1] ** 2 + H_ * m[2] ** 2 - 1.0]
    s = _scene([(A.PVX_QUADRIC, 0), (A.PVX_QUADRIC, 1), (A.PVX_QUADRIC, 2)], {},
               quadric_coef=np.array([coef, [1.0, 1.0, 0.0, 0, 0, 0, 0, 0, 0, -4.0],
                                      [1.0, 1.0, 1.0, 0.5, 0, 0, 0, 0, 0, -1.0]]),
               quadric_transform=np.array([-1, -1, -1], np.int32))
    o = OracleScene(s)
    ll, size, inf = o.object_bbox(0)
    assert not inf
    np.testing.assert_allclose(ll, m - [a, b, c], rtol=1e-6)
    np.testing.assert_allclose(ll + size, m + [a, b, c], rtol=1e-6)
    # z-cylinder of radius 2: bounded in x and y only, its z extent is
    # longer than CRITICAL_LENGTH (1e6), so the default box stays: infinite
    ll, size, inf = o.object_bbox(1)
    assert inf and size[0] == np.float32(2e10) and size[2] == np.float32(2e10)
    # a mixed term: the default box
    ll, size, inf = o.object_bbox(2)
    assert inf and size[0] > 1e9


def test_polygon_box_widens_only_thin_axes():
    pts = np.array([[0.1, 0.2, 2.0], [1.3, 0.2, 2.0], [1.3, 1.7, 2.0], [0.1, 1.7, 2.0]])
    s = _scene([(A.PVX_POLYGON, 0)], {}, polygon_first=np.array([0, 4], np.uint32), polygon_points=pts,
               polygon_transform=np.array([-1], np.int32))
    ll, size, inf = OracleScene(s).object_bbox(0)
    np.testing.assert_allclose([ll[0], ll[1]], [0.1, 0.2], rtol=1e-6)
    np.testing.assert_allclose([size[0], size[1]], [1.2, 1.5], rtol=1e-6)
    assert ll[2] == np.float32(np.float32(2.0) - 1e-3) and size[2] == np.float32(2e-3)


@pytest.mark.parametrize("kind", ["polygon", "disc"])
def test_grazing_hit_outside_fp32_box_is_culled(kind):
    """The box's lower-left corner is the FP32 rounding of the shape's
    smallest x, and float(-0.7) > -0.7: a ray perpendicular to the shape
    hitting between -0.7 and float(-0.7) is inside the shape but misses the
    box, so the reference's pretest culls it; one just inside the box is
    kept."""
    edge = -0.7
    f = float(np.float32(edge))
    assert f > edge
    if kind == "polygon":
        pts = np.array([[edge, -0.5, 0.0], [0.5, -0.5, 0.0], [0.5, 0.5, 0.0], [edge, 0.5, 0.0]])
        s = _scene([(A.PVX_POLYGON, 0)], {}, polygon_first=np.array([0, 4], np.uint32), polygon_points=pts,
                   polygon_transform=np.array([-1], np.int32))
    else:
        s = _scene([(A.PVX_DISC, 0)], {}, disc_center=np.zeros((1, 3)), disc_normal=np.array([[0.0, 0.0, 1.0]]),
                   disc_radius=np.array([[-edge, 0.0]]), disc_transform=np.array([-1], np.int32))
    o = OracleScene(s)
    ll, _, _ = o.object_bbox(0)
    assert ll[0] == np.float32(edge)
    x_out = 0.5 * (f + edge)  # inside the shape, outside the FP32 box
    x_in = f + 1e-7
    rays = np.array([[x_out, 0.0, -1.0, 0.0, 0.0, 1.0], [x_in, 0.0, -1.0, 0.0, 0.0, 1.0]])
    hits = o.find_intersection(rays)
    assert hits["obj"][0] == -1 and hits["obj"][1] == 0
    # the shape's own test does hit the first ray (All_Intersections)
    d, _, cnt = o.all_intersections(rays[:1], [0])
    assert cnt[0] == 1 and d[0, 0] == pytest.approx(1.0)


# ---------------------------------------------------------------------------
# Torus without `sturm`: Ferrari quartic, difficult_coeffs, d1 clamp
# (torus.cpp:241-329, polynomialsolver.cpp:1047-1087, 1325-1450)
# ---------------------------------------------------------------------------
def _torus_residual(s, obj, p):
    """|(|q|^2 + R^2 - r^2)^2 - 4 R^2 (qx^2 + qz^2)| / (4 R^2 r^2) at the
    torus-space point q of world point p (0 on the surface)."""
    idx = s.objects[obj, 1]
    inv = s.transforms[s.torus_transform[idx], 16:].reshape(4, 4)
    q = np.c_[p, np.ones(len(p))] @ inv
    R, r = s.torus_major[idx], s.torus_minor[idx]
    k = (q[:, :3] ** 2).sum(1) + R * R - r * r
    return np.abs(k * k - 4.0 * R * R * (q[:, 0] ** 2 + q[:, 2] ** 2)) / (4.0 * R * R * r * r)


def test_torus_without_sturm_reaches_every_branch_and_hits_the_surface():
    s = scenes.torus_scene(width=160, height=120)
    rays = np.concatenate([scenes.camera_rays_np(s.camera), scenes.torus_plane_rays(s)])
    o = OracleScene(s)
    O.solver_branches()
    h = o.find_intersection(rays, method=0)
    br = O.solver_branches()
    assert br["quartic"] > 0 and br["difficult"] > 0 and br["d1_clamp"] > 0 and br["sturm"] == 0, br
    hit = h["obj"] >= 0
    assert hit.sum() > 500
    p = rays[hit, :3] + h["t"][hit, None] * rays[hit, 3:]
    res = np.array([_torus_residual(s, ob, pp[None])[0] for ob, pp in zip(h["obj"][hit], p)])
    # distance to the surface is about res * r / 2: within 5e-6 of the tube radius
    assert res.max() < 1e-5, res.max()
    # the closed-form and Sturm solvers find the same first crossings
    s2 = scenes.torus_scene(width=160, height=120, sturm=True)
    h2 = OracleScene(s2).find_intersection(rays, method=0)
    assert O.solver_branches()["sturm"] > 0
    same = h["obj"] == h2["obj"]
    assert same.mean() > 0.999
    both = same & hit
    np.testing.assert_allclose(h["t"][both], h2["t"][both], rtol=1e-6)


# ---------------------------------------------------------------------------
# Placed mesh2 (Mesh::Transform / Mesh::Intersect / test_hit / Mesh::Normal,
# mesh.cpp:179-230, 342-380, 568-580, 1218-1243)
# ---------------------------------------------------------------------------
def _moller_trumbore(o, d, a, b, c):
    """Nearest t > 1e-9 of the ray against triangle abc (FP64, world space)."""
    e1, e2 = b - a, c - a
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-14:
        return np.inf
    s = o - a
    u = (s @ p) / det
    q = np.cross(s, e1)
    v = (d @ q) / det
    t = (e2 @ q) / det
    if u < 0 or v < 0 or u + v > 1 or t <= 1e-9:
        return np.inf
    return t


def test_placed_mesh_hits_and_normals_in_world_space():
    s = scenes.placed_mesh_scene(level=2, width=64, height=48)
    o = OracleScene(s)
    rays = scenes.camera_rays_np(s.camera)
    h = o.find_intersection(rays, method=1)
    assert np.array_equal(h["obj"], o.find_intersection(rays, method=0)["obj"])
    v, f, _ = s.meshes[0]
    v = v.astype(np.float64)
    checked = 0
    for i in np.nonzero((h["obj"] >= 1) & (h["obj"] <= 3))[0][::7]:
        obj, tri = int(h["obj"][i]), int(h["prim"][i])
        m = s.transforms[s.mesh_transform[obj]][:16].reshape(4, 4)
        w = v[f[tri]] @ m[:3, :3] + m[3, :3]  # the triangle in world space
        t = _moller_trumbore(rays[i, :3], rays[i, 3:], *w)
        if not np.isfinite(t):  # edge-ambiguous: decided by the FP32 mesh-space test
            continue
        # the reference's plane is FP32 (Distance, normal): 1e-6 relative
        assert h["t"][i] == pytest.approx(t, rel=1e-6), (i, h["t"][i], t)
        checked += 1
    assert checked > 100
    # a placed flat mesh is lit through its face normal sent out by the inverse
    # transpose (MTransNormal): one triangle scaled non-uniformly and rotated,
    # a point light, diffuse 1: the pixel is N.L
    tri = np.array([[0, 1, 2]], np.int32)
    pv = np.array([[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [0.0, 1.0, 0.0]], np.float32)
    place = scenes.compose((np.diag([1.0, 3.0, 1.0, 1.0]), np.diag([1.0, 1.0 / 3.0, 1.0, 1.0])),
                           scenes.rotation_transform((35.0, 20.0, 0.0)))
    light = np.array([3.0, 4.0, -6.0])
    flat = _scene([(A.PVX_MESH, 0)], dict(diffuse=[1.0]), meshes=[(pv, tri, None)],
                  lights=[dict(kind="point", position=light, colour=np.ones(3))],
                  transforms=np.concatenate([place[0].ravel(), place[1].ravel()])[None],
                  mesh_transform=np.array([0], np.int32))
    fo = OracleScene(flat)
    target = np.array([0.1, -0.2, 0.0]) @ place[0][:3, :3] + place[0][3, :3]
    eye = np.array([0.5, 0.5, -6.0])
    ray = _rays(eye, target - eye).reshape(1, 6)
    hh = fo.find_intersection(ray, method=0)
    assert hh["obj"][0] == 0
    hp = ray[0, :3] + hh["t"][0] * ray[0, 3:]
    np.testing.assert_allclose(hp, target, atol=1e-6)
    n = place[1][:3, :3] @ np.array([0.0, 0.0, 1.0])
    n /= np.linalg.norm(n)
    if n @ ray[0, 3:] > 0:
        n = -n
    lv = _unit(light - hp)
    rgb = fo.trace_rays(flat, ray)
    np.testing.assert_allclose(rgb[0], np.full(3, max(n @ lv, 0.0)), rtol=2e-6, atol=1e-7)
