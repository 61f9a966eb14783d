"""Lathes (core/shape/lathe.cpp) and surfaces of revolution (core/shape/sor.cpp)
with their bounding cylinders (core/bounding/boundingcylinder.cpp).

CPU: the oracle is pinned against the surfaces themselves, evaluated here
from the textbook forms of the splines (Bernstein polynomials for Bezier
lathes, the Catmull-Rom and quadratic interpolants for cubic and quadratic
ones, a cubic Hermite interpolant of r^2 in the height for sors) and never
from the oracle's segment coefficients: the first crossing along each ray
is found by a dense march and bisection on the implicit function
F(p) = radius(p) - profile(height(p)), the membership test is F < 0 inside
the height range, and the normal is F's numerical gradient.  A linear lathe
shaped as a capped cylinder is checked against the closed form.
GPU: closest hits, light rays, intersection lists, Inside, shaded frames,
the kd-tree walk and scene batches against the oracle."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from povray_amd import _abi as A
from povray_amd import scenes

PROFILES = scenes.REVOLUTION_PROFILES


def _single(kind, pts=None, flags=0, transform=None):
    """One lathe or sor (the scene's camera and light), optionally placed."""
    pts = PROFILES[kind] if pts is None else pts
    s = scenes.revolution_scene(n=1, width=8, height=8)
    s.revolution_first = np.array([0, len(pts)], np.uint32)
    s.revolution_points = np.array(pts, np.float64)
    s.revolution_kind = np.array([kind], np.uint8)
    s.revolution_flags = np.array([flags], np.uint8)
    if transform is None:
        s.revolution_transform = np.array([-1], np.int32)
        s.transforms = None
    else:
        s.revolution_transform = np.array([0], np.int32)
        s.transforms = scenes._xf_row(transform).reshape(1, 32)
    s.objects = np.array([[A.PVX_SOR if kind == A.PVX_SOR_CUBIC else A.PVX_LATHE, 0, 0, 0]], np.uint32)
    s.csg = s.csg_children = None
    s.box_corner1 = s.box_corner2 = s.sphere_center = s.sphere_radius = None
    s.plane_normal = s.plane_distance = None
    s.materials = {k: v[:1] for k, v in s.materials.items()}
    return s


# --- the surfaces from textbook spline forms ---------------------------------------
def _segments(kind, pts):
    """Per segment: a function w -> (x(w), y(w)) on [0, 1] (lathes), or the
    (y0, y1, r^2(h)) of a sor segment."""
    P = np.asarray(pts, float)
    out = []
    if kind == A.PVX_LATHE_LINEAR:
        for i in range(len(P) - 1):
            out.append(lambda w, a=P[i], b=P[i + 1]: (1 - w) * a + w * b)
    elif kind == A.PVX_LATHE_QUADRATIC:  # through P[i+1] (w = 0) and P[i+2] (w = 1), slope (P[i+2]-P[i])/2
        for i in range(len(P) - 2):
            out.append(lambda w, p0=P[i], p1=P[i + 1], p2=P[i + 2]:
                       p1 + 0.5 * (p2 - p0) * w + 0.5 * (p0 - 2 * p1 + p2) * w * w)
    elif kind == A.PVX_LATHE_CUBIC:  # Catmull-Rom
        for i in range(len(P) - 3):
            out.append(lambda w, p0=P[i], p1=P[i + 1], p2=P[i + 2], p3=P[i + 3]:
                       0.5 * (2 * p1 + (p2 - p0) * w + (2 * p0 - 5 * p1 + 4 * p2 - p3) * w * w
                              + (3 * p1 - p0 - 3 * p2 + p3) * w ** 3))
    elif kind == A.PVX_LATHE_BEZIER:  # Bernstein form
        for i in range(0, len(P), 4):
            out.append(lambda w, q=P[i:i + 4]: ((1 - w) ** 3 * q[0] + 3 * (1 - w) ** 2 * w * q[1]
                                                + 3 * (1 - w) * w * w * q[2] + w ** 3 * q[3]))
    else:  # sor: cubic Hermite of r^2 with end slopes 2 r (chord slope)
        for i in range(len(P) - 3):
            y0, y1 = P[i + 1, 1], P[i + 2, 1]
            f0, f1 = P[i + 1, 0] ** 2, P[i + 2, 0] ** 2
            m0 = 2 * P[i + 1, 0] * (P[i + 2, 0] - P[i, 0]) / (P[i + 2, 1] - P[i, 1])
            m1 = 2 * P[i + 2, 0] * (P[i + 3, 0] - P[i + 1, 0]) / (P[i + 3, 1] - P[i + 1, 1])

            def r2(h, y0=y0, y1=y1, f0=f0, f1=f1, m0=m0, m1=m1):
                L = y1 - y0
                s = (h - y0) / L
                return ((2 * s ** 3 - 3 * s * s + 1) * f0 + (s ** 3 - 2 * s * s + s) * L * m0
                        + (-2 * s ** 3 + 3 * s * s) * f1 + (s ** 3 - s * s) * L * m1)
            out.append((y0, y1, r2))
    return out


def _profile_at(kind, segs, h):
    """All profile radii at height h (lathes: every segment point with
    y(w) = h, w from the roots of the cubic y(w) fitted through four samples
    of the textbook form); sor: sqrt(r^2(h))."""
    if kind == A.PVX_SOR_CUBIC:
        for y0, y1, r2 in segs:
            if y0 <= h <= y1:
                return [np.sqrt(max(r2(h), 0.0))]
        return []
    rs = []
    w4 = np.array([0.0, 1.0 / 3.0, 2.0 / 3.0, 1.0])
    for f in segs:
        yc = np.polyfit(w4, f(w4[:, None])[:, 1], 3)  # y(w) is a cubic (or less) in w
        yc[-1] -= h
        yc = np.trim_zeros(np.where(np.abs(yc) < 1e-14, 0.0, yc), "f")
        if len(yc) < 2:
            continue
        for w in np.roots(yc):
            if abs(w.imag) < 1e-9 and -1e-12 <= w.real <= 1.0 + 1e-12:
                rs.append(abs(f(min(max(w.real, 0.0), 1.0))[0]))
    return rs


def _inside_local(kind, segs, p):
    """Point membership of the solid: the horizontal ray from p outwards
    crosses the profile an odd number of times (a sor: r < profile)."""
    r = np.hypot(p[0], p[2])
    rs = _profile_at(kind, segs, p[1])
    if kind == A.PVX_SOR_CUBIC:
        return bool(rs) and r < rs[0]
    return sum(1 for x in rs if x > r) % 2 == 1


def _inside_many(kind, segs, P):
    """_inside_local for many points at once, with the lathe profile as a
    polyline of 400 pieces per segment (for the march; the crossing found
    is then refined with the exact form)."""
    r = np.hypot(P[:, 0], P[:, 2])
    h = P[:, 1]
    if kind == A.PVX_SOR_CUBIC:
        out = np.zeros(len(P), bool)
        done = np.zeros(len(P), bool)
        for y0, y1, r2 in segs:
            m = (h >= y0) & (h <= y1) & ~done
            out[m] = r[m] ** 2 < r2(h[m])
            done |= m
        return out
    cnt = np.zeros(len(P), int)
    for f in segs:
        xy = f(np.linspace(0.0, 1.0, 401)[:, None])
        y0, y1, x0, x1 = xy[:-1, 1], xy[1:, 1], xy[:-1, 0], xy[1:, 0]
        g0, g1 = y0[None, :] - h[:, None], y1[None, :] - h[:, None]
        cross = (g0 < 0) != (g1 < 0)
        lam = np.where(cross, g0 / np.where(cross, g0 - g1, 1.0), 0.0)
        xr = np.abs(x0[None, :] + lam * (x1 - x0)[None, :])
        cnt += (cross & (xr > r[:, None])).sum(1)
    return cnt % 2 == 1


def _first_hit(kind, segs, o, d, tmax):
    """Depth of the first surface crossing of o + t d (shape space) on
    (1e-4, tmax): a march on a membership change, refined by bisection with
    the exact form and kept if the point lies on the profile (an open end of
    a lathe changes the membership without a surface); None without one."""
    ts = np.linspace(1e-4, tmax, 1401)
    inside = _inside_many(kind, segs, o[None, :] + ts[:, None] * d[None, :])
    for j in np.nonzero(inside[:-1] != inside[1:])[0]:
        a, b = ts[max(j - 1, 0)], ts[min(j + 2, len(ts) - 1)]
        ia = _inside_local(kind, segs, o + a * d)
        if ia == _inside_local(kind, segs, o + b * d):
            continue
        for _ in range(50):
            m = 0.5 * (a + b)
            if _inside_local(kind, segs, o + m * d) == ia:
                a = m
            else:
                b = m
        t = 0.5 * (a + b)
        if kind == A.PVX_SOR_CUBIC or _near_surface(kind, segs, o + t * d, 1e-6):
            return t
    return None


def _rays_at(s, n, seed):
    """Rays from a sphere of radius 7 around the shape towards points near
    its axis."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    o = 7.0 * u / np.linalg.norm(u, axis=1, keepdims=True) + np.array([0.0, 1.0, 0.0])
    tgt = np.column_stack([rng.uniform(-0.8, 0.8, n), rng.uniform(-0.3, 2.6, n), rng.uniform(-0.8, 0.8, n)])
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.hstack([o, d])


# --- CPU: the oracle pinned ------------------------------------------------------
def test_linear_lathe_cylinder_closed_form():
    """A linear lathe through (0,0) (1,0) (1,2) (0,2) is the closed cylinder
    r = 1, 0 <= y <= 2: hits equal the closed form, also when the lathe is
    rotated, scaled non-uniformly and moved (depths of the world ray)."""
    pts = [(0.0, 0.0), (1.0, 0.0), (1.0, 2.0), (0.0, 2.0)]
    tr = scenes.compose(scenes.compose((np.diag([1.5, 0.7, 1.2, 1.0]), np.diag([1 / 1.5, 1 / 0.7, 1 / 1.2, 1.0])),
                                       scenes.rotation_transform([20.0, 35.0, -10.0])),
                        scenes.translation_transform([0.3, -0.2, 0.5]))
    for t in (None, tr):
        s = _single(A.PVX_LATHE_LINEAR, pts, transform=t)
        rays = _rays_at(s, 600, 1)
        if t is not None:  # aim at the placed shape
            m = t[0]
            tgt = rays[:, :3] + 7.0 * rays[:, 3:]
            tgt = np.column_stack([tgt, np.ones(len(tgt))]) @ m
            rays[:, 3:] = tgt[:, :3] - rays[:, :3]
            rays[:, 3:] /= np.linalg.norm(rays[:, 3:], axis=1, keepdims=True)
        h = OracleScene(s).find_intersection(rays, method=0)
        inv = np.eye(4) if t is None else t[1]
        o = np.column_stack([rays[:, :3], np.ones(len(rays))]) @ inv
        d = rays[:, 3:] @ inv[:3, :3]
        n_hit = 0
        for i in range(len(rays)):
            best = None
            a = d[i, 0] ** 2 + d[i, 2] ** 2
            b = 2 * (o[i, 0] * d[i, 0] + o[i, 2] * d[i, 2])
            c = o[i, 0] ** 2 + o[i, 2] ** 2 - 1.0
            if a > 0 and b * b - 4 * a * c >= 0:
                for tt in ((-b - np.sqrt(b * b - 4 * a * c)) / (2 * a), (-b + np.sqrt(b * b - 4 * a * c)) / (2 * a)):
                    if tt > 1e-4 and 0.0 <= o[i, 1] + tt * d[i, 1] <= 2.0:
                        best = tt if best is None else min(best, tt)
            if abs(d[i, 1]) > 1e-12:
                for yc in (0.0, 2.0):
                    tt = (yc - o[i, 1]) / d[i, 1]
                    q = o[i] + tt * np.append(d[i], 0.0)
                    if tt > 1e-4 and q[0] ** 2 + q[2] ** 2 <= 1.0:
                        best = tt if best is None else min(best, tt)
            if best is None:
                assert h["obj"][i] == -1, i
            else:
                n_hit += 1
                assert h["obj"][i] == 0, i
                assert abs(h["t"][i] - best) <= 1e-9 * best, (i, h["t"][i], best)
        assert n_hit > 300


@pytest.mark.parametrize("kind", [A.PVX_LATHE_QUADRATIC, A.PVX_LATHE_CUBIC, A.PVX_LATHE_BEZIER, A.PVX_SOR_CUBIC])
def test_spline_surfaces_march(kind):
    """The closest hit of each spline kind (with `sturm`: roots to 1e-12
    relative) equals the first crossing of the surface found by marching the
    textbook spline form along the ray; the normal there is parallel to the
    gradient of the implicit function."""
    s = _single(kind, flags=A.PVX_REV_STURM)
    segs = _segments(kind, PROFILES[kind])
    rays = _rays_at(s, 60, 2 + kind)
    o = OracleScene(s)
    h = o.find_intersection(rays, method=0)
    nrm = o.hit_normals(rays)
    checked = 0
    for i in range(len(rays)):
        org, dirn = rays[i, :3], rays[i, 3:]
        t = _first_hit(kind, segs, org, dirn, 14.0)
        if t is None:
            assert h["obj"][i] == -1, i
            continue
        if h["obj"][i] == -1:  # a grazing touch the march can see and a root can miss
            assert _grazing(kind, segs, org, dirn, t), i
            continue
        assert abs(h["t"][i] - t) <= 1e-8 * t, (i, h["t"][i], t)
        p = org + h["t"][i] * dirn
        g = _gradient(kind, segs, p)
        if g is not None:
            c = abs(np.dot(g, nrm[i])) / np.linalg.norm(g)
            assert c > 1 - 1e-4, (i, c)
        checked += 1
    assert checked > 20


def _grazing(kind, segs, o, d, t):
    """The membership changes within 1e-3 around t and back: a tangent ray."""
    return _first_hit(kind, segs, o + (t + 1e-3) * d, d, 1e-2) is not None


def _gradient(kind, segs, p, e=1e-6):
    """Numerical gradient of F = r - profile(y) (sor: r^2 - r^2(y)) at p
    (None at a cap or a profile corner)."""
    def F(q):
        r = np.hypot(q[0], q[2])
        rs = _profile_at(kind, segs, q[1])
        if not rs:
            return None
        x = min(rs, key=lambda v: abs(v - r))
        return r * r - x * x if kind == A.PVX_SOR_CUBIC else r - x
    g = []
    for k in range(3):
        dp = np.zeros(3)
        dp[k] = e
        a, b = F(p + dp), F(p - dp)
        if a is None or b is None:
            return None
        g.append((a - b) / (2 * e))
    g = np.array(g)
    return None if np.linalg.norm(g) < 1e-3 else g


@pytest.mark.parametrize("kind", [A.PVX_LATHE_LINEAR, A.PVX_LATHE_QUADRATIC, A.PVX_LATHE_CUBIC,
                                  A.PVX_LATHE_BEZIER, A.PVX_SOR_CUBIC])
def test_inside_matches_profile(kind):
    """Lathe::Inside's crossing count and Sor::Inside's radius test equal
    the membership of the solid of the textbook profile."""
    s = _single(kind)
    segs = _segments(kind, PROFILES[kind])
    rng = np.random.default_rng(7 + kind)
    pts = np.column_stack([rng.uniform(-1.4, 1.4, 600), rng.uniform(-0.7, 3.0, 600), rng.uniform(-1.4, 1.4, 600)])
    got = OracleScene(s).inside(pts, np.zeros(len(pts), np.uint32))
    want = np.array([_inside_local(kind, segs, p) for p in pts])
    # points within 1e-6 of the surface may go either way
    near = np.array([_near_surface(kind, segs, p) for p in pts])
    if kind != A.PVX_SOR_CUBIC:
        # Lathe::Inside counts crossings only for points whose distance from
        # the axis lies in the profile's radius range (lathe.cpp:400-405)
        xs = np.concatenate([np.abs(f(np.linspace(0.0, 1.0, 20001)[:, None])[:, 0]) for f in segs])
        r = np.hypot(pts[:, 0], pts[:, 2])
        want &= (r >= xs.min()) & (r <= xs.max())
        near |= (np.abs(r - xs.min()) < 1e-6) | (np.abs(r - xs.max()) < 1e-6)
    assert np.array_equal(got.astype(bool)[~near], want[~near])
    assert want.sum() > 30


def _near_surface(kind, segs, p, e=1e-6):
    r = np.hypot(p[0], p[2])
    return any(abs(x - r) < e for x in _profile_at(kind, segs, p[1]))


def test_sor_open_and_closed():
    """A closed sor's base and cap discs are hit along the axis (depth to
    the plane y = first / last segment height); an open one is hit on its
    inside wall instead, or missed."""
    pts = PROFILES[A.PVX_SOR_CUBIC]
    rays = np.array([[0.0, -3.0, 0.0, 0.0, 1.0, 0.0], [0.0, 5.0, 0.0, 0.0, -1.0, 0.0],
                     [0.2, -3.0, 0.1, 0.0, 1.0, 0.0]])
    closed = OracleScene(_single(A.PVX_SOR_CUBIC)).find_intersection(rays, method=0)
    assert np.allclose(closed["t"], [3.0 + pts[1][1], 5.0 - pts[-2][1], 3.0 + pts[1][1]], rtol=1e-12)
    opened = OracleScene(_single(A.PVX_SOR_CUBIC, flags=A.PVX_REV_OPEN)).find_intersection(rays, method=0)
    assert np.all((opened["obj"] == -1) | (opened["t"] > closed["t"] + 1e-6))


@pytest.mark.parametrize("kind", [A.PVX_LATHE_QUADRATIC, A.PVX_LATHE_CUBIC, A.PVX_LATHE_BEZIER, A.PVX_SOR_CUBIC])
def test_sturm_agrees(kind):
    """The default roots (solve_quartic for the quadratic lathe, solve_cubic
    for the sor, Sturm sequences anyway for the order-6 cubic and Bezier
    lathes) agree with `sturm` up to the closed forms' own accuracy.  The
    sor's cubic has the leading coefficient A D.y^3: for rays within ~3
    degrees of the horizontal it is tiny and the reference's solve_cubic
    loses most digits (the reason the reference's users give a sor `sturm`);
    such rays are left out of the comparison."""
    rays = _rays_at(_single(kind), 3000, 11)
    a = OracleScene(_single(kind)).find_intersection(rays, method=0)
    b = OracleScene(_single(kind, flags=A.PVX_REV_STURM)).find_intersection(rays, method=0)
    steep = np.abs(rays[:, 4]) > 0.05
    assert np.array_equal(a["obj"][steep], b["obj"][steep])
    m = steep & (a["obj"] >= 0)
    rel = np.abs(a["t"][m] - b["t"][m]) / b["t"][m]
    assert np.mean(rel <= 1e-6) >= 0.995 and rel.max() < 1e-3
    if kind in (A.PVX_LATHE_CUBIC, A.PVX_LATHE_BEZIER):  # the same Sturm roots either way
        assert np.array_equal(a["t"], b["t"])


def test_scene_tree_equals_brute_force():
    s = scenes.revolution_scene()
    o = OracleScene(s)
    rays = scenes.camera_rays_np(s.camera)
    h1, h0 = o.find_intersection(rays, method=1), o.find_intersection(rays, method=0)
    assert np.array_equal(h1["obj"], h0["obj"]) and np.array_equal(h1["t"], h0["t"])
    assert len(np.unique(h1["obj"])) == s.n_objects + 1


@pytest.mark.parametrize("kind,pts", [
    (A.PVX_LATHE_LINEAR, [(0.0, 0.0), (-0.5, 1.0)]),           # negative radius
    (A.PVX_LATHE_BEZIER, [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0)]),  # not four points per segment
    (A.PVX_SOR_CUBIC, [(0.0, 0.0), (1.0, 1.0), (1.0, 0.5), (1.0, 2.0), (1.0, 3.0)]),  # height decreases
    (A.PVX_LATHE_CUBIC, [(0.0, 0.0), (1.0, 1.0), (1.0, 2.0)]),  # too few points
])
def test_bad_points_rejected(kind, pts):
    with pytest.raises(RuntimeError):
        OracleScene(_single(kind, pts))


# --- GPU --------------------------------------------------------------------------
def _flat_default_sor_hits(s, rays, obj):
    """Rays whose hit object is a top-level sor without `sturm` and whose
    direction, in the sor's own space, is within 0.05 of horizontal."""
    out = np.zeros(rays.shape[0], bool)
    for o in np.unique(obj[obj >= 0]):
        t, i = s.objects[o, 0], s.objects[o, 1]
        if t != A.PVX_SOR or s.revolution_flags[i] & A.PVX_REV_STURM:
            continue
        m = obj == o
        d = rays[m, 3:]
        xf = s.revolution_transform[i]
        if xf >= 0:
            d = d @ s.transforms[xf, 16:].reshape(4, 4)[:3, :3]
        out[m] = np.abs(d[:, 1]) / np.linalg.norm(d, axis=1) < 0.05
    return out


@pytest.mark.gpu
def test_gpu_revolutions_match_oracle(gpu):
    s = scenes.revolution_scene()
    gs, o = gpu.GpuScene(s), OracleScene(s)
    rays = scenes.camera_rays_np(s.camera)
    extra = np.vstack([_rays_at(_single(A.PVX_LATHE_CUBIC), 2000, 3) + np.array([x, 0, z, 0, 0, 0])
                       for x, z in ((-6.0, 0.0), (0.0, 3.0), (3.0, 6.0))])
    rays = np.vstack([rays, extra])
    md = np.random.default_rng(5).uniform(1.0, 20.0, rays.shape[0])
    for kind, m in ((A.PVX_RAY_IMAGE, None), (A.PVX_RAY_SHADOW, md)):
        g, c = gs.find_intersection(rays, m, kind), o.find_intersection(rays, m, kind, method=1)
        bad = (g["obj"] != c["obj"]) | (g["prim"] != c["prim"])
        print(f"[revolutions kind {kind}] {rays.shape[0]} rays, id mismatches {int(bad.sum())}", flush=True)
        assert not bad.any()
        # depths on a sor without `sturm` for rays within 3 degrees of its
        # horizontal come from the reference's solve_cubic with a tiny leading
        # coefficient (A D.y^3), which amplifies the last-bit differences of
        # acos / cos / pow between two math libraries (test_sturm_agrees);
        # those are compared by id only
        hit = (c["obj"] >= 0) & ~_flat_default_sor_hits(s, rays, c["obj"])
        rel = np.abs(g["t"][hit] - c["t"][hit]) / np.abs(c["t"][hit])
        worst = np.argsort(rel)[-5:]
        print(f"[revolutions kind {kind}] t: {int((rel > 1e-5).sum())} above 1e-5, bit-identical "
              f"{np.mean(rel == 0):.5f}; worst objects {c['obj'][hit][worst].tolist()} rel {rel[worst].tolist()}",
              flush=True)
        assert np.all(rel <= 1e-5)
    rng = np.random.default_rng(6)
    obj = rng.integers(0, s.n_objects - 1, 4000).astype(np.uint32)
    rr = rays[rng.integers(0, rays.shape[0], 4000)]
    (gd, gp, gc), (cd, cp, cc) = gs.all_intersections(rr, obj, 8), o.all_intersections(rr, obj, 8)
    assert np.array_equal(gc, cc) and np.array_equal(gp, cp)
    fin = np.isfinite(cd)
    assert np.array_equal(np.isfinite(gd), fin)
    assert np.all(np.abs(gd[fin] - cd[fin]) <= 1e-5 * np.abs(cd[fin]))
    pts = rng.uniform(-7.0, 7.0, (20000, 3)) * np.array([1.0, 0.3, 1.0]) + np.array([0.0, 1.2, 3.0])
    po = rng.integers(0, s.n_objects - 1, 20000).astype(np.uint32)
    assert np.array_equal(gs.inside(pts, po), o.inside(pts, po))


@pytest.mark.gpu
def test_gpu_render_revolutions_matches_oracle(gpu):
    s = scenes.revolution_scene(width=320, height=240)
    gs = gpu.GpuScene(s)
    gs.set_shading()
    img, st = gs.render()
    ref, (n_closest, n_shadow) = OracleScene(s).render(s)
    q = lambda a: np.round(np.clip(a, 0.0, 1.0) * 255.0).astype(np.int32)  # noqa: E731
    ok = float(np.mean(np.abs(q(img) - q(ref)).max(-1) <= 1))
    print(f"[revolutions shaded] {ok * 100:.3f}% of pixels within 1/255", flush=True)
    assert ok >= 0.999


@pytest.mark.gpu
def test_gpu_revolutions_kd_and_batch(gpu):
    """The kd-tree walk and a scene batch (with a scene of other shapes)
    give the single-scene hits."""
    s = scenes.revolution_scene()
    rays = scenes.camera_rays_np(s.camera)
    ref = gpu.GpuScene(s).find_intersection(rays)
    kd = gpu.GpuScene(s, accel=A.PVX_ACCEL_BSP).find_intersection(rays)
    assert np.array_equal(kd["obj"], ref["obj"]) and np.array_equal(kd["t"], ref["t"])
    other = scenes.cone_scene(width=64, height=48)
    batch = gpu.GpuBatch([gpu.GpuScene(s), gpu.GpuScene(other)])
    hb = batch.find_intersection([rays, scenes.camera_rays_np(other.camera)])
    assert np.array_equal(hb[0]["obj"], ref["obj"]) and np.array_equal(hb[0]["t"], ref["t"])
    ho = gpu.GpuScene(other).find_intersection(scenes.camera_rays_np(other.camera))
    assert np.array_equal(hb[1]["obj"], ho["obj"])
