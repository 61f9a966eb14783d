"""Stand-alone triangle / smooth_triangle (core/shape/triangle.cpp): the
oracle pinned by closed forms on the CPU, then the GPU against the oracle."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from povray_amd import _abi as A
from povray_amd import scenes


def _scene(verts, normals=None, smooth=None):
    verts = np.asarray(verts, np.float64).reshape(-1, 9)
    n = verts.shape[0]
    smooth = np.zeros(n, bool) if smooth is None else np.asarray(smooth, bool)
    objects = np.zeros((n, 4), np.uint32)
    objects[:, 0] = np.where(smooth, A.PVX_SMOOTH_TRIANGLE, A.PVX_TRIANGLE)
    objects[:, 1] = np.arange(n)
    return scenes.Scene(name="tri", objects=objects, triangle_vertices=verts,
                        triangle_normals=None if normals is None else np.asarray(normals, np.float64).reshape(-1, 9))


def _down(x, y, z=5.0):
    return np.array([[x, y, z, 0.0, 0.0, -1.0]])


def test_oracle_triangle_closed_forms():
    s = _scene([0, 0, 0, 1, 0, 0, 0, 1, 0])
    o = OracleScene(s)
    h = o.find_intersection(_down(0.25, 0.25))
    assert h["obj"][0] == 0 and h["t"][0] == 5.0  # plane depth, exact
    assert o.find_intersection(_down(0.6, 0.6))["obj"][0] == -1       # beyond the hypotenuse
    assert o.find_intersection(_down(0.5, 0.5))["obj"][0] == 0        # on the edge: inside
    assert o.find_intersection(_down(-0.01, 0.5))["obj"][0] == -1
    assert o.find_intersection(_down(0.25, 0.25, -5.0))["obj"][0] == -1  # behind the origin
    par = np.array([[0.25, 0.25, 5.0, 1.0, 0.0, 0.0]])
    assert o.find_intersection(par)["obj"][0] == -1                    # parallel to the plane
    # the flat normal is the unit face normal, whatever the vertex order
    nrm = o.triangle_normal(0, [[0.25, 0.25, 0.0]])[0]
    assert np.allclose(np.abs(nrm), [0, 0, 1], atol=0) and np.linalg.norm(nrm) == 1.0
    # collinear vertices: degenerate, never hit
    d = OracleScene(_scene([0, 0, 0, 1, 1, 0, 2, 2, 0]))
    assert d.find_intersection(_down(1.0, 1.0))["obj"][0] == -1


def test_oracle_smooth_triangle_normal_interpolates_vertex_normals():
    P = np.array([0, 0, 0, 2, 0, 0, 0, 2, 0], np.float64)
    N = np.array([0.2, 0.1, 1.0, -0.1, 0.3, 1.0, 0.0, -0.2, 1.0], np.float64)
    o = OracleScene(_scene(P, N, [True]))
    got = o.triangle_normal(0, P.reshape(3, 3))
    want = N.reshape(3, 3) / np.linalg.norm(N.reshape(3, 3), axis=1, keepdims=True)
    # at each vertex the interpolation returns that vertex's (normalised) normal
    assert np.allclose(got, want, atol=1e-12)
    mid = o.triangle_normal(0, [[2 / 3, 2 / 3, 0.0]])[0]
    assert abs(np.linalg.norm(mid) - 1.0) < 1e-12 and mid[2] > 0.9
    # a vertex normal pointing the other way: marked degenerate, never hit
    Nbad = N.copy()
    Nbad[8] = -1.0
    bad = OracleScene(_scene(P, Nbad, [True]))
    assert bad.find_intersection(_down(0.5, 0.5))["obj"][0] == -1
    assert o.find_intersection(_down(0.5, 0.5))["obj"][0] == 0


def test_triangle_scene_generator():
    s = scenes.triangle_scene(n=64, n_spheres=4)
    d, keep = s.desc()
    assert d.n_triangles == 64 and s.n_objects == 64 + 4 + 1
    kinds = s.objects[:64, 0]
    assert set(kinds.tolist()) == {A.PVX_TRIANGLE, A.PVX_SMOOTH_TRIANGLE}


# ---------------------------------------------------------------- GPU
def _cmp(g, r, what):
    bad = (g["obj"] != r["obj"]) | (g["prim"] != r["prim"])
    assert not bad.any(), f"{what}: {int(bad.sum())} id mismatches"
    hit = r["obj"] >= 0
    assert np.array_equal(g["t"][hit], r["t"][hit]), what  # same FP64 formula: identical depths
    print(f"[{what}] rays {len(r)} hit frac {hit.mean():.3f}", flush=True)


@pytest.mark.gpu
def test_gpu_triangles_match_oracle(gpu):
    s = scenes.triangle_scene()
    rays = scenes.camera_rays_np(s.camera)
    gs, o = gpu.GpuScene(s), OracleScene(s)
    _cmp(gs.find_intersection(rays), o.find_intersection(rays, method=1), "triangles primary")
    _cmp(gs.find_intersection(rays), o.find_intersection(rays, method=0), "triangles vs brute force")
    md = np.random.default_rng(2).uniform(5.0, 30.0, rays.shape[0])
    _cmp(gs.find_intersection(rays, md, A.PVX_RAY_SHADOW), o.find_intersection(rays, md, A.PVX_RAY_SHADOW),
         "triangles shadow + max_depth")
    # All_Intersections and Inside of the triangle objects
    rng = np.random.default_rng(3)
    obj = rng.integers(0, 400, rays.shape[0]).astype(np.uint32)
    gd, gp, gc = gs.all_intersections(rays, obj, 4)
    od, op, oc = o.all_intersections(rays, obj, 4)
    assert np.array_equal(gc, oc) and np.array_equal(gd[gc > 0, 0], od[oc > 0, 0])
    pts = rng.uniform(-6, 6, (1000, 3))
    assert not gs.inside(pts, obj[:1000]).any()


@pytest.mark.gpu
def test_gpu_render_triangles_matches_oracle(gpu):
    s = scenes.triangle_scene(width=160, height=120)
    gs = gpu.GpuScene(s)
    gs.set_shading()
    img, _ = gs.render()
    ref, _ = OracleScene(s).render(s)
    q = lambda a: np.round(np.clip(a, 0.0, 1.0) * 255.0).astype(np.int32)
    within = np.all(np.abs(q(img) - q(ref)) <= 1, axis=-1)
    print(f"[render-triangles] within 1/255: {within.mean():.5f}", flush=True)
    assert within.mean() >= 0.999
