"""Host-side scene generation: determinism, sizes, camera basis, CSG encoding."""
import numpy as np

from povray_amd import _abi as A
from povray_amd import scenes


def test_rng_is_deterministic_and_uniform():
    a = scenes.Rng(42).uniform(100000)
    b = scenes.Rng(42).uniform(100000)
    assert np.array_equal(a, b)
    assert 0.0 <= a.min() and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.005
    assert not np.array_equal(a, scenes.Rng(43).uniform(100000))


def test_icosphere_counts():
    for level in range(5):
        v, f = scenes.icosphere(level)
        assert f.shape[0] == 20 * 4 ** level
        assert v.shape[0] == 10 * 4 ** level + 2  # Euler: V - E + F = 2
        assert np.allclose(np.linalg.norm(v, axis=1), 1.0)
        # outward winding: face normal along the centroid
        n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
        assert np.all(np.sum(n * v[f].mean(1), 1) > 0)


def test_mesh_scene_sizes_and_displacement():
    s = scenes.mesh_scene(level=4)
    v, f, n = s.meshes[0]
    assert v.dtype == np.float32 and f.dtype == np.int32 and n.dtype == np.float32
    r = np.linalg.norm(v.astype(np.float64), axis=1)
    assert 0.84 < r.min() and r.max() < 1.16  # amplitude 0.15 about the unit sphere
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-6)


def test_look_at_camera_basis():
    cam = scenes.look_at_camera((0.0, 0.0, -30.0), (0.0, 0.0, 0.0), 640, 480)
    d, r, u = cam["direction"], cam["right"], cam["up"]
    assert np.allclose(d, [0, 0, 1])
    assert np.allclose(r, [640 / 480, 0, 0])  # left-handed: right = sky x direction
    assert np.allclose(u, [0, 1, 0])
    rays = scenes.camera_rays_np(cam, 4, 2)
    assert np.allclose(np.linalg.norm(rays[:, 3:], axis=1), 1.0)


def test_csg_scene_encoding():
    s = scenes.csg_scene(n_objects=400)
    assert s.n_objects == 50
    # 3-level full binary trees: 7 nodes and 8 leaves per tree
    assert s.csg.shape[0] == 50 * 7
    leaves = s.csg_children[s.csg_children[:, 0] != A.PVX_CSG]
    assert leaves.shape[0] == 400
    counts = np.bincount(leaves[:, 0], minlength=4)
    assert abs(counts[A.PVX_BOX] - counts[A.PVX_SPHERE]) <= 1
    assert abs(counts[A.PVX_TORUS] - counts[A.PVX_SPHERE]) <= 1
    ops = set(s.csg[:, 0].tolist())
    assert ops <= {A.PVX_CSG_UNION, A.PVX_CSG_INTERSECTION, A.PVX_CSG_MERGE}
    # inverted children only come from differences, never in the first slot
    inv = (s.csg_children[:, 2] & A.PVX_FLAG_INVERTED) != 0
    assert inv.any()
    # every transform's inverse is its inverse
    for t in s.transforms[:20]:
        m, mi = t[:16].reshape(4, 4), t[16:].reshape(4, 4)
        assert np.allclose(m @ mi, np.eye(4), atol=1e-12)
    # torus minor radius in [0.1, 0.3] of the major radius
    ratio = s.torus_minor / s.torus_major
    assert ratio.min() >= 0.1 and ratio.max() <= 0.3


def test_scene_desc_points_at_the_arrays():
    s = scenes.sphere_scene(n=7)
    d, keep = s.desc()
    assert d.n_objects == 7 and d.n_spheres == 7 and d.n_meshes == 0
    assert d.sphere_radius[3] == s.sphere_radius[3]
    assert d.sphere_center[3 * 2 + 1] == s.sphere_center[2, 1]


def test_batch_plan_ranges_and_determinism():
    plan = scenes.batch_plan()
    assert len(plan) == 64 and plan == scenes.batch_plan()
    tris = np.array([p["n_tris"] for p in plan])
    assert tris.min() >= 1000 and tris.max() <= 500000
    # log-uniform: about half of the draws below the geometric mean
    assert 20 <= int(np.sum(tris < np.sqrt(1000 * 500000))) <= 44
    sph = np.array([p["n_spheres"] for p in plan])
    assert sph.min() >= 0 and sph.max() <= 2000
    w = np.array([p["width"] for p in plan])
    h = np.array([p["height"] for p in plan])
    assert w.min() >= 256 and w.max() <= 1920 and h.min() >= 256 and h.max() <= 1080
    assert len({p["seed"] for p in plan}) == 64


def test_terrain_mesh_counts_and_winding():
    for n in (1000, 4321, 20000):
        v, f, nrm = scenes.terrain_mesh(n, seed=5)
        assert abs(f.shape[0] - n) <= 0.02 * n + 4
        assert f.min() == 0 and f.max() == v.shape[0] - 1
        fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
        assert np.all(fn[:, 1] > 0)  # faces point up
        assert np.all(nrm[:, 1] > 0) and np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-6)
    s = scenes.batch_scene(seed=3, n_tris=2000, n_spheres=17, width=300, height=260)
    assert s.n_objects == 18 and s.objects[0, 0] == A.PVX_MESH
    assert np.all(s.objects[1:, 0] == A.PVX_SPHERE) and s.camera["width"] == 300
    s0 = scenes.batch_scene(seed=3, n_tris=2000, n_spheres=0, width=300, height=260)
    assert s0.n_objects == 1 and s0.sphere_center is None
