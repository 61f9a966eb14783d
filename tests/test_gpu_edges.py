"""Closest-hit edge cases against the FP64 oracle: scenes without a tree,
axis-aligned rays (zero direction components), rays that start inside
objects, ragged ray counts, max_depth at and below EPSILON, and zero-area
mesh triangles."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from povray_amd import _abi as A
from povray_amd import scenes

pytestmark = pytest.mark.gpu


def _same(g, r, what):
    bad = (g["obj"] != r["obj"]) | (g["prim"] != r["prim"])
    assert not bad.any(), f"{what}: {int(bad.sum())} id mismatches, first {np.nonzero(bad)[0][:4]}"
    hit = r["obj"] >= 0
    assert np.array_equal(g["t"][hit], r["t"][hit]), f"{what}: depths differ"
    assert np.all(g["t"][~hit] == r["t"][~hit]), f"{what}: miss depths differ"
    print(f"[{what}] n={len(r)} hits={int(hit.sum())}", flush=True)


def _rays(n, seed, origin_box=8.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-origin_box, origin_box, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1)


def test_scene_without_objects_and_planes_only(gpu):
    empty = scenes.Scene(name="empty", objects=np.zeros((0, 4), np.uint32))
    r = _rays(300, 1)
    g = gpu.GpuScene(empty).find_intersection(r)
    assert np.all(g["obj"] == -1)
    planes = scenes.Scene(name="planes", objects=np.array([[A.PVX_PLANE, 0, 0, 0], [A.PVX_PLANE, 1, 0, 0]],
                                                          np.uint32),
                          plane_normal=np.array([[0.0, 1.0, 0.0], [0.6, 0.0, 0.8]]),
                          plane_distance=np.array([2.0, -3.0]))
    _same(gpu.GpuScene(planes).find_intersection(r), OracleScene(planes).find_intersection(r), "planes only")


def test_axis_aligned_and_inside_rays(gpu):
    s = scenes.sphere_scene(n=300)
    rng = np.random.default_rng(7)
    n = 4001  # not a multiple of 64
    o = rng.uniform(-10, 10, (n, 3))
    axes = np.eye(3)[rng.integers(0, 3, n)] * rng.choice([-1.0, 1.0], (n, 1))
    ax = np.concatenate([o, axes], 1)
    # origins at sphere centres: every such ray starts inside its sphere
    c = s.sphere_center[rng.integers(0, 300, n)]
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    inside = np.concatenate([c, d], 1)
    gs, o_ = gpu.GpuScene(s), OracleScene(s)
    _same(gs.find_intersection(ax), o_.find_intersection(ax), "axis-aligned rays")
    _same(gs.find_intersection(inside), o_.find_intersection(inside), "rays from inside spheres")
    room = scenes.radiosity_scene(n_boxes=200, width=8, height=8)
    cb = np.array([[0.0, 1.0, 0.0]]) + rng.uniform(-0.01, 0.01, (500, 3))
    rb = np.concatenate([cb, d[:500]], 1)
    _same(gpu.GpuScene(room).find_intersection(rb), OracleScene(room).find_intersection(rb), "rays in the room")


def test_ragged_counts_and_max_depth_edges(gpu):
    s = scenes.mesh_scene(level=4, width=64, height=48)
    gs, o = gpu.GpuScene(s), OracleScene(s)
    full = scenes.camera_rays_np(s.camera)
    for n in (1, 63, 64, 65, 257, 3071):
        _same(gs.find_intersection(full[:n]), o.find_intersection(full[:n]), f"ragged n={n}")
    n = full.shape[0]
    md = np.where(np.arange(n) % 3 == 0, 0.0, np.where(np.arange(n) % 3 == 1, 0.5e-10, 1.5))
    for kind in (A.PVX_RAY_IMAGE, A.PVX_RAY_SHADOW):
        _same(gs.find_intersection(full, md, kind), o.find_intersection(full, md, kind), f"max_depth edges {kind}")


def test_zero_area_mesh_triangles(gpu):
    v, f, nrm = scenes.displaced_icosphere(3, seed=9)
    # collapse every 7th triangle onto an edge (zero area) and every 11th to a point
    v = np.concatenate([v, v[f[::7, 0]] * 0.5 + v[f[::7, 1]] * 0.5], 0).astype(np.float32)
    f = f.copy()
    k7 = np.arange(0, f.shape[0], 7)
    f[k7, 2] = np.arange(len(v) - len(k7), len(v))
    f[::11, 1] = f[::11, 0]
    f[::11, 2] = f[::11, 0]
    s = scenes.Scene(name="degenerate_mesh", objects=np.array([[A.PVX_MESH, 0, 0, 0]], np.uint32),
                     meshes=[(v, f.astype(np.int32), None)],
                     camera=scenes.look_at_camera((0.0, 0.3, -3.0), (0.0, 0.0, 0.0), 160, 120))
    r = scenes.camera_rays_np(s.camera)
    _same(gpu.GpuScene(s).find_intersection(r), OracleScene(s).find_intersection(r), "zero-area triangles")


def test_overlapping_launches_on_two_streams(gpu):
    """Two streams tracing the same scene at once (each with its own stack
    spill area) give the single-stream results."""
    from povray_amd import hip

    s = scenes.mesh_scene(level=6, width=640, height=360)
    gs = gpu.GpuScene(s)
    rays = scenes.camera_rays_np(s.camera)
    n = rays.shape[0]
    ref = gs.find_intersection(rays)
    d_r = hip.DeviceBuffer.from_array(rays)
    outs = [hip.DeviceBuffer(n * 16) for _ in range(2)]
    sts = [hip.Stream(), hip.Stream()]
    for _ in range(3):
        for k in range(2):
            gs.trace_closest(d_r, n, outs[k], stream=sts[k])
    for st in sts:
        st.synchronize()
    for k in range(2):
        got = outs[k].download(np.uint8, n * 16).view(gpu.HIT_DTYPE)
        assert np.array_equal(got, ref)
