"""Multi-scene batch: one closest-hit launch over the rays of many scenes.

Per scene the batch must return exactly what the single-scene entry point
returns (same hit ids, bit-identical depths), and both must match the FP64
oracle to the closest-hit bar (ids identical, depth within 1e-5 relative).
Covers ragged ray counts (not multiples of 64), a scene with no rays, scenes
whose content needs different kernels (triangles/spheres, boxes, tori and
CSG in one batch), shadow rays with max_depth, and the per-scene clocks.
"""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from povray_amd import _abi as A
from povray_amd import scenes

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def _small_plan(seed, n):
    return scenes.batch_plan(seed=seed, n_scenes=n, tris=(1000, 20000), max_spheres=200,
                             res_lo=(17, 13), res_hi=(97, 71))


def _equal(a, b):
    return (np.array_equal(a["obj"], b["obj"]) and np.array_equal(a["prim"], b["prim"])
            and np.array_equal(a["t"], b["t"]))


def _vs_oracle(g, r, what):
    bad = (g["obj"] != r["obj"]) | (g["prim"] != r["prim"])
    assert not bad.any(), f"{what}: {int(bad.sum())} id mismatches"
    hit = r["obj"] >= 0
    if hit.any():
        rel = np.abs(g["t"][hit] - r["t"][hit]) / np.abs(r["t"][hit])
        assert rel.max() <= REL_TOL, f"{what}: depth rel err {rel.max():.3g}"


def test_batch_matches_single_scene_and_oracle(gpu):
    plan = _small_plan(11, 6)
    sc = scenes.batch_scenes(plan)
    # content that needs the box / torus / CSG kernels, and an empty ray set
    sc.append(scenes.csg_scene(n_objects=64, width=40, height=30))
    sc.append(scenes.sphere_scene(n=50, width=33, height=21))
    sc.append(scenes.radiosity_scene(n_boxes=40, width=31, height=19))
    gs = [gpu.GpuScene(s) for s in sc]
    rays = [scenes.camera_rays_np(s.camera) for s in sc]
    rays[7] = rays[7][:0]
    batch = gpu.GpuBatch(gs)
    got = batch.find_intersection(rays)
    for k, (s, g, r) in enumerate(zip(sc, gs, rays)):
        assert got[k].shape[0] == r.shape[0]
        if r.shape[0] == 0:
            continue
        single = g.find_intersection(r)
        assert _equal(got[k], single), f"scene {k} ({s.name}): batch differs from the single-scene call"
        _vs_oracle(got[k], OracleScene(s).find_intersection(r, method=1), f"scene {k} ({s.name})")
    print(f"[batch] {len(sc)} scenes, rays {[r.shape[0] for r in rays]}", flush=True)


def test_batch_triangle_scenes_shadow_rays(gpu):
    """Level-0 batch (while-while kernel for camera rays, the plain loop for
    shadow rays) and light rays bounded by max_depth."""
    sc = scenes.batch_scenes(_small_plan(23, 5))
    gs = [gpu.GpuScene(s) for s in sc]
    batch = gpu.GpuBatch(gs)
    cam = [scenes.camera_rays_np(s.camera) for s in sc]
    prim = batch.find_intersection(cam)
    srays, sdep = [], []
    for s, r, h in zip(sc, cam, prim):
        hit = h["obj"] >= 0
        p = r[hit, :3] + r[hit, 3:] * h["t"][hit, None]
        L = np.asarray(s.lights[0]["position"], np.float64)
        d = L[None, :] - p
        dist = np.linalg.norm(d, axis=1)
        srays.append(np.concatenate([p, d / dist[:, None]], 1))
        sdep.append(dist)
    got = batch.find_intersection(srays, ray_kind=A.PVX_RAY_SHADOW, max_depth_per_scene=sdep)
    for k, s in enumerate(sc):
        single = gs[k].find_intersection(srays[k], max_depth=sdep[k], ray_kind=A.PVX_RAY_SHADOW)
        assert _equal(got[k], single), f"scene {k}: shadow batch differs"
        ref = OracleScene(s).find_intersection(srays[k], max_depth=sdep[k], ray_kind=A.PVX_RAY_SHADOW, method=1)
        _vs_oracle(got[k], ref, f"shadow scene {k}")
    print(f"[batch-shadow] occluded fractions {[round(float(np.mean(g['obj'] >= 0)), 3) for g in got]}",
          flush=True)


def test_batch_clock_and_reuse(gpu):
    sc = scenes.batch_scenes(_small_plan(31, 4))
    batch = gpu.GpuBatch([gpu.GpuScene(s) for s in sc])
    rays = [scenes.camera_rays_np(s.camera) for s in sc]
    rays[2] = rays[2][:0]
    a, clk = batch.find_intersection(rays, with_clock=True)
    assert clk[0] > 0 and clk[3] == 0
    for k in (0, 1, 3):
        assert clk[1 + k] >= clk[0]
    assert gpu.clock_rate() > 0
    # same ranges again (cached chunk table), then different ranges
    b = batch.find_intersection(rays)
    rays[1] = rays[1][:77]
    c = batch.find_intersection(rays)
    for k in range(4):
        assert _equal(a[k], b[k])
    assert _equal(c[1], a[1][:77]) and _equal(c[3], a[3])


def test_batch_rejects_bad_ranges(gpu):
    from povray_amd import hip

    sc = scenes.batch_scenes(_small_plan(41, 2))
    batch = gpu.GpuBatch([gpu.GpuScene(s) for s in sc])
    d_rays = hip.DeviceBuffer(48 * 64)
    d_hits = hip.DeviceBuffer(16 * 64)
    with pytest.raises(gpu.PvxError):
        batch.trace_closest(d_rays, [0, 10, 5], d_hits)
    with pytest.raises(gpu.PvxError):
        batch.trace_closest(d_rays, [1, 10, 20], d_hits)
    with pytest.raises(ValueError):
        batch.trace_closest(d_rays, [0, 10], d_hits)
