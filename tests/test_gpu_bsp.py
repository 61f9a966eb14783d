"""The kd-tree acceleration structure (PVX_ACCEL_BSP, the reference's +BM2
BSPTree): every closest hit must be the one the BVH finds — same object,
primitive and depth — and the FP64 oracle's."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from povray_amd import _abi as A
from povray_amd import scenes

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    bad = (a["obj"] != b["obj"]) | (a["prim"] != b["prim"])
    assert not bad.any(), f"{what}: {int(bad.sum())} id mismatches, first {np.nonzero(bad)[0][:4]}"
    assert np.array_equal(a["t"], b["t"]), f"{what}: depths differ"


@pytest.mark.parametrize("name", ["spheres", "mesh", "csg", "triangles", "room"])
def test_bsp_matches_bvh_and_oracle(gpu, name):
    s = {"spheres": lambda: scenes.sphere_scene(width=200, height=150),
         "mesh": lambda: scenes.mesh_scene(level=5, width=240, height=135),
         "csg": lambda: scenes.csg_scene(n_objects=400, width=160, height=90),
         "triangles": lambda: scenes.triangle_scene(width=160, height=120),
         "room": lambda: scenes.radiosity_scene(n_boxes=300, width=160, height=90)}[name]()
    rays = scenes.camera_rays_np(s.camera)
    bvh = gpu.GpuScene(s)
    bsp = gpu.GpuScene(s, accel=A.PVX_ACCEL_BSP)
    st = bsp.stats()
    hb, hk = bvh.find_intersection(rays), bsp.find_intersection(rays)
    print(f"[bsp {name}] kd nodes {st['n_nodes']} depth {st['bvh_depth']} list {st['n_prims']}; "
          f"hit frac {np.mean(hk['obj'] >= 0):.3f}", flush=True)
    assert st["node_bytes"] == 8 and st["bvh_depth"] <= 64
    _same(hk, hb, f"{name}: BSP vs BVH")
    _same(hk, OracleScene(s).find_intersection(rays, method=1), f"{name}: BSP vs oracle")
    # light rays bounded by max_depth, through the same trees
    md = np.random.default_rng(1).uniform(3.0, 30.0, rays.shape[0])
    _same(bsp.find_intersection(rays, md, A.PVX_RAY_SHADOW), bvh.find_intersection(rays, md, A.PVX_RAY_SHADOW),
          f"{name}: BSP vs BVH, shadow")


def test_bsp_depth_limit_and_flags(gpu):
    s = scenes.sphere_scene(n=500, width=120, height=90)
    s.objects[:, 2] = np.random.default_rng(4).integers(0, 16, s.n_objects).astype(np.uint32)
    rays = scenes.camera_rays_np(s.camera)
    bvh = gpu.GpuScene(s)
    shallow = gpu.GpuScene(s, accel=A.PVX_ACCEL_BSP, bsp_max_depth=3)
    assert shallow.stats()["bvh_depth"] <= 3
    for kind in (A.PVX_RAY_IMAGE, A.PVX_RAY_REFLECTION, A.PVX_RAY_SHADOW):
        _same(shallow.find_intersection(rays, None, kind), bvh.find_intersection(rays, None, kind),
              f"depth-3 BSP kind {kind}")
