"""The traversal autotuner (pvx_scene_autotune / pvx_scene_set_traversal):
every variant and grid a scene admits gives the same hits, bit for bit, and a
tuned scene traces exactly as the untuned one."""
import numpy as np
import pytest

from povray_amd import _abi as A
from povray_amd import pvx, scenes

pytestmark = pytest.mark.gpu

V_ALL = (A.PVX_TRAVERSAL_WIDE_WW, A.PVX_TRAVERSAL_WIDE_STEP, A.PVX_TRAVERSAL_BINARY, A.PVX_TRAVERSAL_BINARY_WW)


def _same(a, b, what):
    bad = (a["obj"] != b["obj"]) | (a["prim"] != b["prim"])
    assert not bad.any(), f"{what}: {int(bad.sum())} id mismatches, first {np.nonzero(bad)[0][:4]}"
    assert np.array_equal(a["t"], b["t"]), f"{what}: depths differ"


def _light_rays(s, rays, seed=3):
    return np.random.default_rng(seed).uniform(2.0, 20.0, rays.shape[0])


def test_every_variant_same_hits(gpu):
    s = scenes.mesh_scene(level=6, width=320, height=180)
    rays = scenes.camera_rays_np(s.camera)
    md = _light_rays(s, rays)
    gs = gpu.GpuScene(s)
    ref_img = gs.find_intersection(rays)
    ref_sh = gs.find_intersection(rays, md, A.PVX_RAY_SHADOW)
    for v in V_ALL:
        for g in (0, 1, 37):
            print(f"[autotune] variant {v} grid {g}", flush=True)
            gs.set_traversal(v, g)
            gs.set_traversal(v, g, ray_kind=A.PVX_RAY_SHADOW)
            _same(gs.find_intersection(rays), ref_img, f"variant {v} grid {g}")
            _same(gs.find_intersection(rays, md, A.PVX_RAY_SHADOW), ref_sh, f"variant {v} grid {g} shadow")
    gs.set_traversal()
    gs.set_traversal(ray_kind=A.PVX_RAY_SHADOW)


def test_autotune_keeps_hits(gpu):
    s = scenes.mesh_scene(level=7, width=480, height=270)
    rays = scenes.camera_rays_np(s.camera)
    md = _light_rays(s, rays)
    gs = gpu.GpuScene(s)
    ref_img = gs.find_intersection(rays)
    ref_sh = gs.find_intersection(rays, md, A.PVX_RAY_SHADOW)
    r = gs.autotune(rays)
    rs = gs.autotune(rays, ray_kind=A.PVX_RAY_SHADOW)
    print(f"[autotune] camera {r}\n[autotune] light {rs}", flush=True)
    for t in (r, rs):
        assert t["variant"] in V_ALL
        assert t["grid_blocks"] > 0 and t["best_ms"] > 0 and t["default_ms"] > 0
        # the default is among the timed candidates, so the pick is no slower
        # than it up to timing noise
        assert t["best_ms"] <= t["default_ms"] * 1.5
    _same(gs.find_intersection(rays), ref_img, "tuned camera rays")
    _same(gs.find_intersection(rays, md, A.PVX_RAY_SHADOW), ref_sh, "tuned light rays")


@pytest.mark.parametrize("name", ["room", "csg"])
def test_autotune_other_scene_kinds(gpu, name):
    s = {"room": lambda: scenes.radiosity_scene(n_boxes=300, width=160, height=90),
         "csg": lambda: scenes.csg_scene(n_objects=400, width=160, height=90)}[name]()
    rays = scenes.camera_rays_np(s.camera)
    gs = gpu.GpuScene(s)
    ref = gs.find_intersection(rays)
    r = gs.autotune(rays)
    assert r["variant"] == (A.PVX_TRAVERSAL_BINARY if name == "room" else A.PVX_TRAVERSAL_DEFAULT)
    _same(gs.find_intersection(rays), ref, f"{name} tuned")
    # the 4-wide walks are refused where the scene has no 4-wide tree
    with pytest.raises(pvx.PvxError):
        gs.set_traversal(A.PVX_TRAVERSAL_WIDE_WW)


def test_autotune_bad_arguments(gpu):
    s = scenes.sphere_scene(n=50, width=32, height=24)
    gs = gpu.GpuScene(s)
    with pytest.raises(pvx.PvxError):
        gs.set_traversal(A.PVX_TRAVERSAL_WIDE_WW, 1 << 30)
    with pytest.raises(pvx.PvxError):
        gs.set_traversal(99)
    with pytest.raises(pvx.PvxError):
        gs.autotune(np.zeros((0, 6)))


@pytest.mark.parametrize("name", ["mesh", "spheres"])
def test_render_autotune_same_image(gpu, name):
    """pvx_render_autotune tunes on the frame's own camera and light rays;
    the frame, and every later frame, is the untuned one bit for bit."""
    s = {"mesh": lambda: scenes.mesh_scene(level=5, width=320, height=180),
         "spheres": lambda: scenes.sphere_scene(width=320, height=240)}[name]()
    gs = gpu.GpuScene(s)
    gs.set_shading()
    ref, st0 = gs.render()
    img, st, tune = gs.render_autotune()
    print(f"[render autotune {name}] camera {tune[0]} light {tune[1]}", flush=True)
    assert np.array_equal(img, ref)
    assert st["closest_rays"] == st0["closest_rays"] and st["shadow_rays"] == st0["shadow_rays"]
    assert tune[0]["variant"] in V_ALL and tune[0]["best_ms"] > 0
    # light rays of these opaque triangle / sphere scenes go in packets
    assert tune[1]["variant"] in (A.PVX_TRAVERSAL_PACKET, A.PVX_TRAVERSAL_FUSED, A.PVX_TRAVERSAL_REFILL,
                                  A.PVX_TRAVERSAL_LIGHT_BINARY, A.PVX_TRAVERSAL_REFILL_BINARY) and tune[1]["best_ms"] > 0
    again, _ = gs.render()
    assert np.array_equal(again, ref)
