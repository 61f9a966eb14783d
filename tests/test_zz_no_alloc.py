"""pvx_trace_closest_noalloc against pvx_trace_closest, through the C ABI:
identical hits on two calls in a row, the second with a smaller input (the
shared scratch array keeps the first call's contents)."""
import numpy as np
import pytest

from povray_amd import _abi as A
from povray_amd import hip, scenes

pytestmark = pytest.mark.gpu


def _trace(gpu, g, entry, rays, max_depth, kind):
    n = rays.shape[0]
    d_r = hip.DeviceBuffer.from_array(rays)
    d_h = hip.DeviceBuffer(n * 16)
    d_m = hip.DeviceBuffer.from_array(max_depth) if max_depth is not None else None
    fn = getattr(gpu.lib(), entry)
    gpu._check(fn(g.h, d_r.value, d_m.value if d_m else None, kind, n, d_h.value, None), entry)
    hip.synchronize()
    return d_h.download(np.uint8, n * 16).view(gpu.HIT_DTYPE).copy()


@pytest.mark.parametrize("which", ["mesh", "csg"])
def test_noalloc_matches_trace_closest(gpu, which):
    if which == "mesh":
        s = scenes.mesh_scene(level=6, width=640, height=360)
    else:
        s = scenes.csg_scene(n_objects=400, width=320, height=180)
    print(f"[noalloc] {s.name}", flush=True)
    g = gpu.GpuScene(s)
    rays = scenes.camera_rays_np(s.camera)
    md = np.random.default_rng(1).uniform(5.0, 60.0, rays.shape[0])
    for k, (sel, kind) in enumerate([(slice(None), A.PVX_RAY_IMAGE), (slice(0, rays.shape[0] // 3), A.PVX_RAY_SHADOW)]):
        r = np.ascontiguousarray(rays[sel])
        m = np.ascontiguousarray(md[sel]) if k else None
        want = _trace(gpu, g, "pvx_trace_closest", r, m, kind)
        got = _trace(gpu, g, "pvx_trace_closest_noalloc", r, m, kind)
        print(f"[noalloc] {s.name} call {k} rays={r.shape[0]} hit-frac={np.mean(want['obj'] >= 0):.3f}", flush=True)
        assert np.array_equal(got["obj"], want["obj"]) and np.array_equal(got["prim"], want["prim"])
        assert np.array_equal(got["t"], want["t"])


def test_noalloc_rejects_bad_arguments(gpu):
    s = scenes.sphere_scene(n=10, width=8, height=8)
    g = gpu.GpuScene(s)
    lib = gpu.lib()
    assert lib.pvx_trace_closest_noalloc(None, None, None, 0, 1, None, None) == A.PVX_EINVAL
    assert lib.pvx_trace_closest_noalloc(g.h, None, None, 0, 1, None, None) == A.PVX_EINVAL
    assert lib.pvx_trace_closest_noalloc(g.h, None, None, 1 << 9, 1, None, None) == A.PVX_EINVAL
    assert lib.pvx_trace_closest_noalloc(g.h, None, None, 0, 0, None, None) == A.PVX_OK
