"""Light rays made in the traversal kernel (pvx_render on triangle / sphere
scenes whose materials are all opaque): from the shaded points and the
light's grid, tested for occlusion only, with the 4-wide tree walked per lane
(PVX_TRAVERSAL_FUSED, the default; PVX_TRAVERSAL_REFILL, lanes taking new
rays as theirs end) or by each wavefront as one packet with a shared stack
(PVX_TRAVERSAL_PACKET).  Both give the visibility of the light
rays written out and traced one by one (a per-ray walk chosen with
pvx_scene_set_traversal), so the frames are identical bit for bit, and they
match the oracle."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from povray_amd import _abi as A
from povray_amd import scenes

pytestmark = pytest.mark.gpu


def _q(a):
    return np.round(np.clip(a, 0.0, 1.0) * 255.0).astype(np.int32)


@pytest.mark.parametrize("name", ["mesh_area_lights", "spheres", "mesh_flags"])
def test_packet_light_rays_equal_per_ray_walk(gpu, name):
    if name == "spheres":
        s = scenes.sphere_scene(width=320, height=240)
    else:
        s = scenes.mesh_scene(level=5, width=320, height=180)
    if name == "mesh_flags":  # the ground plane casts no shadow: the precondition path
        s.objects = s.objects.copy()
        s.objects[-1, 2] |= A.PVX_FLAG_NO_SHADOW
    gs = gpu.GpuScene(s)
    gs.set_shading()
    img, st = gs.render()
    _, st2, tune = gs.render_autotune()
    print(f"[packet {name}] light tune {tune[1]}; shadow {st['shadow_rays']} rays in {st['shadow_ms']:.3f} ms",
          flush=True)
    assert tune[1]["variant"] in (A.PVX_TRAVERSAL_PACKET, A.PVX_TRAVERSAL_FUSED, A.PVX_TRAVERSAL_REFILL,
                                  A.PVX_TRAVERSAL_LIGHT_BINARY, A.PVX_TRAVERSAL_REFILL_BINARY)
    frames = {}
    for v in (A.PVX_TRAVERSAL_PACKET, A.PVX_TRAVERSAL_FUSED, A.PVX_TRAVERSAL_REFILL, A.PVX_TRAVERSAL_LIGHT_BINARY,
              A.PVX_TRAVERSAL_REFILL_BINARY, A.PVX_TRAVERSAL_WIDE_WW):
        gs.set_traversal(v, 0, A.PVX_RAY_SHADOW)
        frames[v] = gs.render()
        print(f"[packet {name}] walk {v}: light rays {frames[v][1]['shadow_ms']:.3f} ms", flush=True)
    for v, (f, stv) in frames.items():
        assert stv["shadow_rays"] == st["shadow_rays"] and stv["closest_rays"] == st["closest_rays"]
        assert np.array_equal(f, img), v
    orc, _ = OracleScene(s).render(s)
    ok = float(np.mean(np.abs(_q(img) - _q(orc)).max(-1) <= 1))
    print(f"[packet {name}] {ok * 100:.4f}% of pixels within 1/255 of the oracle", flush=True)
    assert ok >= 0.999
