"""Camera rays in tile order (pvx_camera_rays_tiled): the same rays as the
row-major pvx_camera_rays, bit for bit, each at the pixel the call reports;
the pixel indices are a permutation of the image, the tiles come row by row
of tiles and, inside a tile, pixels row by row; and a shaded frame (whose
ray queue uses 8 x 8 tiles) still matches the oracle."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from povray_amd import hip, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size,tile", [((1920, 1080), (8, 8)), ((37, 23), (8, 8)), ((37, 23), (3, 5)),
                                       ((8, 8), (8, 8)), ((5, 3), (16, 4)), ((64, 1), (8, 8))])
def test_tiled_rays_are_the_row_major_rays_permuted(gpu, size, tile):
    W, H = size
    tw, th = tile
    s = scenes.sphere_scene(width=W, height=H)
    st = hip.Stream()
    rm, n = gpu.camera_rays(s, stream=st)
    pix = hip.DeviceBuffer(n * 4)
    tl, _ = gpu.camera_rays(s, stream=st, tile=tile, pixel=pix)
    st.synchronize()
    a = rm.download(np.float64, n * 6).reshape(n, 6)
    b = tl.download(np.float64, n * 6).reshape(n, 6)
    p = pix.download(np.uint32, n).astype(np.int64)
    print(f"[tiles] {W}x{H} tiles {tw}x{th}", flush=True)
    assert np.array_equal(np.sort(p), np.arange(n))
    assert np.array_equal(b, a[p])
    x, y = p % W, p // W
    ty, tx = y // th, x // tw
    key = (ty * ((W + tw - 1) // tw) + tx) * (tw * th) + (y % th) * tw + (x % tw)
    assert np.all(np.diff(key) > 0)  # tile rows, tiles, then rows inside a tile, in order


def test_shaded_frame_in_tile_order_matches_oracle(gpu):
    s = scenes.mesh_scene(level=4, width=203, height=117)  # partial tiles on both edges
    gs = gpu.GpuScene(s)
    gs.set_shading()
    img, _ = gs.render()
    orc, _ = OracleScene(s).render(s)
    q = lambda a: np.round(np.clip(a, 0.0, 1.0) * 255.0).astype(np.int32)
    ok = float(np.mean(np.abs(q(img) - q(orc)).max(-1) <= 1))
    print(f"[tiles] frame: {ok * 100:.4f}% of pixels within 1/255 of the oracle", flush=True)
    assert ok >= 0.999


def test_render_tile_choice_leaves_the_frame_unchanged(gpu):
    """pvx_scene_set_render_tile: every pixel order of the ray queue gives
    the same frame, bit for bit; pvx_render_autotune reports the tile it
    keeps (one of its candidates) and renders that same frame."""
    s = scenes.mesh_scene(level=4, width=203, height=117)
    gs = gpu.GpuScene(s)
    gs.set_shading()
    ref, st = gs.render()
    for tw, th in ((16, 4), (0, 1), (3, 5), (203, 117), (8, 8)):
        gs.set_render_tile(tw, th)
        img, st2 = gs.render()
        print(f"[tiles] queue tile {tw}x{th}: closest {st2['closest_rays']} light {st2['shadow_rays']}", flush=True)
        assert np.array_equal(img, ref), (tw, th)
        assert st2["closest_rays"] == st["closest_rays"] and st2["shadow_rays"] == st["shadow_rays"]
    for bad in ((0, 0), (0, 2), (4, 0), (5000, 1)):
        with pytest.raises(gpu.PvxError):
            gs.set_render_tile(*bad)
    img, _, tune = gs.render_autotune()
    print(f"[tiles] autotune kept tile {tune[0]['tile_w']}x{tune[0]['tile_h']}, walk {tune[0]['variant']} "
          f"grid {tune[0]['grid_blocks']}", flush=True)
    assert (tune[0]["tile_w"], tune[0]["tile_h"]) in ((8, 8), (16, 4), (32, 2), (0, 1))
    assert np.array_equal(img, ref)
    assert np.array_equal(gs.render()[0], ref)
