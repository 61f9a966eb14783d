"""Radiosity frames of a scene batch (pvx_batch_radiosity_render): each
scene's image and statistics equal its own pvx_radiosity_render frame
(plain and learned modes are deterministic), whatever stream it ran on;
one scene's plain frame agrees with the FP64 oracle within the oracle's own
gather-order noise."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from povray_amd import _abi as A
from povray_amd import scenes

pytestmark = pytest.mark.gpu

PLAN = [dict(seed=11, n_tris=2000, n_spheres=40, width=96, height=64),
        dict(seed=12, n_tris=9000, n_spheres=0, width=128, height=96),
        dict(seed=13, n_tris=1200, n_spheres=300, width=64, height=64),
        dict(seed=14, n_tris=20000, n_spheres=120, width=160, height=90),
        dict(seed=15, n_tris=4000, n_spheres=800, width=80, height=120)]


def _scenes():
    return [scenes.batch_radiosity_scene(p["seed"], p["n_tris"], p["n_spheres"], p["width"], p["height"],
                                         name=f"rb{k}") for k, p in enumerate(PLAN)]


def _gpu(gpu, sc):
    g = gpu.GpuScene(sc)
    g.set_shading()
    return g


@pytest.mark.parametrize("mode", [A.PVX_RADIOSITY_PLAIN, A.PVX_RADIOSITY_LEARNED])
def test_batch_frames_equal_per_scene_frames(gpu, mode):
    scs = _scenes()
    gs = [_gpu(gpu, s) for s in scs]
    batch = gpu.GpuBatch(gs)
    imgs, sst, done, bst = batch.radiosity_render(mode=mode, n_streams=3)
    print(f"[batch-rad mode {mode}] {bst} done {np.round(done, 1)}", flush=True)
    assert bst["streams"] == 3 and bst["pixels"] == sum(p["width"] * p["height"] for p in PLAN)
    for k, g in enumerate(gs):
        img, st = g.radiosity_render(mode=mode)
        assert np.array_equal(img, imgs[k]), f"scene {k}"
        assert st["gathers"] == sst[k]["gathers"] and st["gather_rays"] == sst[k]["gather_rays"]
        assert img.mean() > 0.02
    if mode == A.PVX_RADIOSITY_LEARNED:
        assert bst["surrogate_queries"] > 0
        assert bst["fallbacks"] <= bst["final_misses"]


def test_batch_scene_plain_frame_vs_oracle(gpu):
    s = _scenes()[0]
    over = dict(count=50)
    o = OracleScene(s)
    ref, _ = o.radiosity_render(s, **over)
    alt, _ = o.radiosity_render(s, pretrace_start=0.09, **over)
    floor = float(np.sqrt(np.mean((np.clip(alt, 0, 1) - np.clip(ref, 0, 1)) ** 2)))
    img, _ = _gpu(gpu, s).radiosity_render(**over)
    rmse = float(np.sqrt(np.mean((np.clip(img, 0, 1) - np.clip(ref, 0, 1)) ** 2)))
    print(f"[batch-rad oracle] rmse {rmse * 255:.2f}/255, oracle order noise {floor * 255:.2f}/255", flush=True)
    assert rmse <= 1.5 * floor + 1.0 / 255
    assert abs(float(img.mean() / ref.mean()) - 1.0) < 0.05


def test_batch_radiosity_rejects_bad_arguments(gpu):
    scs = _scenes()[:2]
    batch = gpu.GpuBatch([_gpu(gpu, s) for s in scs])
    with pytest.raises(gpu.PvxError, match="n_streams"):
        batch.radiosity_render(n_streams=0)
