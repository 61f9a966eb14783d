"""The radiosity boundary: pvx_radiosity_begin / pvx_radiosity_compute_ambient
against the oracle's sequential RadiosityFunction::ComputeAmbient
(core/lighting/radiosity.cpp:385-461) over points the caller chooses.

With count 1600 every gather uses each direction of the Halton pool once,
and with recursion_limit 1 and a vanishing error bound no point reuses
another's record, so GPU and oracle must agree to rounding whatever order
the points are gathered in.  Otherwise the cache depends on the gather
order and the two are compared statistically.
"""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from povray_amd import _abi as A
from povray_amd import scenes

pytestmark = pytest.mark.gpu


def _room(n_boxes=100):
    s = scenes.radiosity_scene(n_boxes=n_boxes, width=64, height=36)
    s.materials["pigment"] = np.maximum(s.materials["pigment"], 0.9)
    return s


def _points(n, seed=1):
    """Floor points (y = 0, normal +y) and left-wall points (x = -5, normal +x)."""
    rng = np.random.default_rng(seed)
    nf = n - n // 4
    floor = np.c_[rng.uniform(-4.8, 4.8, nf), np.zeros(nf), rng.uniform(-4.8, 4.8, nf)]
    nw = n - nf
    wall = np.c_[np.full(nw, -5.0), rng.uniform(0.2, 3.8, nw), rng.uniform(-4.8, 4.8, nw)]
    p = np.concatenate([floor, wall])
    nrm = np.concatenate([np.tile([0.0, 1.0, 0.0], (nf, 1)), np.tile([1.0, 0.0, 0.0], (nw, 1))])
    return p, nrm


def _gscene(gpu, s):
    g = gpu.GpuScene(s)
    g.set_shading()
    return g


def test_order_free_ambient_matches_oracle(gpu):
    s = _room()
    over = dict(count=1600, recursion_limit=1, error_bound=1e-6)
    p, nrm = _points(300)
    w = np.linspace(0.3, 1.0, p.shape[0])
    ref_f = OracleScene(s).radiosity_function(s, **over)
    ref = ref_f.compute_ambient(p, nrm, w)
    f = _gscene(gpu, s).radiosity_function(**over)
    got = f.compute_ambient(p, nrm, nrm, w)
    st = f.last_stats
    print(f"[ambient-pin] oracle {ref_f.counters} gpu gathers={st['gathers']} rays={st['gather_rays']}",
          flush=True)
    err = np.abs(got - ref)
    print(f"[ambient-pin] max abs err {err.max():.3g} mean {ref.mean():.4f}", flush=True)
    assert ref.mean() > 0.05
    assert err.max() <= 1e-5 * max(1.0, float(ref.max()))
    assert st["gathers"][0] == p.shape[0] == ref_f.counters["gathers0"]


def test_cache_persists_between_calls(gpu):
    """Records filed by one call serve the next: the same points asked again
    in the final trace are all answered from the cache."""
    s = _room(300)
    p, nrm = _points(400, seed=2)
    f = _gscene(gpu, s).radiosity_function(count=100)
    first = f.compute_ambient(p, nrm, final=False)
    s1 = f.last_stats
    again = f.compute_ambient(p, nrm, final=True)
    s2 = f.last_stats
    print(f"[ambient-cache] first gathers={s1['gathers']} records={s1['records']}; "
          f"again gathers={s2['gathers']}", flush=True)
    assert 0 < s1["gathers"][0] <= p.shape[0]
    # only the points whose gather filed no record (a hit closer than
    # maximum_reuse * 1e-4, GatherLight's filing test) gather again
    assert s2["gathers"][0] == s1["gathers"][0] - s1["records"][0]
    assert np.all(np.isfinite(again)) and abs(float(again.mean() / first.mean()) - 1.0) < 0.05
    # a depth-1 query (inside a depth-0 gather) is served by the depth-1 cache
    d1 = f.compute_ambient(p[:50], nrm[:50], depth=1, final=True)
    assert np.all(np.isfinite(d1)) and f.last_stats["gathers"][0] == 0


def test_ambient_statistically_matches_oracle(gpu):
    """Default reuse (error_bound 0.5): the two caches hold different records
    (gather order), so the answers agree in the mean and to within the
    oracle's own order noise."""
    s = _room(300)
    over = dict(count=50)
    p, nrm = _points(2000, seed=3)
    o = OracleScene(s)
    ref = o.radiosity_function(s, **over).compute_ambient(p, nrm, final=True)
    perm = np.random.default_rng(4).permutation(p.shape[0])
    alt = np.empty_like(ref)
    alt[perm] = o.radiosity_function(s, **over).compute_ambient(p[perm], nrm[perm], final=True)
    noise = float(np.sqrt(np.mean((alt - ref) ** 2)))
    got = _gscene(gpu, s).radiosity_function(**over).compute_ambient(p, nrm, final=True)
    rmse = float(np.sqrt(np.mean((got - ref) ** 2)))
    rel = float(got.mean() / ref.mean() - 1.0)
    print(f"[ambient-stat] rmse {rmse:.4f} oracle order noise {noise:.4f} mean-rel {rel:+.4f}", flush=True)
    assert rmse <= 1.5 * noise + 1e-3
    assert abs(rel) < 0.04


def test_ambient_errors(gpu):
    s = _room(20)
    g = _gscene(gpu, s)
    p, nrm = _points(8)
    f = g.radiosity_function(recursion_limit=1)
    with pytest.raises(gpu.PvxError, match="recursion_limit"):
        f.compute_ambient(p, nrm, depth=1)
    bent = nrm + np.array([0.0, 0.0, 0.1])
    with pytest.raises(gpu.PvxError, match="perturbed"):
        f.compute_ambient(p, nrm, layer_normal=bent)
    # a radiosity frame closes the per-query cache
    g.radiosity_render(width=16, height=9, count=20)
    with pytest.raises(gpu.PvxError, match="pvx_radiosity_begin"):
        f.compute_ambient(p, nrm)
    # the surrogate works per frame: per-query ambient takes the plain gather
    with pytest.raises(gpu.PvxError, match="plain"):
        _begin_raw(gpu, g, s, A.radiosity_settings(s, mode=A.PVX_RADIOSITY_LEARNED))


def _begin_raw(gpu, g, s, rs):
    import ctypes as C

    cam = s.camera_struct()
    gpu._check(gpu.lib().pvx_radiosity_begin(g.h, C.byref(cam), C.byref(rs), None), "pvx_radiosity_begin")
