"""Radiosity cache files (.rca; RadiosityCache::Load / ot_write_block,
core/lighting/radiosity.cpp:909-1000, core/support/octree.cpp:1345-1370):
the file holds one line per irradiance record in the reference's layout, a
loaded file seeds later frames so that they gather far less and render the
same picture up to the file's rounding, and foreign lines are skipped."""
import re

import numpy as np
import pytest

from povray_amd import scenes

pytestmark = pytest.mark.gpu

LINE = re.compile(r"^C([1-4])\t(\S+)\t(\S+)\t(\S+)\t([0-9a-f]{6})\t(\d+\.\d{4})\t(\d+\.\d{4})\t(\d+\.\d{4})\t(\S+)\t(\S+)"
                  r"\t([0-9a-f]{6})$")


def _room(gpu):
    s = scenes.radiosity_scene(n_boxes=300, width=240, height=135)
    g = gpu.GpuScene(s)
    g.set_shading()
    return s, g


def test_rca_round_trip(gpu, tmp_path):
    s, g = _room(gpu)
    ref, st = g.radiosity_render(count=60)
    path = tmp_path / "room.rca"
    n = g.radiosity_save(path)
    assert n == sum(st["records"]) and n > 100
    lines = path.read_text().splitlines()
    assert len(lines) == n
    depths = []
    for ln in lines:
        m = LINE.match(ln)
        assert m, ln
        depths.append(int(m.group(1)))
        assert float(m.group(9)) > 0  # harmonic mean distance
    assert set(depths) <= {1, 2}  # recursion_limit 2: bounce depths 1 and 2

    # a fresh scene seeded from the file gathers far less, same picture
    g2 = gpu.GpuScene(s)
    g2.set_shading()
    assert g2.radiosity_load(path) == n
    img, st2 = g2.radiosity_render(count=60)
    print(f"[rca] {n} records; gathers {st['gathers']} -> {st2['gathers']}", flush=True)
    assert sum(st2["gathers"]) * 4 < sum(st["gathers"])
    # Different gathers mean different Monte Carlo noise (two plain frames of
    # this count differ by ~6/255 RMS), so the seeded frame is compared with
    # a converged frame: it must be as close to it as the frame that wrote the
    # file, and as bright.
    conv, _ = g.radiosity_render(count=1600, error_bound=0.2, nearest_count=10)
    e_ref = np.sqrt(np.mean((ref - conv) ** 2)) * 255
    e_img = np.sqrt(np.mean((img - conv) ** 2)) * 255
    print(f"[rca] RMSE vs converged: writer {e_ref:.2f}/255, seeded {e_img:.2f}/255", flush=True)
    assert e_img <= e_ref + 1.0
    assert abs(img.mean() / ref.mean() - 1.0) < 0.02
    # every later frame starts from the file until it is dropped
    img_b, st_b = g2.radiosity_render(count=60)
    assert np.array_equal(img_b, img)
    g2.radiosity_load(None)
    _, st3 = g2.radiosity_render(count=60)
    assert st3["gathers"] == st["gathers"]


def test_rca_skips_foreign_lines(gpu, tmp_path):
    _, g = _room(gpu)
    g.radiosity_render(count=40)
    good = tmp_path / "good.rca"
    n = g.radiosity_save(good)
    lines = good.read_text().splitlines()
    mixed = tmp_path / "mixed.rca"
    mixed.write_text("B1.0\nP\n" + "\n".join(lines[:10]) + "\nC1 1 2 3\nhello\nC9\t0\t0\t0\t7f7f7f\t1\t1\t1\t1\t1\t7f7f7f\n"
                     + "\n".join(lines[10:20]) + "\n")
    assert g.radiosity_load(mixed) == 20
    with pytest.raises(Exception):
        g.radiosity_load(tmp_path / "missing.rca")
    assert n >= 20
