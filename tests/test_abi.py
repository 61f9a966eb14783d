"""C-ABI boundary checks that need no GPU: the library loads, exports every
entry point include/pvx.h declares, and the ctypes mirror matches the C
struct layouts byte for byte."""
import ctypes as C
import os
import re
import subprocess

import pytest

from povray_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pvx.h")
LIB = os.path.join(ROOT, "povray_amd", "libpvx.so")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pvx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = _declared_functions()
    for must in ("pvx_scene_create", "pvx_trace_closest", "pvx_all_intersections", "pvx_inside"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libpvx.so not built"
    lib = C.CDLL(LIB)
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    from povray_amd import pvx

    assert set(pvx.EXPORTED_SYMBOLS) == set(_declared_functions())


def test_no_device_is_reported_not_faked():
    """Without a GPU the library reports PVX_EDEVICE; it never computes on
    the CPU instead."""
    from povray_amd import hip, pvx

    if hip.device_count() > 0:
        pytest.skip("a GPU is present")
    lib = pvx.lib()
    assert lib.pvx_init(0) == A.PVX_EDEVICE
    assert b"device" in lib.pvx_last_error()
    from povray_amd import scenes

    s = scenes.sphere_scene(n=4)
    desc, keep = s.desc()
    h = C.c_void_p()
    assert lib.pvx_scene_create(C.byref(desc), None, C.byref(h)) == A.PVX_EDEVICE
    assert not h.value


_SIZE_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "pvx.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(pvx_object), sizeof(pvx_csg_node),
         sizeof(pvx_mesh_desc), sizeof(pvx_scene_desc), sizeof(pvx_ray), sizeof(pvx_hit),
         sizeof(pvx_build_opts), sizeof(pvx_scene_stats), sizeof(pvx_camera),
         offsetof(pvx_scene_desc, csg_children));
  return 0;
}
"""


def test_ctypes_layout_matches_header(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(_SIZE_PROBE)
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)])
    got = list(map(int, subprocess.check_output([str(exe)]).split()))
    want = [C.sizeof(A.PvxObject), C.sizeof(A.PvxCsgNode), C.sizeof(A.PvxMeshDesc),
            C.sizeof(A.PvxSceneDesc), C.sizeof(A.PvxRay), C.sizeof(A.PvxHit),
            C.sizeof(A.PvxBuildOpts), C.sizeof(A.PvxSceneStats), C.sizeof(A.PvxCamera),
            A.PvxSceneDesc.csg_children.offset]
    assert got == want
