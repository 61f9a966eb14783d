"""Scene-description reader (povray_amd.sdl): the statements it covers are
read as the reference parser reads them, anything else is an error naming
the line, and a parsed scene renders on the GPU as the oracle renders it."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from povray_amd import _abi as A
from povray_amd import scenes, sdl

SCENE = r"""
#include "colors.inc"
// a small test scene
#declare R = 0.75;
#declare Glass = texture { pigment { color rgbt <1, 1, 1, 0.9> } finish { ambient 0 diffuse 0.1 reflection 0.1 } }
global_settings { max_trace_level 6 ambient_light rgb <1, 1, 1> }
camera { location <0, 2, -8> look_at <0, 0.5, 0> angle 60 }
light_source { <10, 20, -15> color White }
light_source { <-6, 8, -6> color rgb 0.5 area_light <2, 0, 0>, <0, 0, 2>, 3, 3 }
background { color rgb <0.1, 0.2, 0.3> }
plane { y, -1 pigment { color Gray50 } }
sphere { <0, 0, 0>, R pigment { Red } finish { phong 0.5 phong_size 40 } translate <-2, 0, 0> scale 1 }
box { <-0.5, -0.5, -0.5>, <0.5, 0.5, 0.5> rotate <0, 30, 0> translate <2, 0, 0> pigment { rgb <0.2, 0.8, 0.2> } }
cylinder { <0, -1, 2>, <0, 1.5, 2>, 0.4 open pigment { Blue } }
cone { <1, -1, 3>, 0.6, <1, 1, 3>, 0.1 pigment { Yellow } }
disc { <-1, 1, 3>, <0, 0, -1>, 0.7, 0.2 pigment { Orange } }
torus { 0.6, 0.15 sturm rotate x * 90 translate <0, 2, 1> pigment { Magenta } }
quadric { <1, 1, 4>, <0, 0, 0>, <0, 0, 0>, -0.25 translate <2, 2, 2> pigment { Cyan } }
triangle { <-3, -1, 1>, <-2, -1, 1>, <-2.5, 0, 1> pigment { White } }
smooth_triangle { <3, -1, 1>, <0, 0, -1>, <4, -1, 1>, <0.2, 0, -1>, <3.5, 0, 1>, <0, 0.2, -1> pigment { White } }
difference {
  sphere { <0, 0, 0>, 1 }
  box { <-2, -0.2, -2>, <2, 2, 2> }
  translate <0, 0, 4>
  texture { Glass }
  interior { ior 1.5 }
}
merge { sphere { <-3, 1, 3>, 0.5 } sphere { <-2.6, 1, 3>, 0.5 } pigment { rgb 0.7 } no_shadow }
polygon { 5, <-4, 2, 2>, <-3, 2, 2>, <-3, 3, 2>, <-4, 3, 2>, <-4, 2, 2> pigment { rgb <0.3, 0.3, 0.9> } }
mesh2 {
  vertex_vectors { 4, <0, 2.5, 0>, <1, 2.5, 0>, <1, 3.5, 0>, <0, 3.5, 0> }
  face_indices { 2, <0, 1, 2>, <0, 2, 3> }
  pigment { rgb <0.9, 0.5, 0.1> }
  translate <-0.5, 0, 0>
}
"""


def test_sdl_reads_the_scene():
    s = sdl.parse(SCENE, 160, 120)
    types = [int(t) for t in s.objects[:, 0]]
    assert types == [A.PVX_PLANE, A.PVX_SPHERE, A.PVX_BOX, A.PVX_CONE, A.PVX_CONE, A.PVX_DISC, A.PVX_TORUS,
                     A.PVX_QUADRIC, A.PVX_TRIANGLE, A.PVX_SMOOTH_TRIANGLE, A.PVX_CSG, A.PVX_CSG, A.PVX_POLYGON,
                     A.PVX_MESH]
    # Parse_Plane: unit normal, Distance = -d
    assert np.allclose(s.plane_normal[0], [0, 1, 0]) and s.plane_distance[0] == 1.0
    # the sphere moved by its translate, radius from the declared value
    assert np.allclose(s.sphere_center[0], [-2, 0, 0]) and s.sphere_radius[0] == 0.75
    # cylinder: first point is the apex, one radius, open; cone: both radii
    assert s.cone_flags[0] == A.PVX_CONE_CYLINDER | A.PVX_CONE_OPEN and s.cone_radius[0][0] == 0.4
    assert s.cone_flags[1] == 0 and tuple(s.cone_radius[1]) == (0.6, 0.1)
    assert tuple(s.disc_radius[0]) == (0.7, 0.2)
    assert s.torus_sturm[0] == 1 and s.torus_transform[0] >= 0
    # the difference: an intersection whose second operand is inverted
    node = s.csg[s.objects[10, 1]]
    assert node[0] == A.PVX_CSG_INTERSECTION and node[2] == 2
    kids = s.csg_children[node[1]:node[1] + 2]
    assert kids[0][2] & A.PVX_FLAG_INVERTED == 0 and kids[1][2] & A.PVX_FLAG_INVERTED
    assert s.csg[s.objects[11, 1]][0] == A.PVX_CSG_MERGE and s.objects[11, 2] & A.PVX_FLAG_NO_SHADOW
    # materials: named colours, rgbt through a declared texture, ior
    assert np.allclose(s.materials["pigment"][1], [1, 0, 0])
    assert s.materials["transmit"][10] == 0.9 and s.materials["ior"][10] == 1.5
    assert s.materials["reflection"][10] == 0.1
    # mesh2 moved by its translate
    v = s.meshes[0][0]
    assert np.allclose(v[0], [-0.5, 2.5, 0]) and s.meshes[0][1].shape == (2, 3)
    # lights, camera (angle 60: direction length 0.5 * |right| / tan 30)
    assert s.lights[0]["kind"] == "point" and s.lights[1]["kind"] == "area" and s.lights[1]["size"] == (3, 3)
    cam = s.camera
    assert np.isclose(np.linalg.norm(cam["direction"]), 0.5 * np.linalg.norm(cam["right"]) / np.tan(np.radians(30)))
    assert s.settings["max_trace_level"] == 6 and np.allclose(s.settings["background"], [0.1, 0.2, 0.3])


@pytest.mark.parametrize("bad, what", [
    ('sphere { <0,0,0>, 1 pigment { checker } }', "checker"),
    ('#include "shapes.inc"', "shapes.inc"),
    ('blob { }', "blob"),
    ('sphere { <0,0,0>, 1 scale <1, 2, 1> }', "uniform scale"),
    ('light_source { <0,0,0> color White area_light x, y, 2, 2 jitter }', "jitter"),
    ('camera { location <0,0,0> fisheye }', "fisheye"),
])
def test_sdl_rejects_what_it_does_not_read(bad, what):
    with pytest.raises(sdl.SdlError, match=what):
        sdl.parse("\n\n" + bad, 32, 24)


def test_sdl_scene_through_the_oracle():
    s = sdl.parse(SCENE, 160, 120)
    o = OracleScene(s)
    rays = scenes.camera_rays_np(s.camera)
    h1, h0 = o.find_intersection(rays, method=1), o.find_intersection(rays, method=0)
    assert np.array_equal(h1["obj"], h0["obj"]) and np.array_equal(h1["t"], h0["t"])
    seen = set(int(x) for x in np.unique(h1["obj"]) if x >= 0)
    assert len(seen) >= 10  # most objects are in view
    assert list(s.polygon_first) == [0, 5] and s.polygon_points.shape == (5, 3)


@pytest.mark.gpu
def test_gpu_renders_parsed_scene_as_oracle(gpu):
    s = sdl.parse(SCENE, 240, 180)
    gs = gpu.GpuScene(s)
    rays = scenes.camera_rays_np(s.camera)
    g, c = gs.find_intersection(rays), OracleScene(s).find_intersection(rays, method=1)
    assert np.array_equal(g["obj"], c["obj"]) and np.array_equal(g["t"], c["t"])
    gs.set_shading()
    img, st = gs.render()
    ref, (n_closest, n_shadow) = OracleScene(s).render(s)
    ok = np.mean(np.abs(img - ref).max(axis=2) <= 1.0 / 255.0)
    assert ok >= 0.999, ok
