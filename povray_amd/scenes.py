"""Seeded synthetic scenes for the hot-path workloads.

Every scene is generated from a splitmix64 stream so that the GPU engine and
the CPU oracle see bit-identical inputs.  Geometry follows
POV-Ray's conventions: left-handed camera basis built like the `look_at`
keyword (parser/parser.cpp:1926-1969), mesh2 vertices stored as FP32
(core/shape/mesh.h: MeshVector = SnglVector3d), planes as N.P + D = 0
(core/shape/plane.cpp:92-170).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class Rng:
    """splitmix64 counter stream; uniform() returns doubles in [0, 1)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def bits(self, n: int) -> np.ndarray:
        idx = np.arange(self.count + 1, self.count + n + 1, dtype=np.uint64)
        self.count += n
        with np.errstate(over="ignore"):
            return _mix64(self.state + idx * _GOLDEN)

    def uniform(self, n: int, lo=0.0, hi=1.0) -> np.ndarray:
        u = (self.bits(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u


# --------------------------------------------------------------------------
# camera
# --------------------------------------------------------------------------
def look_at_camera(location, look_at, width, height, sky=(0.0, 1.0, 0.0)):
    """Camera vectors as the `look_at` keyword leaves them, with the default
    direction <0,0,1>, up <0,1,0>, right <w/h,0,0> (parser/parser.cpp:1926-1969,
    core/scene/camera.cpp:228-231)."""
    loc = np.asarray(location, np.float64)
    d_len, up_len, right_len = 1.0, 1.0, width / height
    direction = np.asarray(look_at, np.float64) - loc
    direction /= np.linalg.norm(direction)
    right = np.cross(np.asarray(sky, np.float64), direction)
    right /= np.linalg.norm(right)
    up = np.cross(direction, right)
    # handedness of the default basis is positive: cross(up, dir) . right > 0
    return dict(location=loc, direction=direction * d_len, right=right * right_len,
                up=up * up_len, width=int(width), height=int(height))


# --------------------------------------------------------------------------
# scene container
# --------------------------------------------------------------------------
def _f64(a, shape):
    return np.ascontiguousarray(np.asarray(a, np.float64).reshape(shape))


@dataclass
class Scene:
    name: str
    objects: np.ndarray                      # (n, 4) uint32: type, index, flags, pad
    sphere_center: np.ndarray = None
    sphere_radius: np.ndarray = None
    plane_normal: np.ndarray = None
    plane_distance: np.ndarray = None
    box_corner1: np.ndarray = None
    box_corner2: np.ndarray = None
    box_transform: np.ndarray = None
    torus_major: np.ndarray = None
    torus_minor: np.ndarray = None
    torus_transform: np.ndarray = None
    torus_sturm: np.ndarray = None
    transforms: np.ndarray = None            # (n, 32) float64
    meshes: list = field(default_factory=list)  # (verts f32 (n,3), idx i32 (m,3), normals f32|None)
    csg: np.ndarray = None                   # (n, 4) uint32: op, first, count, pad
    csg_children: np.ndarray = None          # (n, 4) uint32 like objects
    triangle_vertices: np.ndarray = None     # (n, 9) float64: P1, P2, P3
    triangle_normals: np.ndarray = None      # (n, 9) float64: N1, N2, N3 (smooth triangles)
    cone_apex: np.ndarray = None             # (n, 3) float64: first point of cone{} / cylinder{}
    cone_base: np.ndarray = None             # (n, 3) float64: second point
    cone_radius: np.ndarray = None           # (n, 2) float64: first, second radius
    cone_flags: np.ndarray = None            # (n,) uint8: PVX_CONE_OPEN | PVX_CONE_CYLINDER
    cone_transform: np.ndarray = None        # (n,) int32: transform index or -1
    disc_center: np.ndarray = None           # (n, 3) float64
    disc_normal: np.ndarray = None           # (n, 3) float64
    disc_radius: np.ndarray = None           # (n, 2) float64: radius, hole radius
    disc_transform: np.ndarray = None        # (n,) int32: transform index or -1
    quadric_coef: np.ndarray = None          # (n, 10) float64: A..J as quadric { <A,B,C>, <D,E,F>, <G,H,I>, J }
    quadric_transform: np.ndarray = None     # (n,) int32: transform index or -1
    polygon_first: np.ndarray = None         # (n + 1,) uint32: offsets into polygon_points (in points)
    polygon_points: np.ndarray = None        # (m, 3) float64: the statements' points
    polygon_transform: np.ndarray = None     # (n,) int32: transform index or -1
    mesh_transform: np.ndarray = None        # (len(meshes),) int32: transform index or -1
    revolution_first: np.ndarray = None      # (n + 1,) uint32: offsets into revolution_points (in points)
    revolution_points: np.ndarray = None     # (m, 2) float64: lathe / sor points (radius, height)
    revolution_kind: np.ndarray = None       # (n,) uint8: PVX_LATHE_* or PVX_SOR_CUBIC
    revolution_flags: np.ndarray = None      # (n,) uint8: PVX_REV_STURM | PVX_REV_OPEN
    revolution_transform: np.ndarray = None  # (n,) int32: transform index or -1
    camera: dict = None
    lights: list = field(default_factory=list)
    materials: dict = field(default_factory=dict)  # per top-level object arrays
    settings: dict = field(default_factory=dict)

    @property
    def n_objects(self):
        return int(self.objects.shape[0])

    _ARRAYS = ("objects", "sphere_center", "sphere_radius", "plane_normal", "plane_distance", "box_corner1",
               "box_corner2", "box_transform", "torus_major", "torus_minor", "torus_transform", "torus_sturm",
               "transforms", "csg", "csg_children", "triangle_vertices", "triangle_normals", "cone_apex",
               "cone_base", "cone_radius", "cone_flags", "cone_transform", "disc_center", "disc_normal",
               "disc_radius", "disc_transform", "quadric_coef", "quadric_transform", "polygon_first",
               "polygon_points", "polygon_transform", "mesh_transform", "revolution_first", "revolution_points",
               "revolution_kind", "revolution_flags", "revolution_transform")

    def save(self, path):
        """Write the scene to a compressed .npz (arrays, camera, lights,
        materials, settings; no pickled objects)."""
        out = {"name": np.array(self.name)}
        for k in self._ARRAYS:
            v = getattr(self, k)
            if v is not None:
                out["a_" + k] = np.asarray(v)
        for i, (v, ix, nrm) in enumerate(self.meshes):
            out[f"mesh{i}_v"], out[f"mesh{i}_f"] = np.asarray(v), np.asarray(ix)
            if nrm is not None:
                out[f"mesh{i}_n"] = np.asarray(nrm)
        for k, v in (self.camera or {}).items():
            out["cam_" + k] = np.asarray(v)
        for i, lt in enumerate(self.lights):
            for k, v in lt.items():
                out[f"light{i}_{k}"] = np.asarray(v)
        for k, v in self.materials.items():
            out["mat_" + k] = np.asarray(v)
        for k, v in self.settings.items():
            if k == "radiosity":
                for k2, v2 in v.items():
                    out["rad_" + k2] = np.asarray(v2)
            elif k != "warnings":
                out["set_" + k] = np.asarray(v)
        np.savez_compressed(path, **out)

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        kw = {k[2:]: z[k] for k in z.files if k.startswith("a_")}
        meshes, i = [], 0
        while f"mesh{i}_v" in z.files:
            meshes.append((z[f"mesh{i}_v"], z[f"mesh{i}_f"], z[f"mesh{i}_n"] if f"mesh{i}_n" in z.files else None))
            i += 1
        cam = {k[4:]: z[k] for k in z.files if k.startswith("cam_")}
        for k in ("width", "height"):
            if k in cam:
                cam[k] = int(cam[k])
        lights, i = [], 0
        while any(k.startswith(f"light{i}_") for k in z.files):
            lt = {k[len(f"light{i}_"):]: z[k] for k in z.files if k.startswith(f"light{i}_")}
            lt["kind"] = str(lt["kind"])
            if "size" in lt:
                lt["size"] = tuple(int(x) for x in lt["size"])
            lights.append(lt)
            i += 1
        mats = {k[4:]: z[k] for k in z.files if k.startswith("mat_")}
        settings = {k[4:]: (z[k].item() if z[k].ndim == 0 else z[k]) for k in z.files if k.startswith("set_")}
        rad = {k[4:]: z[k].item() for k in z.files if k.startswith("rad_")}
        if rad:
            settings["radiosity"] = rad
        return cls(name=str(z["name"]), meshes=meshes, camera=cam or None, lights=lights, materials=mats,
                   settings=settings, **kw)

    @property
    def n_triangles(self):
        return int(sum(m[1].shape[0] for m in self.meshes))

    def desc(self):
        """Build a pvx_scene_desc that points into this scene's arrays.
        Returns (desc, keepalive); keep both alive while the desc is used."""
        keep = []

        def arr(a, dtype, ctype):
            if a is None or len(a) == 0:
                return None, 0
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype)), a.shape[0]

        d = A.PvxSceneDesc()
        objs = np.ascontiguousarray(self.objects, np.uint32).reshape(-1, 4)
        keep.append(objs)
        d.n_objects = objs.shape[0]
        d.objects = objs.ctypes.data_as(C.POINTER(A.PvxObject))
        if self.sphere_center is not None and len(self.sphere_center):
            d.sphere_center, d.n_spheres = arr(self.sphere_center, np.float64, C.c_double)
            d.sphere_radius, _ = arr(self.sphere_radius, np.float64, C.c_double)
        if self.plane_normal is not None and len(self.plane_normal):
            d.plane_normal, d.n_planes = arr(self.plane_normal, np.float64, C.c_double)
            d.plane_distance, _ = arr(self.plane_distance, np.float64, C.c_double)
        if self.box_corner1 is not None and len(self.box_corner1):
            d.box_corner1, d.n_boxes = arr(self.box_corner1, np.float64, C.c_double)
            d.box_corner2, _ = arr(self.box_corner2, np.float64, C.c_double)
            d.box_transform, _ = arr(self.box_transform, np.int32, C.c_int32)
        if self.torus_major is not None and len(self.torus_major):
            d.torus_major, d.n_tori = arr(self.torus_major, np.float64, C.c_double)
            d.torus_minor, _ = arr(self.torus_minor, np.float64, C.c_double)
            d.torus_transform, _ = arr(self.torus_transform, np.int32, C.c_int32)
            d.torus_sturm, _ = arr(self.torus_sturm, np.uint8, C.c_uint8)
        if self.transforms is not None and len(self.transforms):
            d.transforms, d.n_transforms = arr(self.transforms, np.float64, C.c_double)
        if self.meshes:
            md = (A.PvxMeshDesc * len(self.meshes))()
            for i, (v, ix, nrm) in enumerate(self.meshes):
                v = np.ascontiguousarray(v, np.float32)
                ix = np.ascontiguousarray(ix, np.int32)
                keep += [v, ix]
                md[i].n_vertices = v.shape[0]
                md[i].n_triangles = ix.shape[0]
                md[i].vertices = v.ctypes.data_as(A.c_fp)
                md[i].indices = ix.ctypes.data_as(A.c_i32p)
                if nrm is not None:
                    nrm = np.ascontiguousarray(nrm, np.float32)
                    keep.append(nrm)
                    md[i].n_normals = nrm.shape[0]
                    md[i].normals = nrm.ctypes.data_as(A.c_fp)
            keep.append(md)
            d.n_meshes = len(self.meshes)
            d.meshes = C.cast(md, C.POINTER(A.PvxMeshDesc))
            if self.mesh_transform is not None:
                d.mesh_transform, _ = arr(self.mesh_transform, np.int32, C.c_int32)
        if self.triangle_vertices is not None and len(self.triangle_vertices):
            d.triangle_vertices, _ = arr(np.asarray(self.triangle_vertices).reshape(-1), np.float64, C.c_double)
            d.n_triangles = int(np.asarray(self.triangle_vertices).reshape(-1, 9).shape[0])
            if self.triangle_normals is not None:
                d.triangle_normals, _ = arr(np.asarray(self.triangle_normals).reshape(-1), np.float64, C.c_double)
        if self.cone_apex is not None and len(self.cone_apex):
            d.cone_apex, d.n_cones = arr(np.asarray(self.cone_apex).reshape(-1, 3), np.float64, C.c_double)
            d.cone_base, _ = arr(np.asarray(self.cone_base).reshape(-1, 3), np.float64, C.c_double)
            d.cone_radius, _ = arr(np.asarray(self.cone_radius).reshape(-1, 2), np.float64, C.c_double)
            if self.cone_flags is not None:
                d.cone_flags, _ = arr(self.cone_flags, np.uint8, C.c_uint8)
            if self.cone_transform is not None:
                d.cone_transform, _ = arr(self.cone_transform, np.int32, C.c_int32)
        if self.disc_center is not None and len(self.disc_center):
            d.disc_center, d.n_discs = arr(np.asarray(self.disc_center).reshape(-1, 3), np.float64, C.c_double)
            d.disc_normal, _ = arr(np.asarray(self.disc_normal).reshape(-1, 3), np.float64, C.c_double)
            d.disc_radius, _ = arr(np.asarray(self.disc_radius).reshape(-1, 2), np.float64, C.c_double)
            if self.disc_transform is not None:
                d.disc_transform, _ = arr(self.disc_transform, np.int32, C.c_int32)
        if self.quadric_coef is not None and len(self.quadric_coef):
            d.quadric_coef, d.n_quadrics = arr(np.asarray(self.quadric_coef).reshape(-1, 10), np.float64, C.c_double)
            if self.quadric_transform is not None:
                d.quadric_transform, _ = arr(self.quadric_transform, np.int32, C.c_int32)
        if self.polygon_first is not None and len(self.polygon_first) > 1:
            d.polygon_first, n1 = arr(self.polygon_first, np.uint32, C.c_uint32)
            d.n_polygons = n1 - 1
            d.polygon_points, _ = arr(np.asarray(self.polygon_points).reshape(-1, 3), np.float64, C.c_double)
            if self.polygon_transform is not None:
                d.polygon_transform, _ = arr(self.polygon_transform, np.int32, C.c_int32)
        if self.revolution_first is not None and len(self.revolution_first) > 1:
            d.revolution_first, n1 = arr(self.revolution_first, np.uint32, C.c_uint32)
            d.n_revolutions = n1 - 1
            d.revolution_points, _ = arr(np.asarray(self.revolution_points).reshape(-1, 2), np.float64, C.c_double)
            d.revolution_kind, _ = arr(self.revolution_kind, np.uint8, C.c_uint8)
            if self.revolution_flags is not None:
                d.revolution_flags, _ = arr(self.revolution_flags, np.uint8, C.c_uint8)
            if self.revolution_transform is not None:
                d.revolution_transform, _ = arr(self.revolution_transform, np.int32, C.c_int32)
        if self.csg is not None and len(self.csg):
            c = np.ascontiguousarray(self.csg, np.uint32).reshape(-1, 4)
            ch = np.ascontiguousarray(self.csg_children, np.uint32).reshape(-1, 4)
            keep += [c, ch]
            d.n_csg = c.shape[0]
            d.csg = c.ctypes.data_as(C.POINTER(A.PvxCsgNode))
            d.n_csg_children = ch.shape[0]
            d.csg_children = ch.ctypes.data_as(C.POINTER(A.PvxObject))
        return d, keep

    def shading(self):
        """pvx_shading_desc for this scene's materials and lights.
        Returns (desc, keepalive)."""
        m = self.materials
        n = self.n_objects
        mats = (A.PvxMaterial * n)()
        for i in range(n):
            t = mats[i]
            t.pigment[:] = [float(x) for x in m["pigment"][i]]
            t.filter = float(m["filter"][i]) if "filter" in m else 0.0
            t.transmit = float(m.get("transmit", np.zeros(n))[i])
            amb = float(m.get("ambient", np.full(n, 0.1))[i])
            t.ambient[:] = [amb, amb, amb]
            em = m.get("emission", None)
            t.emission[:] = [float(x) for x in em[i]] if em is not None else [0.0, 0.0, 0.0]
            t.diffuse = float(m.get("diffuse", np.full(n, 0.6))[i])
            t.brilliance = float(m["brilliance"][i]) if "brilliance" in m else 1.0
            t.phong = float(m.get("phong", np.zeros(n))[i])
            t.phong_size = float(m.get("phong_size", np.full(n, 40.0))[i])
            r = float(m.get("reflection", np.zeros(n))[i])
            t.reflection_min[:] = [r, r, r]
            t.reflection_max[:] = [r, r, r]
            t.reflection_falloff = 1.0
            ior = float(m.get("ior", np.ones(n))[i])
            t.ior = ior if (t.transmit > 0.0 or t.filter > 0.0) else 0.0
        lights = (A.PvxLight * max(1, len(self.lights)))()
        for i, L in enumerate(self.lights):
            lights[i].kind = A.PVX_LIGHT_AREA if L["kind"] == "area" else A.PVX_LIGHT_POINT
            lights[i].position[:] = [float(x) for x in L["position"]]
            if L["kind"] == "area":
                lights[i].axis1[:] = [float(x) for x in L["axis1"]]
                lights[i].axis2[:] = [float(x) for x in L["axis2"]]
                lights[i].size1, lights[i].size2 = L["size"]
            else:
                lights[i].size1 = lights[i].size2 = 1
            lights[i].colour[:] = [float(x) for x in L["colour"]]
        d = A.PvxShadingDesc()
        d.n_materials = n
        d.materials = C.cast(mats, C.POINTER(A.PvxMaterial))
        d.n_lights = len(self.lights)
        d.lights = C.cast(lights, C.POINTER(A.PvxLight))
        d.ambient_light[:] = [float(x) for x in self.settings.get("ambient_light", (1.0, 1.0, 1.0))]
        d.background[:] = [float(x) for x in self.settings.get("background", (0.0, 0.0, 0.0))]
        d.max_trace_level = int(self.settings.get("max_trace_level", 5))
        d.adc_bailout = float(self.settings.get("adc_bailout", 1.0 / 255.0))
        return d, [mats, lights]

    def camera_struct(self, jitter=(0.5, 0.5), width=None, height=None):
        c = self.camera
        cam = A.PvxCamera()
        for k in ("location", "direction", "right", "up"):
            getattr(cam, k)[:] = [float(x) for x in c[k]]
        cam.width = int(width or c["width"])
        cam.height = int(height or c["height"])
        cam.jitter_x, cam.jitter_y = jitter
        return cam


def camera_rays_np(cam: dict, width=None, height=None, jitter=(0.5, 0.5)):
    """Camera rays on the host, same formula as pvx_camera_rays
    (core/render/tracepixel.cpp:357-372, 924): (n, 6) float64 [o, d]."""
    w = int(width or cam["width"])
    h = int(height or cam["height"])
    ys, xs = np.meshgrid(np.arange(h, dtype=np.float64), np.arange(w, dtype=np.float64),
                         indexing="ij")
    x0 = (xs + jitter[0]) / w - 0.5
    y0 = 0.5 - (ys + jitter[1]) / h
    d = (cam["direction"][None, None, :] + x0[..., None] * cam["right"][None, None, :]
         + y0[..., None] * cam["up"][None, None, :])
    d /= np.sqrt((d * d).sum(-1))[..., None]
    o = np.broadcast_to(cam["location"], d.shape)
    return np.ascontiguousarray(np.concatenate([o, d], -1).reshape(-1, 6))


# --------------------------------------------------------------------------
# 1,000 spheres (the CPU-runnable workload)
# --------------------------------------------------------------------------
def sphere_scene(seed: int = 0xC0FFEE + 0, n: int = 1000, width=640, height=480) -> Scene:
    """1,000 spheres, centres U[-10,10]^3, radii U[0.2,0.8], diffuse+phong
    finish, 5% reflective (0.3); point light (20,30,-20); camera (0,0,-30)
    looking at the origin; max_trace_level 5."""
    r = Rng(seed)
    centre = r.uniform(3 * n, -10.0, 10.0).reshape(n, 3)
    radius = r.uniform(n, 0.2, 0.8)
    albedo = r.uniform(3 * n).reshape(n, 3)
    reflect = np.where(r.uniform(n) < 0.05, 0.3, 0.0)
    objects = np.zeros((n, 4), np.uint32)
    objects[:, 0] = A.PVX_SPHERE
    objects[:, 1] = np.arange(n)
    return Scene(
        name="spheres1000", objects=objects, sphere_center=centre, sphere_radius=radius,
        camera=look_at_camera((0.0, 0.0, -30.0), (0.0, 0.0, 0.0), width, height),
        lights=[dict(kind="point", position=np.array([20.0, 30.0, -20.0]),
                     colour=np.ones(3))],
        materials=dict(pigment=albedo, diffuse=np.full(n, 0.6), ambient=np.full(n, 0.1),
                       phong=np.full(n, 0.5), phong_size=np.full(n, 40.0),
                       reflection=reflect, ior=np.ones(n), transmit=np.zeros(n)),
        settings=dict(max_trace_level=5, adc_bailout=1.0 / 255.0),
    )


# --------------------------------------------------------------------------
# procedural icosphere mesh
# --------------------------------------------------------------------------
def icosphere(level: int):
    """Unit icosphere: 20 * 4**level faces, outward (counter-clockwise seen
    from outside) winding.  Returns (verts float64 (n,3), faces int32 (m,3))."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1)[:, None]
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(level):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], 0)
        lo = e.min(1)
        hi = e.max(1)
        key = lo * nv + hi
        uk, inv = np.unique(key, return_inverse=True)
        a = uk // nv
        b = uk % nv
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1)[:, None]
        m = inv.reshape(3, -1) + nv  # midpoint ids for edges 01, 12, 20
        v = np.concatenate([v, mid], 0)
        m01, m12, m20 = m[0], m[1], m[2]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1),
            np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1),
            np.stack([m01, m12, m20], 1)], 0)
    return v, f.astype(np.int32)


def _lattice_hash(ix, iy, iz, seed):
    with np.errstate(over="ignore"):
        k = (ix.astype(np.int64).astype(np.uint64) * np.uint64(0x8DA6B343)
             ^ iy.astype(np.int64).astype(np.uint64) * np.uint64(0xD8163841)
             ^ iz.astype(np.int64).astype(np.uint64) * np.uint64(0xCB1AB31F)
             ^ np.uint64(seed & 0xFFFFFFFFFFFFFFFF))
        h = _mix64(k)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def value_noise(p: np.ndarray, seed: int, octaves: int = 4, base_freq: float = 2.0):
    """Seeded multi-octave value noise in [-1, 1]: trilinear interpolation of
    hashed lattice values with a smoothstep fade; octave k has frequency
    base_freq * 2**k and weight 0.5**k (normalised by the weight sum)."""
    total = np.zeros(p.shape[0])
    wsum = 0.0
    for k in range(octaves):
        q = p * (base_freq * 2.0 ** k)
        i0 = np.floor(q)
        fr = q - i0
        s = fr * fr * (3.0 - 2.0 * fr)
        ix, iy, iz = (i0[:, 0].astype(np.int64), i0[:, 1].astype(np.int64),
                      i0[:, 2].astype(np.int64))
        acc = np.zeros(p.shape[0])
        for dx in (0, 1):
            wx = s[:, 0] if dx else 1.0 - s[:, 0]
            for dy in (0, 1):
                wy = s[:, 1] if dy else 1.0 - s[:, 1]
                for dz in (0, 1):
                    wz = s[:, 2] if dz else 1.0 - s[:, 2]
                    acc += wx * wy * wz * _lattice_hash(ix + dx, iy + dy, iz + dz, seed + k)
        w = 0.5 ** k
        total += w * (2.0 * acc - 1.0)
        wsum += w
    return total / wsum


def vertex_normals(v: np.ndarray, f: np.ndarray) -> np.ndarray:
    """Smoothed per-vertex normals: area-weighted sum of face normals."""
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    n = np.zeros_like(v)
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    n /= np.linalg.norm(n, axis=1)[:, None]
    return n


def displaced_icosphere(level: int, seed: int, amplitude: float = 0.15, radius: float = 1.0):
    v, f = icosphere(level)
    disp = value_noise(v, seed) * amplitude
    v = v * (radius * (1.0 + disp))[:, None]
    v32 = v.astype(np.float32)
    n32 = vertex_normals(v32.astype(np.float64), f).astype(np.float32)
    return v32, f, n32


def mesh_scene(seed: int = 0xC0FFEE + 1, level: int = 8, width=1920, height=1080) -> Scene:
    """Icosphere subdivided `level` times (20*4**level triangles; 1,310,720
    at level 8) displaced along the normal by 4-octave value noise
    (amplitude 0.15) with smoothed vertex normals, as mesh2 with an index
    buffer; ground plane y = -1.25; two 4x4 area lights; 1920x1080."""
    v, f, n = displaced_icosphere(level, seed)
    objects = np.array([[A.PVX_MESH, 0, 0, 0], [A.PVX_PLANE, 0, 0, 0]], np.uint32)
    return Scene(
        name=f"icosphere_L{level}", objects=objects, meshes=[(v, f, n)],
        plane_normal=np.array([[0.0, 1.0, 0.0]]), plane_distance=np.array([1.25]),
        camera=look_at_camera((0.0, 0.3, -1.95), (0.0, -0.05, 0.0), width, height),
        lights=[dict(kind="area", position=np.array([-6.0, 8.0, -6.0]),
                     axis1=np.array([2.0, 0.0, 0.0]), axis2=np.array([0.0, 0.0, 2.0]),
                     size=(4, 4), colour=np.full(3, 0.6)),
                dict(kind="area", position=np.array([7.0, 6.0, -3.0]),
                     axis1=np.array([0.0, 2.0, 0.0]), axis2=np.array([0.0, 0.0, 2.0]),
                     size=(4, 4), colour=np.full(3, 0.5))],
        materials=dict(pigment=np.array([[0.8, 0.6, 0.4], [0.7, 0.7, 0.7]]),
                       diffuse=np.array([0.7, 0.7]), ambient=np.array([0.1, 0.1]),
                       phong=np.array([0.4, 0.0]), phong_size=np.array([60.0, 40.0]),
                       reflection=np.array([0.25, 0.1]), ior=np.ones(2),
                       transmit=np.zeros(2)),
        settings=dict(max_trace_level=5, adc_bailout=1.0 / 255.0),
    )


# --------------------------------------------------------------------------
# transforms (core/math/matrix.cpp conventions: row vectors, matrix+inverse)
# --------------------------------------------------------------------------
def rotation_transform(deg):
    """Compute_Rotation_Transform (core/math/matrix.cpp): X, then Y, then Z;
    the inverse is the product of the transposes.  Returns (matrix, inverse)."""
    rx, ry, rz = np.radians(np.asarray(deg, np.float64))
    cx, sx, cy, sy, cz, sz = np.cos(rx), np.sin(rx), np.cos(ry), np.sin(ry), np.cos(rz), np.sin(rz)
    mx = np.eye(4)
    mx[1, 1] = mx[2, 2] = cx
    mx[1, 2] = sx
    mx[2, 1] = -sx
    my = np.eye(4)
    my[0, 0] = my[2, 2] = cy
    my[0, 2] = -sy
    my[2, 0] = sy
    mz = np.eye(4)
    mz[0, 0] = mz[1, 1] = cz
    mz[0, 1] = sz
    mz[1, 0] = -sz
    m = mx @ my @ mz
    inv = mz.T @ my.T @ mx.T
    return m, inv


def translation_transform(v):
    m = np.eye(4)
    m[3, :3] = v
    inv = np.eye(4)
    inv[3, :3] = -np.asarray(v, np.float64)
    return m, inv


def compose(a, b):
    """Compose_Transforms(a, b): apply a, then b."""
    return a[0] @ b[0], b[1] @ a[1]


def _xf_row(t):
    return np.concatenate([t[0].reshape(-1), t[1].reshape(-1)])


# --------------------------------------------------------------------------
# CSG and quartic mix
# --------------------------------------------------------------------------
_OPS = ("union", "intersection", "difference", "merge")


def csg_scene(seed: int = 0xC0FFEE + 2, n_objects: int = 10000, width=1280, height=720,
              extent: float | None = None, sturm: bool = True) -> Scene:
    """`n_objects` leaves in equal thirds boxes (extents U[0.1,1]), spheres
    and tori (R U[0.5,1.5], r = U[0.1,0.3]*R), grouped into full binary CSG
    trees of depth 3 (8 leaves each) with random union / intersection /
    difference / merge operators and random rotations; half of the trees are
    glass (ior 1.5); max_trace_level 10.  Differences are stored the way the
    parser builds them: an intersection whose children 2..n are inverted
    (parser/parser.cpp:2223-2225, csg.cpp:629-645)."""
    r = Rng(seed)
    n_trees = max(1, n_objects // 8)
    if extent is None:  # keep the object density of the 1,250-tree field
        extent = 24.0 * (n_trees / 1250.0) ** (1.0 / 3.0)
    kinds = np.array([0, 1, 2] * ((n_trees * 8 + 2) // 3))[: n_trees * 8]
    perm = np.argsort(r.uniform(kinds.shape[0]))
    kinds = kinds[perm]
    centres = r.uniform(3 * n_trees, -extent, extent).reshape(n_trees, 3)
    centres[:, 2] = r.uniform(n_trees, 0.0, 2.0 * extent)

    sph_c, sph_r = [], []
    box_c1, box_c2, box_tr = [], [], []
    tor_R, tor_r, tor_tr = [], [], []
    xforms = []
    csg_nodes, csg_children = [], []

    def leaf(kind, pos):
        if kind == 1:
            sph_c.append(pos)
            sph_r.append(float(r.uniform(1, 0.2, 0.6)[0]))
            return [A.PVX_SPHERE, len(sph_c) - 1, 0, 0]
        rot = rotation_transform(r.uniform(3, 0.0, 360.0))
        t = compose(rot, translation_transform(pos))
        xforms.append(_xf_row(t))
        if kind == 0:
            e = r.uniform(3, 0.1, 1.0)
            box_c1.append(-0.5 * e)
            box_c2.append(0.5 * e)
            box_tr.append(len(xforms) - 1)
            return [A.PVX_BOX, len(box_c1) - 1, 0, 0]
        R = float(r.uniform(1, 0.5, 1.5)[0])
        tor_R.append(R)
        tor_r.append(R * float(r.uniform(1, 0.1, 0.3)[0]))
        tor_tr.append(len(xforms) - 1)
        return [A.PVX_TORUS, len(tor_R) - 1, 0, 0]

    # build nested trees as python structures first, then flatten
    def invert(node):
        """Object::Invert: every node toggles INVERTED; CSG operators change
        union -> intersection, intersection -> merge, merge -> intersection
        (CompoundObject::Invert, csg.cpp:629-645), children inverted too."""
        if node[0] == "leaf":
            node[1][2] ^= A.PVX_FLAG_INVERTED
            return node
        op, kids, flags = node[1], node[2], node[3]
        kids = [invert(k) for k in kids]
        new_op = {"union": "intersection", "intersection": "merge", "merge": "intersection"}[op]
        return ("csg", new_op, kids, flags ^ A.PVX_FLAG_INVERTED)

    def make(level, base, pos):
        if level == 3:
            return ("leaf", leaf(int(kinds[base]), pos + r.uniform(3, -0.6, 0.6)))
        op = _OPS[int(r.uniform(1)[0] * 4) % 4]
        span = 1 << (3 - level - 1)
        kids = [make(level + 1, base, pos), make(level + 1, base + span, pos)]
        if op == "difference":
            kids = [kids[0]] + [invert(k) for k in kids[1:]]
            op = "intersection"
        return ("csg", op, kids, 0)

    op_code = {"union": A.PVX_CSG_UNION, "intersection": A.PVX_CSG_INTERSECTION,
               "merge": A.PVX_CSG_MERGE}

    def flatten(node):
        idx = len(csg_nodes)
        csg_nodes.append(None)
        first = len(csg_children)
        kids = node[2]
        csg_children.extend([None] * len(kids))
        for k, kid in enumerate(kids):
            if kid[0] == "leaf":
                csg_children[first + k] = kid[1]
            else:
                csg_children[first + k] = [A.PVX_CSG, flatten(kid), kid[3], 0]
        csg_nodes[idx] = [op_code[node[1]], first, len(kids), 0]
        return idx

    roots = []
    for t in range(n_trees):
        tree = make(0, t * 8, centres[t])
        roots.append(flatten(tree))
    objects = np.zeros((n_trees, 4), np.uint32)
    objects[:, 0] = A.PVX_CSG
    objects[:, 1] = roots
    glass = r.uniform(n_trees) < 0.5
    alb = r.uniform(3 * n_trees).reshape(n_trees, 3)
    return Scene(
        name=f"csg{n_objects}", objects=objects,
        sphere_center=np.array(sph_c).reshape(-1, 3), sphere_radius=np.array(sph_r),
        box_corner1=np.array(box_c1).reshape(-1, 3), box_corner2=np.array(box_c2).reshape(-1, 3),
        box_transform=np.array(box_tr, np.int32),
        torus_major=np.array(tor_R), torus_minor=np.array(tor_r),
        torus_transform=np.array(tor_tr, np.int32),
        torus_sturm=np.full(len(tor_R), 1 if sturm else 0, np.uint8),
        transforms=np.array(xforms).reshape(-1, 32),
        csg=np.array(csg_nodes, np.uint32).reshape(-1, 4),
        csg_children=np.array(csg_children, np.uint32).reshape(-1, 4),
        camera=look_at_camera((0.0, 0.0, -extent * 1.2), (0.0, 0.0, extent * 0.5), width, height),
        lights=[dict(kind="point", position=np.array([30.0, 40.0, -30.0]), colour=np.ones(3))],
        materials=dict(pigment=alb, diffuse=np.full(n_trees, 0.6), ambient=np.full(n_trees, 0.1),
                       phong=np.full(n_trees, 0.3), phong_size=np.full(n_trees, 40.0),
                       reflection=np.where(glass, 0.1, 0.0), ior=np.where(glass, 1.5, 1.0),
                       transmit=np.where(glass, 0.9, 0.0)),
        settings=dict(max_trace_level=10, adc_bailout=1.0 / 255.0),
    )


def scale_transform(s):
    """Compute_Scaling_Transform for a uniform factor s."""
    m = np.eye(4)
    inv = np.eye(4)
    m[0, 0] = m[1, 1] = m[2, 2] = s
    inv[0, 0] = inv[1, 1] = inv[2, 2] = 1.0 / s
    return m, inv


def placed_mesh_scene(seed: int = 0xC0FFEE + 8, level: int = 4, width=320, height=240) -> Scene:
    """Four instances of one displaced icosphere mesh2 (level 4, smoothed
    normals): one unplaced, three with their own placement (rotation, a
    non-uniform scale, translation), so the rays move into mesh space with a
    direction of length != 1 (Mesh::Transform / Mesh::Intersect); a point
    light and a ground plane."""
    v, f, n = displaced_icosphere(level, seed)
    r = Rng(seed)
    xforms = []
    for k in range(3):
        sc = np.eye(4)
        inv = np.eye(4)
        s3 = r.uniform(3, 0.5, 1.6)
        sc[0, 0], sc[1, 1], sc[2, 2] = s3
        inv[0, 0], inv[1, 1], inv[2, 2] = 1.0 / s3
        t = compose(compose((sc, inv), rotation_transform(r.uniform(3, 0.0, 360.0))),
                    translation_transform([2.4 * (k - 1), 1.9 if k == 1 else 0.2 * k, 1.2 - 0.3 * k]))
        xforms.append(_xf_row(t))
    objects = np.array([[A.PVX_MESH, i, 0, 0] for i in range(4)] + [[A.PVX_PLANE, 0, 0, 0]], np.uint32)
    nobj = objects.shape[0]
    return Scene(
        name=f"placed_meshes_L{level}", objects=objects, meshes=[(v, f, n)] * 4,
        mesh_transform=np.array([-1, 0, 1, 2], np.int32), transforms=np.array(xforms).reshape(-1, 32),
        plane_normal=np.array([[0.0, 1.0, 0.0]]), plane_distance=np.array([1.6]),
        camera=look_at_camera((0.0, 1.2, -4.5), (0.0, 0.0, 1.2), width, height),
        lights=[dict(kind="point", position=np.array([-5.0, 8.0, -6.0]), colour=np.ones(3))],
        materials=dict(pigment=r.uniform(3 * nobj, 0.3, 0.9).reshape(nobj, 3), diffuse=np.full(nobj, 0.7),
                       ambient=np.full(nobj, 0.1), phong=np.full(nobj, 0.4), phong_size=np.full(nobj, 40.0),
                       reflection=np.where(np.arange(nobj) == 2, 0.3, 0.0)),
        settings=dict(max_trace_level=5, adc_bailout=1.0 / 255.0),
    )


# Profiles (radius, height) of the lathe / sor scene, one per spline kind.
REVOLUTION_PROFILES = {
    A.PVX_LATHE_LINEAR: [(0.0, 0.0), (1.0, 0.0), (1.0, 0.2), (0.25, 0.45), (0.2, 1.5), (0.9, 1.9), (1.1, 2.6)],
    A.PVX_LATHE_QUADRATIC: [(0.0, -0.2), (0.8, 0.0), (1.0, 0.5), (0.6, 1.2), (0.9, 1.8), (1.0, 2.2)],
    A.PVX_LATHE_CUBIC: [(0.3, -0.5), (0.8, 0.0), (1.1, 0.6), (0.5, 1.4), (0.7, 2.0), (0.9, 2.4)],
    A.PVX_LATHE_BEZIER: [(0.0, 0.0), (1.2, 0.0), (1.2, 0.8), (0.6, 1.0),
                         (0.6, 1.0), (0.2, 1.2), (0.8, 2.0), (0.9, 2.2)],
    A.PVX_SOR_CUBIC: [(0.0, -0.5), (0.9, 0.0), (1.2, 0.6), (0.7, 1.2), (0.5, 1.8), (0.9, 2.3), (1.0, 2.8)],
}


def revolution_scene(seed: int = 0xC0FFEE + 9, n: int = 20, width=320, height=240) -> Scene:
    """`n` lathes and sors on a 5-wide grid cycling through the five spline
    kinds (linear, quadratic, cubic and Bezier lathes, the sor's cubic), one
    in three with `sturm`, every other sor `open`; each placed by a random
    rotation, a non-uniform scale U[0.5, 1.2] per axis for one in three, and
    a translation.  Two CSG objects: a cubic lathe cut by a box
    (intersection) and a closed sor minus a sphere (difference).  A ground
    plane, one point light; one in five reflective."""
    r = Rng(seed)
    kinds = list(REVOLUTION_PROFILES)
    first, pts, kind, flags, xf_idx, xforms = [0], [], [], [], [], []

    def add(k, fl, t):
        pts.extend(REVOLUTION_PROFILES[k])
        first.append(len(pts))
        kind.append(k)
        flags.append(fl)
        xforms.append(_xf_row(t))
        xf_idx.append(len(xforms) - 1)
        return len(kind) - 1

    for i in range(n):
        k = kinds[i % len(kinds)]
        fl = (A.PVX_REV_STURM if i % 3 == 1 else 0) | (A.PVX_REV_OPEN if k == A.PVX_SOR_CUBIC and i % 2 else 0)
        t = rotation_transform(r.uniform(3, -35.0, 35.0))
        if i % 3 == 2:
            s3 = r.uniform(3, 0.5, 1.2)
            t = compose((np.diag(list(s3) + [1.0]), np.diag(list(1.0 / s3) + [1.0])), t)
        t = compose(t, translation_transform([(i % 5) * 3.0 - 6.0, 0.0, (i // 5) * 3.0]))
        add(k, fl, t)
    lathe_c = add(A.PVX_LATHE_CUBIC, 0, translation_transform([-4.5, 0.0, -3.0]))
    sor_c = add(A.PVX_SOR_CUBIC, 0, translation_transform([4.5, 0.0, -3.0]))
    objects = [[A.PVX_SOR if kind[i] == A.PVX_SOR_CUBIC else A.PVX_LATHE, i, 0, 0] for i in range(n)]
    objects += [[A.PVX_CSG, 0, 0, 0], [A.PVX_CSG, 1, 0, 0], [A.PVX_PLANE, 0, 0, 0]]
    csg = np.array([[A.PVX_CSG_INTERSECTION, 0, 2, 0], [A.PVX_CSG_INTERSECTION, 2, 2, 0]], np.uint32)
    kids = np.array([[A.PVX_LATHE, lathe_c, 0, 0], [A.PVX_BOX, 0, 0, 0],
                     [A.PVX_SOR, sor_c, 0, 0], [A.PVX_SPHERE, 0, A.PVX_FLAG_INVERTED, 0]], np.uint32)
    nobj = len(objects)
    return Scene(
        name=f"revolutions{n}", objects=np.array(objects, np.uint32), csg=csg, csg_children=kids,
        revolution_first=np.array(first, np.uint32), revolution_points=np.array(pts, np.float64),
        revolution_kind=np.array(kind, np.uint8), revolution_flags=np.array(flags, np.uint8),
        revolution_transform=np.array(xf_idx, np.int32), transforms=np.array(xforms).reshape(-1, 32),
        box_corner1=np.array([[-5.6, 0.3, -4.2]]), box_corner2=np.array([[-3.2, 3.0, -3.0]]),
        sphere_center=np.array([[4.9, 2.2, -3.6]]), sphere_radius=np.array([0.9]),
        plane_normal=np.array([[0.0, 1.0, 0.0]]), plane_distance=np.array([0.6]),
        camera=look_at_camera((0.0, 6.0, -13.0), (0.0, 0.8, 3.0), width, height),
        lights=[dict(kind="point", position=np.array([-8.0, 14.0, -10.0]), colour=np.ones(3))],
        materials=dict(pigment=r.uniform(3 * nobj, 0.3, 0.9).reshape(nobj, 3), diffuse=np.full(nobj, 0.7),
                       ambient=np.full(nobj, 0.1), phong=np.full(nobj, 0.4), phong_size=np.full(nobj, 40.0),
                       reflection=np.where(np.arange(nobj) % 5 == 3, 0.3, 0.0)),
        settings=dict(max_trace_level=5, adc_bailout=1.0 / 255.0),
    )


def torus_scene(seed: int = 0xC0FFEE + 7, n: int = 48, width=320, height=240, sturm: bool = False) -> Scene:
    """`n` top-level tori on an 8-wide grid, randomly rotated.  Two in three
    are ordinary (R U[0.5,1.5], r = U[0.1,0.3]*R); every third is a thin ring
    with R = 1e4, r = 200 in its own space scaled by 1e-4 (R 1, r 0.02 in the
    world), whose quartic's coefficients span more than 12 decades, so that
    difficult_coeffs (polynomialsolver.cpp:1047-1087) sends it to the Sturm
    solver even without `sturm`.  In-plane rays through the centres come
    from `torus_plane_rays`."""
    r = Rng(seed)
    R_, r_, xf = [], [], []
    for i in range(n):
        pos = np.array([(i % 8) * 3.0 - 10.5, (i // 8) * 3.0 - 3.0 * (n // 8) / 2.0 + 1.5, 10.0])
        rot = rotation_transform(r.uniform(3, 0.0, 360.0))
        if i % 3 == 2:
            R, rr = 1.0e4, 200.0
            t = compose(compose(scale_transform(1e-4), rot), translation_transform(pos))
        else:
            R = float(r.uniform(1, 0.5, 1.5)[0])
            rr = R * float(r.uniform(1, 0.1, 0.3)[0])
            t = compose(rot, translation_transform(pos))
        R_.append(R)
        r_.append(rr)
        xf.append(_xf_row(t))
    objects = np.zeros((n, 4), np.uint32)
    objects[:, 0] = A.PVX_TORUS
    objects[:, 1] = np.arange(n)
    alb = r.uniform(3 * n).reshape(n, 3)
    return Scene(
        name=f"tori{n}", objects=objects,
        torus_major=np.array(R_), torus_minor=np.array(r_), torus_transform=np.arange(n, dtype=np.int32),
        torus_sturm=np.full(n, 1 if sturm else 0, np.uint8), transforms=np.array(xf).reshape(-1, 32),
        camera=look_at_camera((0.0, 0.0, -8.0), (0.0, 0.0, 10.0), width, height),
        lights=[dict(kind="point", position=np.array([30.0, 40.0, -30.0]), colour=np.ones(3))],
        materials=dict(pigment=alb, diffuse=np.full(n, 0.6), ambient=np.full(n, 0.1),
                       phong=np.full(n, 0.3), phong_size=np.full(n, 40.0)),
    )


def torus_plane_rays(scene: Scene, per_torus: int = 16, distance: float = 6.0) -> np.ndarray:
    """Rays lying in each torus's own plane and passing through its centre,
    from `distance` away.  Their quartic is symmetric about the centre (a
    biquadratic after the depressing shift), so the resolvent cubic's root
    makes d1 = 2z - p vanish up to rounding: the SMALL_ENOUGH clamp of
    solve_quartic (polynomialsolver.cpp:1325-1450)."""
    out = []
    for row in scene.transforms[scene.torus_transform]:
        m = row[:16].reshape(4, 4)
        c = m[3, :3]
        for a in np.linspace(0.0, np.pi, per_torus, endpoint=False):
            u = np.cos(a) * m[0, :3] + np.sin(a) * m[2, :3]
            u = u / np.linalg.norm(u)
            out.append(np.concatenate([c + distance * u, -u]))
    return np.array(out, np.float64)


# --------------------------------------------------------------------------
# radiosity room
# --------------------------------------------------------------------------
def radiosity_scene(seed: int = 0xC0FFEE + 3, n_boxes: int = 5000, width=1280, height=720) -> Scene:
    """A closed 10x4x10 room (six planes) with `n_boxes` seeded boxes
    (extents U[0.05,0.8], albedo U[0.2,0.9]^3, random rotation about y) and
    one emissive ceiling panel; lit only by radiosity
    (count 200, recursion_limit 2, error_bound 0.5)."""
    r = Rng(seed)
    ext = r.uniform(3 * n_boxes, 0.05, 0.8).reshape(n_boxes, 3)
    # boxes clutter the floor (resting on it or slightly raised), leaving a
    # 1.5 m strip in front of the camera free
    pos = np.stack([r.uniform(n_boxes, -4.6, 4.6), 0.5 * ext[:, 1] + r.uniform(n_boxes, 0.0, 0.5),
                    r.uniform(n_boxes, -3.0, 4.6)], 1)
    rot = r.uniform(n_boxes, 0.0, 360.0)
    alb = r.uniform(3 * n_boxes, 0.2, 0.9).reshape(n_boxes, 3)
    xforms = []
    c1, c2, tr = [], [], []
    for i in range(n_boxes):
        t = compose(rotation_transform((0.0, rot[i], 0.0)), translation_transform(pos[i]))
        xforms.append(_xf_row(t))
        c1.append(-0.5 * ext[i])
        c2.append(0.5 * ext[i])
        tr.append(i)
    # the emissive ceiling panel (untransformed)
    c1.append([-1.5, 3.9, -1.5])
    c2.append([1.5, 3.99, 1.5])
    tr.append(-1)
    planes_n = np.array([[0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0], [0, 0, 1], [0, 0, -1]],
                        np.float64)
    planes_d = np.array([0.0, 4.0, 5.0, 5.0, 5.0, 5.0])
    nb = n_boxes + 1
    objects = np.zeros((nb + 6, 4), np.uint32)
    objects[:nb, 0] = A.PVX_BOX
    objects[:nb, 1] = np.arange(nb)
    objects[nb:, 0] = A.PVX_PLANE
    objects[nb:, 1] = np.arange(6)
    n = nb + 6
    pig = np.concatenate([alb, np.ones((1, 3)), np.full((6, 3), 0.75)], 0)
    emission = np.zeros((n, 3))
    emission[n_boxes] = [6.0, 6.0, 6.0]
    diffuse = np.full(n, 0.75)
    diffuse[n_boxes] = 0.0
    return Scene(
        name=f"radiosity_room_{n_boxes}", objects=objects,
        box_corner1=np.array(c1), box_corner2=np.array(c2), box_transform=np.array(tr, np.int32),
        transforms=np.array(xforms).reshape(-1, 32),
        plane_normal=planes_n, plane_distance=planes_d,
        camera=look_at_camera((0.0, 2.4, -4.7), (0.0, 0.5, 0.5), width, height),
        lights=[],
        materials=dict(pigment=pig, diffuse=diffuse, ambient=np.zeros(n), phong=np.zeros(n),
                       phong_size=np.full(n, 40.0), reflection=np.zeros(n), ior=np.ones(n),
                       transmit=np.zeros(n), emission=emission),
        settings=dict(max_trace_level=5, adc_bailout=1.0 / 255.0,
                      radiosity=dict(count=200, recursion_limit=2, error_bound=0.5,
                                     minimum_reuse=0.015, maximum_reuse=0.2, nearest_count=5,
                                     pretrace_start=0.08, pretrace_end=0.04, brightness=1.0,
                                     adc_bailout=0.01, probe_threshold=0.02, train_samples=6144,
                                     train_steps=100, probe_tile=64, probes_per_tile=1)),
    )


# --------------------------------------------------------------------------
# multi-scene batch
# --------------------------------------------------------------------------
def terrain_mesh(n_tris: int, seed: int, half: float = 10.0, height: float = 1.5):
    """Heightfield over [-half, half]^2 (y up): a grid of nx x nz cells, two
    triangles each (nx * nz * 2 ~ n_tris, at least 2), heights from 4-octave
    value noise, smoothed vertex normals.  Returns (verts f32, idx i32,
    normals f32)."""
    cells = max(1, int(round(n_tris / 2)))
    nx = max(1, int(round(np.sqrt(cells))))
    nz = max(1, int(round(cells / nx)))
    xs = np.linspace(-half, half, nx + 1)
    zs = np.linspace(-half, half, nz + 1)
    gx, gz = np.meshgrid(xs, zs, indexing="ij")
    p = np.stack([gx.ravel(), np.zeros(gx.size), gz.ravel()], 1)
    y = value_noise(p * (1.0 / half), seed, octaves=4, base_freq=1.5) * height
    v = np.stack([p[:, 0], y, p[:, 2]], 1)
    i = np.arange(nx)[:, None] * (nz + 1) + np.arange(nz)[None, :]
    a, b, c, d = i.ravel(), i.ravel() + (nz + 1), i.ravel() + 1, i.ravel() + nz + 2
    # counter-clockwise seen from above (normals +y)
    f = np.concatenate([np.stack([a, c, b], 1), np.stack([c, d, b], 1)], 0).astype(np.int32)
    v32 = v.astype(np.float32)
    n32 = vertex_normals(v32.astype(np.float64), f).astype(np.float32)
    return v32, f, n32


def batch_scene(seed: int, n_tris: int, n_spheres: int, width: int, height: int, name: str = None) -> Scene:
    """One scene of the multi-scene batch: a terrain mesh of about `n_tris`
    triangles plus `n_spheres` spheres above it (centres U[-9,9] x U[0.3,4] x
    U[-9,9], radii U[0.1,0.6]), its own camera and one point light."""
    r = Rng(seed)
    v, f, nrm = terrain_mesh(n_tris, seed ^ 0x5EED)
    centre = np.stack([r.uniform(n_spheres, -9.0, 9.0), r.uniform(n_spheres, 0.3, 4.0),
                       r.uniform(n_spheres, -9.0, 9.0)], 1) if n_spheres else np.zeros((0, 3))
    radius = r.uniform(n_spheres, 0.1, 0.6) if n_spheres else np.zeros(0)
    n = 1 + n_spheres
    objects = np.zeros((n, 4), np.uint32)
    objects[0, 0] = A.PVX_MESH
    objects[1:, 0] = A.PVX_SPHERE
    objects[1:, 1] = np.arange(n_spheres)
    ang = r.uniform(1, 0.0, 2.0 * np.pi)[0]
    loc = (13.0 * np.sin(ang), r.uniform(1, 5.0, 9.0)[0], -13.0 * np.cos(ang))
    alb = r.uniform(3 * n, 0.2, 0.9).reshape(n, 3)
    refl = np.where(r.uniform(n) < 0.1, 0.3, 0.0)
    refl[0] = 0.0
    return Scene(
        name=name or f"batch_{n_tris}t_{n_spheres}s_{width}x{height}", objects=objects,
        sphere_center=centre if n_spheres else None, sphere_radius=radius if n_spheres else None,
        meshes=[(v, f, nrm)],
        camera=look_at_camera(loc, (0.0, 0.0, 0.0), width, height),
        lights=[dict(kind="point", position=np.array([20.0, 30.0, -20.0]), colour=np.ones(3))],
        materials=dict(pigment=alb, diffuse=np.full(n, 0.6), ambient=np.full(n, 0.1),
                       phong=np.full(n, 0.4), phong_size=np.full(n, 40.0), reflection=refl,
                       ior=np.ones(n), transmit=np.zeros(n)),
        settings=dict(max_trace_level=5, adc_bailout=1.0 / 255.0),
    )


def batch_plan(seed: int = 0xC0FFEE + 4, n_scenes: int = 64, tris=(1000, 500000), max_spheres: int = 2000,
               res_lo=(256, 256), res_hi=(1920, 1080)):
    """Sizes of the multi-scene batch: per scene (seed, triangles drawn
    log-uniform in `tris`, spheres uniform in 0..max_spheres, width and height
    uniform in [res_lo, res_hi])."""
    r = Rng(seed)
    lt = r.uniform(n_scenes, np.log(tris[0]), np.log(tris[1]))
    ns = np.minimum((r.uniform(n_scenes) * (max_spheres + 1)).astype(np.int64), max_spheres)
    w = np.rint(r.uniform(n_scenes, res_lo[0], res_hi[0])).astype(np.int64)
    h = np.rint(r.uniform(n_scenes, res_lo[1], res_hi[1])).astype(np.int64)
    seeds = (r.bits(n_scenes) >> np.uint64(1)).astype(np.int64)
    return [dict(seed=int(seeds[k]), n_tris=int(round(np.exp(lt[k]))), n_spheres=int(ns[k]),
                 width=int(w[k]), height=int(h[k])) for k in range(n_scenes)]


def batch_radiosity_scene(seed: int, n_tris: int, n_spheres: int, width: int, height: int,
                          name: str = None) -> Scene:
    """A batch scene lit by radiosity: the terrain and spheres of
    `batch_scene` with plain diffuse finishes (diffuse 0.75, no reflection)
    under an emissive ceiling panel 6 units above the terrain (a 12 x 12 box,
    emission 3); count 50, recursion_limit 2, error_bound 0.5, learned mode
    settings of the radiosity room scaled to the frame (train_samples = a
    twentieth of the pixels, at least 1,024)."""
    b = batch_scene(seed, n_tris, n_spheres, width, height, name)
    n0 = b.n_objects
    objects = np.concatenate([b.objects, np.array([[A.PVX_BOX, 0, 0, 0]], np.uint32)])
    n = n0 + 1
    pig = np.concatenate([b.materials["pigment"], np.ones((1, 3))])
    emission = np.zeros((n, 3))
    emission[n0] = 3.0
    diffuse = np.full(n, 0.75)
    diffuse[n0] = 0.0
    px = width * height
    return Scene(
        name=b.name + "_rad", objects=objects, sphere_center=b.sphere_center, sphere_radius=b.sphere_radius,
        meshes=b.meshes, box_corner1=np.array([[-6.0, 6.0, -6.0]]), box_corner2=np.array([[6.0, 6.1, 6.0]]),
        box_transform=np.array([-1], np.int32), camera=b.camera, lights=[],
        materials=dict(pigment=pig, diffuse=diffuse, ambient=np.zeros(n), phong=np.zeros(n),
                       phong_size=np.full(n, 40.0), reflection=np.zeros(n), ior=np.ones(n),
                       transmit=np.zeros(n), emission=emission),
        settings=dict(max_trace_level=5, adc_bailout=1.0 / 255.0,
                      radiosity=dict(count=50, recursion_limit=2, error_bound=0.5, minimum_reuse=0.015,
                                     maximum_reuse=0.2, nearest_count=5, pretrace_start=0.08, pretrace_end=0.04,
                                     brightness=1.0, adc_bailout=0.01, probe_threshold=0.02,
                                     train_samples=max(1024, px // 20), train_steps=100, probe_tile=64,
                                     probes_per_tile=1)),
    )


def batch_scenes(plan=None, **kw):
    """The multi-scene batch: 64 seeded scenes of uneven size (terrain meshes
    of 1k to 500k triangles drawn log-uniform, 0 to 2k spheres, resolutions
    from 256x256 to 1920x1080)."""
    plan = plan if plan is not None else batch_plan(**kw)
    return [batch_scene(p["seed"], p["n_tris"], p["n_spheres"], p["width"], p["height"],
                        name=f"batch{k:02d}") for k, p in enumerate(plan)]


# --------------------------------------------------------------------------
# stand-alone triangles
# --------------------------------------------------------------------------
def triangle_scene(seed: int = 0xC0FFEE + 5, n: int = 400, n_spheres: int = 40, width=320, height=240) -> Scene:
    """`n` seeded stand-alone triangles (every other one a smooth_triangle
    whose vertex normals are the face normal tilted by up to ~35 degrees;
    one in 16 of those with a vertex normal flipped, which the reference
    marks degenerate) in a 12-unit cube, `n_spheres` spheres, a ground
    plane, one point light; 10% of the triangles reflective."""
    r = Rng(seed)
    c = r.uniform(3 * n, -6.0, 6.0).reshape(n, 1, 3)
    e = r.uniform(9 * n, -1.2, 1.2).reshape(n, 3, 3)
    v = (c + e).reshape(n, 9)
    fn = np.cross(e[:, 0] - e[:, 1], e[:, 2] - e[:, 1])
    fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-300)
    tilt = r.uniform(9 * n, -0.7, 0.7).reshape(n, 3, 3)
    nv = fn[:, None, :] + tilt
    flip = (np.arange(n) % 32) == 31
    nv[flip, 2] *= -1.0
    smooth = (np.arange(n) % 2) == 1
    cs = r.uniform(3 * n_spheres, -6.0, 6.0).reshape(n_spheres, 3)
    rs = r.uniform(n_spheres, 0.2, 0.7)
    m = n + n_spheres + 1
    objects = np.zeros((m, 4), np.uint32)
    objects[:n, 0] = np.where(smooth, A.PVX_SMOOTH_TRIANGLE, A.PVX_TRIANGLE)
    objects[:n, 1] = np.arange(n)
    objects[n:n + n_spheres, 0] = A.PVX_SPHERE
    objects[n:n + n_spheres, 1] = np.arange(n_spheres)
    objects[-1, 0] = A.PVX_PLANE
    alb = r.uniform(3 * m, 0.2, 0.95).reshape(m, 3)
    refl = np.where(r.uniform(m) < 0.1, 0.3, 0.0)
    return Scene(
        name=f"triangles{n}", objects=objects, triangle_vertices=v, triangle_normals=nv.reshape(n, 9),
        sphere_center=cs, sphere_radius=rs,
        plane_normal=np.array([[0.0, 1.0, 0.0]]), plane_distance=np.array([7.5]),
        camera=look_at_camera((0.0, 6.0, -20.0), (0.0, 0.0, 0.0), width, height),
        lights=[dict(kind="point", position=np.array([15.0, 25.0, -15.0]), colour=np.ones(3))],
        materials=dict(pigment=alb, diffuse=np.full(m, 0.6), ambient=np.full(m, 0.1), phong=np.full(m, 0.4),
                       phong_size=np.full(m, 40.0), reflection=refl, ior=np.ones(m), transmit=np.zeros(m)),
        settings=dict(max_trace_level=5, adc_bailout=1.0 / 255.0),
    )


# --------------------------------------------------------------------------
# cones and cylinders
# --------------------------------------------------------------------------
def cone_scene(seed: int = 0xC0FFEE + 6, n: int = 300, n_csg: int = 40, n_discs: int = 60, n_quadrics: int = 8,
               n_polygons: int = 10, width=320, height=240) -> Scene:
    """`n` seeded cones and cylinders in a 12-unit cube: a third `cylinder`
    statements, the rest cones with radii U[0, 0.8] (a pointed end at radius
    0 for one in eight; one in eight with equal radii, which the reference
    turns into a cylinder; either end may be the bigger); one in four `open`;
    one in three with a random rotation composed after the shape's own
    placement.  `n_csg` CSG objects: a box minus a cylinder (difference) or
    a cone intersected with a sphere, or (one in four) a sphere minus the
    half-space below a disc (the disc inverted).  `n_discs` discs (radius U[0.3, 1.2], a hole in one
    of three, a rotation composed after one in three).  `n_quadrics`
    ellipsoid quadrics (semi-axes U[0.6, 1.5], rotated and moved by their
    transform) and two CSG objects of unbounded quadrics: an infinite
    cylinder x^2 + z^2 = 0.25 cut by a box, a paraboloid y = x^2 + z^2 cut by
    a sphere.  `n_polygons` planar star polygons (5 to 8 tips, every other
    one with a square hole as a second sub-polygon, one in three rotated by
    its transform).  A ground plane, one point light; 10% reflective, 10%
    glass."""
    r = Rng(seed)
    m = n + 2 * n_csg
    apex = r.uniform(3 * m, -6.0, 6.0).reshape(m, 3)
    axis = r.uniform(3 * m, -1.0, 1.0).reshape(m, 3)
    axis /= np.maximum(np.linalg.norm(axis, axis=1, keepdims=True), 1e-9)
    base = apex + axis * r.uniform(m, 0.5, 2.5)[:, None]
    rad = r.uniform(2 * m, 0.0, 0.8).reshape(m, 2) + 0.05
    k = np.arange(m)
    rad[k % 8 == 3, 1] = 0.0            # pointed cone
    rad[k % 8 == 5, 1] = rad[k % 8 == 5, 0]  # equal radii: the cone becomes a cylinder
    flags = np.where(k % 3 == 0, A.PVX_CONE_CYLINDER, 0).astype(np.uint8)
    flags |= np.where(k % 4 == 1, A.PVX_CONE_OPEN, 0).astype(np.uint8)
    xforms, tr = [], np.full(m, -1, np.int32)
    for i in range(m):
        if i % 3 == 2:
            xforms.append(_xf_row(rotation_transform(r.uniform(3, 0.0, 360.0))))
            tr[i] = len(xforms) - 1
    # discs: the free-standing ones first, CSG ones appended below
    dc = list(r.uniform(3 * n_discs, -6.0, 6.0).reshape(n_discs, 3))
    dn = list(r.uniform(3 * n_discs, -1.0, 1.0).reshape(n_discs, 3))
    drad = r.uniform(n_discs, 0.3, 1.2)
    dr = [(float(drad[i]), float(drad[i]) * 0.5 if i % 3 == 1 else 0.0) for i in range(n_discs)]
    dt = []
    for i in range(n_discs):
        if i % 3 == 2:
            xforms.append(_xf_row(rotation_transform(r.uniform(3, 0.0, 360.0))))
            dt.append(len(xforms) - 1)
        else:
            dt.append(-1)
    # CSG pieces
    box_c1, box_c2, sph_c, sph_r = [], [], [], []
    csg_nodes, csg_children = [], []
    roots = []
    for j in range(n_csg):
        ca, cb = n + 2 * j, n + 2 * j + 1
        c = r.uniform(3, -5.0, 5.0)
        if j % 4 == 3:  # a sphere minus the half-space below a disc
            sph_c.append(c)
            sph_r.append(1.0)
            dc.append(c + r.uniform(3, -0.3, 0.3))
            dn.append(r.uniform(3, -1.0, 1.0))
            dr.append((1.5, 0.0))
            dt.append(-1)
            kids = [[A.PVX_SPHERE, len(sph_c) - 1, 0, 0], [A.PVX_DISC, len(dc) - 1, A.PVX_FLAG_INVERTED, 0]]
            op = A.PVX_CSG_INTERSECTION
        elif j % 2 == 0:  # box minus a cylinder through it
            e = r.uniform(3, 0.6, 1.4)
            box_c1.append(c - e)
            box_c2.append(c + e)
            apex[ca] = c - np.array([0.0, 2.0, 0.0])
            base[ca] = c + np.array([0.0, 2.0, 0.0])
            rad[ca] = 0.4
            flags[ca] = A.PVX_CONE_CYLINDER
            tr[ca] = -1
            kids = [[A.PVX_BOX, len(box_c1) - 1, 0, 0], [A.PVX_CONE, ca, A.PVX_FLAG_INVERTED, 0]]
            op = A.PVX_CSG_INTERSECTION
        else:  # cone intersected with a sphere
            apex[cb] = c + np.array([0.0, 1.2, 0.0])
            base[cb] = c - np.array([0.0, 1.2, 0.0])
            rad[cb] = (0.0, 1.1)
            flags[cb] = 0
            tr[cb] = -1
            sph_c.append(c)
            sph_r.append(1.0)
            kids = [[A.PVX_CONE, cb, 0, 0], [A.PVX_SPHERE, len(sph_c) - 1, 0, 0]]
            op = A.PVX_CSG_INTERSECTION
        first = len(csg_children)
        csg_children.extend(kids)
        csg_nodes.append([op, first, 2, 0])
        roots.append(len(csg_nodes) - 1)
    # quadrics: ellipsoids, then the two unbounded ones used in CSG
    quad, qtr = [], []
    for i in range(n_quadrics):
        ax = r.uniform(3, 0.6, 1.5)
        quad.append([1.0 / ax[0] ** 2, 1.0 / ax[1] ** 2, 1.0 / ax[2] ** 2, 0, 0, 0, 0, 0, 0, -1.0])
        t = compose(rotation_transform(r.uniform(3, 0.0, 360.0)), translation_transform(r.uniform(3, -5.0, 5.0)))
        xforms.append(_xf_row(t))
        qtr.append(len(xforms) - 1)
    qroots = []
    if n_quadrics:
        for k, (coef, c) in enumerate((([1.0, 0, 1.0, 0, 0, 0, 0, 0, 0, -0.25], np.array([-3.0, 1.0, -2.0])),
                                      ([1.0, 0, 1.0, 0, 0, 0, 0, -1.0, 0, 0.0], np.array([3.0, -1.0, -2.0])))):
            quad.append(coef)
            xforms.append(_xf_row(translation_transform(c)))
            qtr.append(len(xforms) - 1)
            if k == 0:
                box_c1.append(c - np.array([1.0, 1.5, 1.0]))
                box_c2.append(c + np.array([1.0, 1.5, 1.0]))
                other = [A.PVX_BOX, len(box_c1) - 1, 0, 0]
            else:
                sph_c.append(c + np.array([0.0, 1.0, 0.0]))
                sph_r.append(1.2)
                other = [A.PVX_SPHERE, len(sph_c) - 1, 0, 0]
            first = len(csg_children)
            csg_children.extend([[A.PVX_QUADRIC, len(quad) - 1, 0, 0], other])
            csg_nodes.append([A.PVX_CSG_INTERSECTION, first, 2, 0])
            qroots.append(len(csg_nodes) - 1)
    n_q = n_quadrics + len(qroots)
    # polygons
    pfirst, ppts, ptr = [0], [], []
    for k in range(n_polygons):
        c = r.uniform(3, -5.0, 5.0)
        a = r.uniform(3, -1.0, 1.0)
        a /= np.linalg.norm(a)
        u = np.cross(a, [0.0, 0.0, 1.0] if abs(a[2]) < 0.9 else [1.0, 0.0, 0.0])
        u /= np.linalg.norm(u)
        v = np.cross(a, u)
        tips = 5 + k % 4
        ang = np.arange(2 * tips) * np.pi / tips
        prad = np.where(np.arange(2 * tips) % 2 == 0, r.uniform(1, 0.8, 1.3)[0], r.uniform(1, 0.35, 0.6)[0])
        outer = [c + rr * (np.cos(t) * u + np.sin(t) * v) for t, rr in zip(ang, prad)]
        pts = outer + [outer[0]]
        if k % 2 == 1:  # a square hole
            sq = [c + 0.15 * (sx * u + sy * v) for sx, sy in ((-1, -1), (1, -1), (1, 1), (-1, 1))]
            pts += sq + [sq[0]]
        ppts.extend(pts)
        pfirst.append(len(ppts))
        if k % 3 == 2:
            xforms.append(_xf_row(rotation_transform(r.uniform(3, 0.0, 30.0))))
            ptr.append(len(xforms) - 1)
        else:
            ptr.append(-1)
    n_obj = n + n_csg + n_discs + n_q + n_polygons + 1
    objects = np.zeros((n_obj, 4), np.uint32)
    objects[:n, 0] = A.PVX_CONE
    objects[:n, 1] = np.arange(n)
    objects[n:n + n_csg, 0] = A.PVX_CSG
    objects[n:n + n_csg, 1] = roots
    objects[n + n_csg:n + n_csg + n_discs, 0] = A.PVX_DISC
    objects[n + n_csg:n + n_csg + n_discs, 1] = np.arange(n_discs)
    q0 = n + n_csg + n_discs
    objects[q0:q0 + n_quadrics, 0] = A.PVX_QUADRIC
    objects[q0:q0 + n_quadrics, 1] = np.arange(n_quadrics)
    objects[q0 + n_quadrics:q0 + n_q, 0] = A.PVX_CSG
    objects[q0 + n_quadrics:q0 + n_q, 1] = qroots
    p0 = q0 + n_q
    objects[p0:p0 + n_polygons, 0] = A.PVX_POLYGON
    objects[p0:p0 + n_polygons, 1] = np.arange(n_polygons)
    objects[-1, 0] = A.PVX_PLANE
    alb = r.uniform(3 * n_obj, 0.2, 0.95).reshape(n_obj, 3)
    u = r.uniform(n_obj)
    refl = np.where(u < 0.1, 0.3, 0.0)
    glass = (u >= 0.1) & (u < 0.2)
    return Scene(
        name=f"cones{n}", objects=objects,
        cone_apex=apex, cone_base=base, cone_radius=rad, cone_flags=flags, cone_transform=tr,
        disc_center=np.array(dc).reshape(-1, 3), disc_normal=np.array(dn).reshape(-1, 3),
        disc_radius=np.array(dr).reshape(-1, 2), disc_transform=np.array(dt, np.int32),
        quadric_coef=np.array(quad, float).reshape(-1, 10), quadric_transform=np.array(qtr, np.int32),
        polygon_first=np.array(pfirst, np.uint32), polygon_points=np.array(ppts, float).reshape(-1, 3),
        polygon_transform=np.array(ptr, np.int32),
        transforms=np.array(xforms).reshape(-1, 32),
        box_corner1=np.array(box_c1).reshape(-1, 3), box_corner2=np.array(box_c2).reshape(-1, 3),
        box_transform=np.full(len(box_c1), -1, np.int32),
        sphere_center=np.array(sph_c).reshape(-1, 3), sphere_radius=np.array(sph_r),
        plane_normal=np.array([[0.0, 1.0, 0.0]]), plane_distance=np.array([7.5]),
        csg=np.array(csg_nodes, np.uint32).reshape(-1, 4),
        csg_children=np.array(csg_children, np.uint32).reshape(-1, 4),
        camera=look_at_camera((0.0, 6.0, -20.0), (0.0, 0.0, 0.0), width, height),
        lights=[dict(kind="point", position=np.array([15.0, 25.0, -15.0]), colour=np.ones(3))],
        materials=dict(pigment=alb, diffuse=np.full(n_obj, 0.6), ambient=np.full(n_obj, 0.1),
                       phong=np.full(n_obj, 0.4), phong_size=np.full(n_obj, 40.0), reflection=refl,
                       ior=np.where(glass, 1.5, 1.0), transmit=np.where(glass, 0.9, 0.0)),
        settings=dict(max_trace_level=5, adc_bailout=1.0 / 255.0),
    )
