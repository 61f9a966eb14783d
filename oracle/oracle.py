"""ctypes wrapper of the FP64 CPU oracle — TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's CPU-baseline leg load this module;
the product (povray_amd) never imports it.  The library is built from the C++
restatement in this directory (`make -C oracle`).
"""
import ctypes as C
import os

import numpy as np

from povray_amd import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

BOUND_HUGE = 2.0e10


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [C.POINTER(A.PvxSceneDesc)]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_find_intersection.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                            C.c_size_t, C.c_int, C.c_void_p, C.c_int]
        L.orc_all_intersections.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                            C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_inside.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
        L.orc_tree_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int32)]
        L.orc_mesh_triangle.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.orc_mesh_normal.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_size_t,
                                      C.c_void_p]
        L.orc_render.argtypes = [C.c_void_p, C.POINTER(A.PvxShadingDesc), C.POINTER(A.PvxCamera),
                                 C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.orc_radiosity_render.argtypes = [C.c_void_p, C.POINTER(A.PvxShadingDesc),
                                           C.POINTER(A.PvxCamera),
                                           C.POINTER(A.PvxRadiositySettings), C.c_void_p,
                                           C.c_void_p]
        L.orc_solve_polynomial.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_void_p]
        L.orc_triangle_normal.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p]
        L.orc_hit_normals.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
        L.orc_trace_rays.argtypes = [C.c_void_p, C.POINTER(A.PvxShadingDesc), C.c_void_p, C.c_size_t,
                                     C.c_void_p, C.c_int]
        L.orc_halton_pool.argtypes = [C.c_void_p]
        L.orc_rad_open.argtypes = [C.c_void_p, C.POINTER(A.PvxShadingDesc), C.POINTER(A.PvxCamera),
                                   C.POINTER(A.PvxRadiositySettings)]
        L.orc_rad_open.restype = C.c_void_p
        L.orc_rad_close.argtypes = [C.c_void_p]
        L.orc_rad_close.restype = None
        L.orc_rad_ambient.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                      C.c_int, C.c_void_p, C.c_void_p]
        L.orc_rad_average.argtypes = [C.c_void_p, C.c_size_t, C.c_double, C.c_void_p, C.c_void_p,
                                      C.c_double, C.c_int, C.c_int, C.c_void_p]
        L.orc_box_intersect.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_object_bbox.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        L.orc_difficult_coeffs.argtypes = [C.c_int, C.c_void_p]
        L.orc_solve_quartic.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_polysolve.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        L.orc_solver_branches.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


HIT_DTYPE = np.dtype([("t", np.float64), ("obj", np.int32), ("prim", np.int32)])


class OracleScene:
    """The oracle's copy of a scene (its own mesh preparation and BBox trees)."""

    def __init__(self, scene):
        L = lib()
        self._desc, self._keep = scene.desc()
        self.h = L.orc_scene_create(C.byref(self._desc))
        if not self.h:
            raise RuntimeError("orc_scene_create failed")

    def close(self):
        if self.h:
            lib().orc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def find_intersection(self, rays, max_depth=None, ray_kind=A.PVX_RAY_IMAGE, method=1,
                          threads=os.cpu_count() or 1):
        """Trace::FindIntersection for (n, 6) float64 rays -> structured hits."""
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
        n = rays.shape[0]
        hits = np.zeros(n, HIT_DTYPE)
        md = None if max_depth is None else np.ascontiguousarray(max_depth, np.float64)
        lib().orc_find_intersection(self.h, _p(rays), None if md is None else _p(md),
                                    ray_kind, n, method, _p(hits), threads)
        return hits

    def all_intersections(self, rays, obj, max_hits=8):
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
        obj = np.ascontiguousarray(obj, np.uint32)
        n = rays.shape[0]
        d = np.zeros((n, max_hits))
        p = np.zeros((n, max_hits), np.int32)
        c = np.zeros(n, np.uint32)
        lib().orc_all_intersections(self.h, _p(rays), _p(obj), n, max_hits, _p(d), _p(p), _p(c))
        return d, p, c

    def inside(self, pts, obj):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        obj = np.ascontiguousarray(obj, np.uint32)
        out = np.zeros(pts.shape[0], np.uint8)
        lib().orc_inside(self.h, _p(pts), _p(obj), pts.shape[0], _p(out))
        return out.astype(bool)

    def tree_info(self, mesh=-1):
        n = C.c_int64()
        d = C.c_int32()
        lib().orc_tree_info(self.h, mesh, C.byref(n), C.byref(d))
        return n.value, d.value

    def mesh_triangle(self, mesh, tri):
        out = np.zeros(10)
        lib().orc_mesh_triangle(self.h, mesh, tri, _p(out))
        return out

    def render(self, scene, width=None, height=None, threads=os.cpu_count() or 1, method=1, offsets=None):
        """Plain-finish TraceRay over the camera; returns (h, w, 3) float32
        linear RGB and (closest rays, shadow rays) counters.  `offsets`
        ((S, 2) sub-pixel positions) averages S samples per pixel."""
        sd, keep = scene.shading()
        cam = scene.camera_struct(width=width, height=height)
        out = np.zeros((cam.height, cam.width, 3), np.float32)
        cnt = np.zeros(2, np.uint64)
        if offsets is None:
            ns, op = 0, None
        else:
            offsets = np.ascontiguousarray(offsets, np.float64).reshape(-1, 2)
            ns, op = offsets.shape[0], _p(offsets)
        lib().orc_render(self.h, C.byref(sd), C.byref(cam), _p(out), threads, _p(cnt), method, ns, op)
        del keep
        return out, (int(cnt[0]), int(cnt[1]))

    def radiosity_render(self, scene, width=None, height=None, **over):
        """Sequential FP64 radiosity (plain gathers, Ward cache) over the
        camera; returns ((h, w, 3) float32 RGB, counters dict)."""
        sd, keep = scene.shading()
        cam = scene.camera_struct(width=width, height=height)
        rs = A.radiosity_settings(scene, **over)
        out = np.zeros((cam.height, cam.width, 3), np.float32)
        cnt = np.zeros(5, np.uint64)
        lib().orc_radiosity_render(self.h, C.byref(sd), C.byref(cam), C.byref(rs), _p(out), _p(cnt))
        del keep
        return out, dict(gathers0=int(cnt[0]), gathers1=int(cnt[1]), rays=int(cnt[2]),
                         records0=int(cnt[3]), records1=int(cnt[4]))

    def radiosity_function(self, scene, **over):
        """A sequential FP64 RadiosityFunction over one cache (see OracleRadiosity)."""
        return OracleRadiosity(self, scene, **over)

    def trace_rays(self, scene, rays, method=1):
        """TraceRay (primary-ray flags) for (n, 6) rays -> (n, 3) float32 RGB."""
        sd, keep = scene.shading()
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
        out = np.zeros((rays.shape[0], 3), np.float32)
        lib().orc_trace_rays(self.h, C.byref(sd), _p(rays), rays.shape[0], _p(out), method)
        del keep
        return out

    def object_bbox(self, obj):
        """(lowerLeft (3,) float32, size (3,) float32, infinite flag) of an object."""
        out = np.zeros(6, np.float32)
        inf = lib().orc_object_bbox(self.h, obj, _p(out))
        return out[:3].copy(), out[3:].copy(), bool(inf)

    def mesh_normal(self, mesh, tri, pts):
        tri = np.ascontiguousarray(tri, np.int32)
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        out = np.zeros_like(pts)
        lib().orc_mesh_normal(self.h, mesh, _p(tri), _p(pts), pts.shape[0], _p(out))
        return out

    def triangle_normal(self, tri, pts):
        """Triangle / SmoothTriangle::Normal of stand-alone triangle `tri`."""
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        out = np.zeros_like(pts)
        lib().orc_triangle_normal(self.h, tri, _p(pts), pts.shape[0], _p(out))
        return out

    def hit_normals(self, rays):
        """Raw Object::Normal at each ray's closest hit (zeros on a miss)."""
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
        out = np.zeros((rays.shape[0], 3))
        lib().orc_hit_normals(self.h, _p(rays), rays.shape[0], _p(out))
        return out


def solve_polynomial(coeffs, sturm=True, epsilon=1e-4):
    """Solve_Polynomial: coeffs in descending powers (order = len-1)."""
    c = np.ascontiguousarray(coeffs, np.float64)
    r = np.zeros(max(4, c.shape[0]))
    n = lib().orc_solve_polynomial(c.shape[0] - 1, _p(c), 1 if sturm else 0, epsilon, _p(r))
    return np.sort(r[:n])


def halton_pool():
    """The radiosity gathers' direction pool, (1600, 3)."""
    out = np.zeros((1600, 3))
    n = lib().orc_halton_pool(_p(out))
    return out[:n]


def rad_average(records, error_bound, point, normal, bound, depth=0, final=False):
    """AverageNearBlock of one query over hand-built records: (n, 15) rows of
    point, normal, irradiance, harmonic-mean distance, nearest distance,
    direction to the nearest surface, depth.  Returns (irradiance (3,), weight sum)."""
    rec = np.ascontiguousarray(records, np.float64).reshape(-1, 15)
    p = np.ascontiguousarray(point, np.float64)
    nrm = np.ascontiguousarray(normal, np.float64)
    out = np.zeros(4)
    lib().orc_rad_average(_p(rec), rec.shape[0], error_bound, _p(p), _p(nrm), bound, depth,
                          1 if final else 0, _p(out))
    return out[:3], out[3]


def box_intersect(c1, c2, ray, transform=None):
    """Box::Intersect: (hit, depth1, depth2, side1, side2); sides 1..6 =
    -x, +x, -y, +y, -z, +z."""
    c1 = np.ascontiguousarray(c1, np.float64)
    c2 = np.ascontiguousarray(c2, np.float64)
    ray = np.ascontiguousarray(ray, np.float64)
    tr = None if transform is None else np.ascontiguousarray(transform, np.float64).reshape(32)
    out = np.zeros(4)
    hit = lib().orc_box_intersect(_p(c1), _p(c2), None if tr is None else _p(tr), _p(ray), _p(out))
    return bool(hit), out[0], out[1], int(out[2]), int(out[3])


def difficult_coeffs(c):
    c = np.ascontiguousarray(c, np.float64)
    return bool(lib().orc_difficult_coeffs(c.shape[0] - 1, _p(c)))


class OracleRadiosity:
    """RadiosityFunction::ComputeAmbient (radiosity.cpp:385-461) point by
    point in the caller's order over one cache that persists between calls."""

    def __init__(self, oscene, scene, **over):
        self.o = oscene
        self._sd, self._keep = scene.shading()
        cam = scene.camera_struct()
        rs = A.radiosity_settings(scene, **over)
        self.h = lib().orc_rad_open(oscene.h, C.byref(self._sd), C.byref(cam), C.byref(rs))

    def compute_ambient(self, ipoint, normal, weight=None, depth=0, final=True):
        p = np.ascontiguousarray(ipoint, np.float64).reshape(-1, 3)
        nr = np.ascontiguousarray(normal, np.float64).reshape(-1, 3)
        w = None if weight is None else np.ascontiguousarray(weight, np.float64)
        out = np.zeros((p.shape[0], 3), np.float32)
        cnt = np.zeros(5, np.uint64)
        lib().orc_rad_ambient(self.h, p.shape[0], _p(p), _p(nr), None if w is None else _p(w), depth,
                              1 if final else 0, _p(out), _p(cnt))
        self.counters = dict(gathers0=int(cnt[0]), gathers1=int(cnt[1]), rays=int(cnt[2]),
                             records0=int(cnt[3]), records1=int(cnt[4]))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_rad_close(self.h)
            self.h = None


def solver_branches():
    """Reads and clears how often the order-4 branches of Solve_Polynomial
    ran since the last call: dict of closed-form quartic, difficult_coeffs
    forcing Sturm, the quartic's d1 clamp to 0, and Sturm by request."""
    out = np.zeros(4, np.uint64)
    lib().orc_solver_branches(_p(out))
    return dict(quartic=int(out[0]), difficult=int(out[1]), d1_clamp=int(out[2]), sturm=int(out[3]))


def solve_quartic_vieta(c):
    """The closed-form quartic branch of Solve_Polynomial (unsorted roots)."""
    c = np.ascontiguousarray(c, np.float64)
    r = np.zeros(4)
    n = lib().orc_solve_quartic(_p(c), _p(r))
    return r[:n].copy()


def polysolve(c):
    """The Sturm-sequence branch of Solve_Polynomial (unsorted roots)."""
    c = np.ascontiguousarray(c, np.float64)
    r = np.zeros(max(4, c.shape[0]))
    n = lib().orc_polysolve(c.shape[0] - 1, _p(c), _p(r))
    return r[:n].copy()
