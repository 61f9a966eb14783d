"""Python host side of the pvx C-ABI (include/pvx.h).

Mirrors the reference's entry points for the hot path in batch form:
`GpuScene.find_intersection` is Trace::FindIntersection
(core/render/trace.cpp:285-344), `all_intersections` is
ObjectBase::All_Intersections and `inside` is ObjectBase::Inside
(core/scene/object.h:238-240).  Every call runs the HIP kernels in
libpvx.so; there is no CPU fallback — a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi as A
from . import hip

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpvx.so")
_L = None

HIT_DTYPE = np.dtype([("t", np.float64), ("obj", np.int32), ("prim", np.int32)])
BOUND_HUGE = 2.0e10


class PvxError(RuntimeError):
    pass


def lib():
    """Load libpvx.so (built in-tree by __graft_entry__.build / make)."""
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise PvxError(f"libpvx.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        hip._lib()  # bind the same HIP runtime first
        L = C.CDLL(LIB_PATH)
        vp, sz, u32 = C.c_void_p, C.c_size_t, C.c_uint32
        L.pvx_last_error.restype = C.c_char_p
        L.pvx_init.argtypes = [C.c_int]
        L.pvx_scene_create.argtypes = [C.POINTER(A.PvxSceneDesc), C.POINTER(A.PvxBuildOpts),
                                       C.POINTER(vp)]
        L.pvx_scene_destroy.argtypes = [vp]
        L.pvx_scene_get_stats.argtypes = [vp, C.POINTER(A.PvxSceneStats)]
        L.pvx_trace_closest.argtypes = [vp, vp, vp, u32, sz, vp, vp]
        L.pvx_trace_closest_counted.argtypes = [vp, vp, vp, u32, sz, vp, vp, vp, vp]
        L.pvx_trace_closest_noalloc.argtypes = [vp, vp, vp, u32, sz, vp, vp]
        L.pvx_all_intersections.argtypes = [vp, vp, vp, sz, u32, vp, vp, vp, vp]
        L.pvx_inside.argtypes = [vp, vp, vp, sz, vp, vp]
        L.pvx_camera_rays.argtypes = [C.POINTER(A.PvxCamera), vp, vp]
        L.pvx_camera_rays_tiled.argtypes = [C.POINTER(A.PvxCamera), C.c_uint32, C.c_uint32, vp, vp, vp]
        L.pvx_scene_set_render_tile.argtypes = [vp, C.c_uint32, C.c_uint32]
        L.pvx_copy.argtypes = [vp, vp, sz, vp]
        L.pvx_scene_set_shading.argtypes = [vp, C.POINTER(A.PvxShadingDesc)]
        L.pvx_render.argtypes = [vp, C.POINTER(A.PvxCamera), vp, C.POINTER(A.PvxRenderStats), vp]
        L.pvx_render_samples.argtypes = [vp, C.POINTER(A.PvxCamera), u32, vp, vp, C.POINTER(A.PvxRenderStats),
                                         vp]
        L.pvx_radiosity_render.argtypes = [vp, C.POINTER(A.PvxCamera), C.POINTER(A.PvxRadiositySettings),
                                           vp, C.POINTER(A.PvxRadiosityStats), vp]
        L.pvx_batch_create.argtypes = [C.POINTER(vp), u32, C.POINTER(vp)]
        L.pvx_batch_destroy.argtypes = [vp]
        L.pvx_batch_trace_closest.argtypes = [vp, vp, vp, vp, u32, vp, vp, vp]
        L.pvx_batch_trace_closest_counted.argtypes = [vp, vp, vp, vp, u32, vp, vp, vp, vp, vp]
        L.pvx_clock_rate.argtypes = [C.POINTER(C.c_double)]
        L.pvx_scene_autotune.argtypes = [vp, vp, sz, u32, C.POINTER(A.PvxTuneResult), vp]
        L.pvx_scene_set_traversal.argtypes = [vp, u32, u32, u32]
        L.pvx_radiosity_begin.argtypes = [vp, C.POINTER(A.PvxCamera), C.POINTER(A.PvxRadiositySettings), vp]
        L.pvx_radiosity_compute_ambient.argtypes = [vp, sz, vp, vp, vp, vp, u32, u32, vp,
                                                    C.POINTER(A.PvxRadiosityStats), vp]
        L.pvx_batch_radiosity_render.argtypes = [vp, C.POINTER(A.PvxCamera), C.POINTER(A.PvxRadiositySettings),
                                                 C.POINTER(vp), u32, C.POINTER(A.PvxRadiosityStats),
                                                 C.POINTER(C.c_double), C.POINTER(A.PvxBatchRadiosityStats)]
        L.pvx_radiosity_save.argtypes = [vp, C.c_char_p, C.POINTER(C.c_uint64)]
        L.pvx_radiosity_load.argtypes = [vp, C.c_char_p, C.POINTER(C.c_uint64)]
        L.pvx_render_autotune.argtypes = [vp, C.POINTER(A.PvxCamera), vp, C.POINTER(A.PvxTuneResult),
                                          C.POINTER(A.PvxRenderStats), vp]
        fp = C.POINTER(C.c_float)
        L.pvx_surrogate_create.argtypes = [fp, C.c_uint64, C.POINTER(vp)]
        L.pvx_surrogate_destroy.argtypes = [vp]
        L.pvx_surrogate_destroy.restype = None
        L.pvx_surrogate_get_params.argtypes = [vp, fp]
        L.pvx_surrogate_grad.argtypes = [vp, vp, vp, sz, fp, fp, vp]
        L.pvx_surrogate_train.argtypes = [vp, vp, vp, sz, u32, u32, C.c_float, fp, vp]
        L.pvx_surrogate_infer.argtypes = [vp, vp, sz, vp, vp]
        _L = L
    return _L


EXPORTED_SYMBOLS = (
    "pvx_last_error", "pvx_init", "pvx_scene_create", "pvx_scene_destroy",
    "pvx_scene_get_stats", "pvx_trace_closest", "pvx_trace_closest_counted",
    "pvx_all_intersections", "pvx_inside",
    "pvx_camera_rays", "pvx_camera_rays_tiled", "pvx_scene_set_render_tile", "pvx_copy", "pvx_scene_set_shading", "pvx_render", "pvx_render_samples",
    "pvx_radiosity_render", "pvx_surrogate_create", "pvx_surrogate_destroy",
    "pvx_surrogate_get_params", "pvx_surrogate_grad", "pvx_surrogate_train", "pvx_surrogate_infer",
    "pvx_batch_create", "pvx_batch_destroy", "pvx_batch_trace_closest", "pvx_batch_trace_closest_counted",
    "pvx_clock_rate", "pvx_scene_autotune", "pvx_scene_set_traversal",
    "pvx_render_autotune", "pvx_radiosity_save", "pvx_radiosity_load",
    "pvx_radiosity_begin", "pvx_radiosity_compute_ambient", "pvx_trace_closest_noalloc",
    "pvx_batch_radiosity_render",
)


def _check(rc, what):
    if rc != 0:
        msg = lib().pvx_last_error().decode()
        raise PvxError(f"{what} failed ({rc}): {msg}")


def init(device: int = 0):
    _check(lib().pvx_init(device), "pvx_init")


def _ptr(buf):
    if buf is None:
        return None
    return buf.value if isinstance(buf, hip.DeviceBuffer) else buf


class GpuScene:
    """A scene uploaded to the GPU with its BVH."""

    def __init__(self, scene, max_leaf_size: int = 0, sah_bins: int = 0, accel: int = A.PVX_ACCEL_BVH,
                 bsp_max_depth: int = 0, builder: int = A.PVX_BUILDER_HOST_SAH):
        L = lib()
        self.scene = scene
        desc, keep = scene.desc()
        opts = A.PvxBuildOpts()
        opts.max_leaf_size = max_leaf_size
        opts.sah_bins = sah_bins
        opts.accel = accel
        opts.bsp_max_depth = bsp_max_depth
        opts.builder = builder
        h = C.c_void_p()
        _check(L.pvx_scene_create(C.byref(desc), C.byref(opts), C.byref(h)), "pvx_scene_create")
        self.h = h
        del keep

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().pvx_scene_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self) -> dict:
        s = A.PvxSceneStats()
        _check(lib().pvx_scene_get_stats(self.h, C.byref(s)), "pvx_scene_get_stats")
        return {k: getattr(s, k) for k, _ in A.PvxSceneStats._fields_}

    # -- device-buffer API (what a C caller does) ---------------------------------
    def trace_closest(self, rays, n, hits, max_depth=None, ray_kind=A.PVX_RAY_IMAGE,
                      stream=None):
        _check(lib().pvx_trace_closest(self.h, _ptr(rays), _ptr(max_depth), ray_kind, n,
                                       _ptr(hits), stream.h if stream else None),
               "pvx_trace_closest")

    def trace_closest_counted(self, rays, n, hits, node_visits, prim_tests, max_depth=None,
                              ray_kind=A.PVX_RAY_IMAGE, stream=None):
        _check(lib().pvx_trace_closest_counted(self.h, _ptr(rays), _ptr(max_depth), ray_kind, n,
                                               _ptr(hits), _ptr(node_visits), _ptr(prim_tests),
                                               stream.h if stream else None),
               "pvx_trace_closest_counted")

    def autotune(self, rays, n=None, ray_kind=A.PVX_RAY_IMAGE, stream=None) -> dict:
        """pvx_scene_autotune over sample rays (a device buffer with n, or an
        (n, 6) array); later closest-hit calls of this ray class use the
        fastest variant and grid.  Returns the choice and its timings."""
        keep = None
        if not isinstance(rays, hip.DeviceBuffer):
            arr = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
            n = arr.shape[0]
            keep = rays = hip.DeviceBuffer.from_array(arr)
        r = A.PvxTuneResult()
        _check(lib().pvx_scene_autotune(self.h, _ptr(rays), n, ray_kind, C.byref(r),
                                        stream.h if stream else None), "pvx_scene_autotune")
        del keep
        return {k: getattr(r, k) for k, _ in A.PvxTuneResult._fields_}

    def set_traversal(self, variant=A.PVX_TRAVERSAL_DEFAULT, grid_blocks=0, ray_kind=A.PVX_RAY_IMAGE):
        _check(lib().pvx_scene_set_traversal(self.h, ray_kind, variant, grid_blocks), "pvx_scene_set_traversal")

    # -- numpy convenience (upload, run, download) -------------------------------
    def find_intersection(self, rays: np.ndarray, max_depth=None, ray_kind=A.PVX_RAY_IMAGE):
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
        n = rays.shape[0]
        out = np.zeros(n, HIT_DTYPE)
        if n == 0:
            return out
        d_rays = hip.DeviceBuffer.from_array(rays)
        d_hits = hip.DeviceBuffer(n * 16)
        d_md = None
        if max_depth is not None:
            d_md = hip.DeviceBuffer.from_array(np.ascontiguousarray(max_depth, np.float64))
        self.trace_closest(d_rays, n, d_hits, d_md, ray_kind)
        hip.synchronize()
        raw = d_hits.download(np.uint8, n * 16)
        return raw.view(HIT_DTYPE).copy()

    def all_intersections(self, rays: np.ndarray, obj, max_hits: int = 8):
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
        n = rays.shape[0]
        obj = np.ascontiguousarray(obj, np.uint32)
        d_r = hip.DeviceBuffer.from_array(rays)
        d_o = hip.DeviceBuffer.from_array(obj)
        d_d = hip.DeviceBuffer(n * max_hits * 8)
        d_p = hip.DeviceBuffer(n * max_hits * 4)
        d_c = hip.DeviceBuffer(n * 4)
        _check(lib().pvx_all_intersections(self.h, d_r.value, d_o.value, n, max_hits, d_d.value,
                                           d_p.value, d_c.value, None), "pvx_all_intersections")
        hip.synchronize()
        return (d_d.download(np.float64).reshape(n, max_hits),
                d_p.download(np.int32).reshape(n, max_hits), d_c.download(np.uint32))

    def set_shading(self, scene=None):
        """Attach the scene's materials and lights (pvx_scene_set_shading)."""
        sd, keep = (scene or self.scene).shading()
        _check(lib().pvx_scene_set_shading(self.h, C.byref(sd)), "pvx_scene_set_shading")
        del keep

    def render_device(self, out, width=None, height=None, stream=None):
        """Render into a device buffer of 3*w*h floats; returns the stats dict."""
        cam = self.scene.camera_struct(width=width, height=height)
        st = A.PvxRenderStats()
        _check(lib().pvx_render(self.h, C.byref(cam), _ptr(out), C.byref(st),
                                stream.h if stream else None), "pvx_render")
        return {k: getattr(st, k) for k, _ in A.PvxRenderStats._fields_ if k != "pad"}

    def render_samples(self, offsets, width=None, height=None, stream=None):
        """Supersampled frame (pvx_render_samples); offsets: (S, 2) sub-pixel
        positions from the pixel corner.  Returns ((h, w, 3) RGB, stats)."""
        cam = self.scene.camera_struct(width=width, height=height)
        offsets = np.ascontiguousarray(offsets, np.float64).reshape(-1, 2)
        out = hip.DeviceBuffer(cam.width * cam.height * 12)
        st = A.PvxRenderStats()
        _check(lib().pvx_render_samples(self.h, C.byref(cam), offsets.shape[0], offsets.ctypes.data, out.value,
                                        C.byref(st), stream.h if stream else None), "pvx_render_samples")
        hip.synchronize()
        img = out.download(np.float32, cam.width * cam.height * 3)
        return (img.reshape(cam.height, cam.width, 3),
                {k: getattr(st, k) for k, _ in A.PvxRenderStats._fields_ if k != "pad"})

    def set_render_tile(self, tile_w=8, tile_h=8):
        """Pixel order of pvx_render's ray queue (pvx_scene_set_render_tile;
        tile_w 0 with tile_h 1 is row-major)."""
        _check(lib().pvx_scene_set_render_tile(self.h, tile_w, tile_h), "pvx_scene_set_render_tile")

    def render_autotune(self, width=None, height=None):
        """pvx_render_autotune: one frame that tunes the queue's pixel tile and
        the closest-hit walk on its camera rays, and the light walks on its
        light rays.  Returns (image, stats, [camera, light] choices)."""
        cam = self.scene.camera_struct(width=width, height=height)
        out = hip.DeviceBuffer(cam.width * cam.height * 12)
        st = A.PvxRenderStats()
        tune = (A.PvxTuneResult * 2)()
        _check(lib().pvx_render_autotune(self.h, C.byref(cam), out.value, tune, C.byref(st), None),
               "pvx_render_autotune")
        hip.synchronize()
        img = out.download(np.float32, cam.width * cam.height * 3)
        return (img.reshape(cam.height, cam.width, 3),
                {k: getattr(st, k) for k, _ in A.PvxRenderStats._fields_ if k != "pad"},
                [{k: getattr(t, k) for k, _ in A.PvxTuneResult._fields_} for t in tune])

    def render(self, width=None, height=None):
        """Render one frame; returns ((h, w, 3) float32 linear RGB, stats)."""
        cam = self.scene.camera_struct(width=width, height=height)
        out = hip.DeviceBuffer(cam.width * cam.height * 12)
        stats = self.render_device(out, width, height)
        hip.synchronize()
        img = out.download(np.float32, cam.width * cam.height * 3)
        return img.reshape(cam.height, cam.width, 3), stats

    def radiosity_render_device(self, out, width=None, height=None, mode=A.PVX_RADIOSITY_PLAIN,
                                stream=None, **over):
        """Radiosity frame into a device buffer of 3*w*h floats (pvx_radiosity_render);
        returns the stats dict."""
        cam = self.scene.camera_struct(width=width, height=height)
        rs = A.radiosity_settings(self.scene, mode=mode, **over)
        st = A.PvxRadiosityStats()
        _check(lib().pvx_radiosity_render(self.h, C.byref(cam), C.byref(rs), _ptr(out), C.byref(st),
                                          stream.h if stream else None), "pvx_radiosity_render")
        d = {}
        for k, _ in A.PvxRadiosityStats._fields_:
            v = getattr(st, k)
            d[k] = list(v) if hasattr(v, "__len__") else v
        return d

    def radiosity_render(self, width=None, height=None, mode=A.PVX_RADIOSITY_PLAIN, **over):
        """Render one radiosity frame; returns ((h, w, 3) float32 RGB, stats)."""
        cam = self.scene.camera_struct(width=width, height=height)
        out = hip.DeviceBuffer(cam.width * cam.height * 12)
        stats = self.radiosity_render_device(out, width, height, mode=mode, **over)
        hip.synchronize()
        img = out.download(np.float32, cam.width * cam.height * 3)
        return img.reshape(cam.height, cam.width, 3), stats

    def radiosity_function(self, **over) -> "RadiosityFunction":
        """Open a per-query irradiance cache on this scene (pvx_radiosity_begin)
        with the scene's radiosity settings and overrides; see RadiosityFunction."""
        return RadiosityFunction(self, **over)

    def radiosity_save(self, path) -> int:
        """Write the last radiosity frame's cache as a .rca file; returns the record count."""
        n = C.c_uint64()
        _check(lib().pvx_radiosity_save(self.h, str(path).encode(), C.byref(n)), "pvx_radiosity_save")
        return n.value

    def radiosity_load(self, path) -> int:
        """Seed later radiosity frames with a .rca file's records (None drops them)."""
        n = C.c_uint64()
        _check(lib().pvx_radiosity_load(self.h, None if path is None else str(path).encode(), C.byref(n)),
               "pvx_radiosity_load")
        return n.value

    def inside(self, pts: np.ndarray, obj):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        n = pts.shape[0]
        obj = np.ascontiguousarray(obj, np.uint32)
        d_p = hip.DeviceBuffer.from_array(pts)
        d_o = hip.DeviceBuffer.from_array(obj)
        d_out = hip.DeviceBuffer(max(n, 1))
        _check(lib().pvx_inside(self.h, d_p.value, d_o.value, n, d_out.value, None), "pvx_inside")
        hip.synchronize()
        return d_out.download(np.uint8, n).astype(bool)


class GpuBatch:
    """Scenes traced together, one launch for the rays of all of them
    (pvx_batch_*).  Keeps the GpuScenes alive."""

    def __init__(self, scenes):
        self.scenes = list(scenes)
        arr = (C.c_void_p * len(self.scenes))(*[s.h.value for s in self.scenes])
        h = C.c_void_p()
        _check(lib().pvx_batch_create(arr, len(self.scenes), C.byref(h)), "pvx_batch_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().pvx_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def radiosity_render(self, mode=A.PVX_RADIOSITY_LEARNED, n_streams: int = 4, images=True, **over):
        """Radiosity frames of every scene (pvx_batch_radiosity_render) with
        each scene's camera and radiosity settings plus the overrides.
        Returns (list of (h, w, 3) images or None, per-scene stats dicts,
        per-scene completion times in ms, batch stats dict)."""
        n = len(self.scenes)
        cams = (A.PvxCamera * n)(*[g.scene.camera_struct() for g in self.scenes])
        sets = (A.PvxRadiositySettings * n)(*[A.radiosity_settings(g.scene, mode=mode, **over)
                                               for g in self.scenes])
        outs = [hip.DeviceBuffer(c.width * c.height * 12) for c in cams]
        ptrs = (C.c_void_p * n)(*[o.value for o in outs])
        sst = (A.PvxRadiosityStats * n)()
        done = (C.c_double * n)()
        bst = A.PvxBatchRadiosityStats()
        _check(lib().pvx_batch_radiosity_render(self.h, cams, sets, ptrs, n_streams, sst, done, C.byref(bst)),
               "pvx_batch_radiosity_render")
        hip.synchronize()

        def todict(st):
            d = {}
            for k, _ in st._fields_:
                v = getattr(st, k)
                d[k] = list(v) if hasattr(v, "__len__") else v
            return d

        imgs = None
        if images:
            imgs = [o.download(np.float32, c.width * c.height * 3).reshape(c.height, c.width, 3)
                    for o, c in zip(outs, cams)]
        return imgs, [todict(x) for x in sst], list(done), todict(bst)

    def _first(self, ray_first):
        first = np.ascontiguousarray(ray_first, np.uint64)
        if first.shape != (len(self.scenes) + 1,):
            raise ValueError("ray_first needs n_scenes + 1 entries")
        return first

    # -- device-buffer API ----------------------------------------------------------
    def trace_closest(self, rays, ray_first, hits, max_depth=None, ray_kind=A.PVX_RAY_IMAGE,
                      scene_clock=None, stream=None):
        first = self._first(ray_first)
        _check(lib().pvx_batch_trace_closest(self.h, _ptr(rays), first.ctypes.data, _ptr(max_depth), ray_kind,
                                             _ptr(hits), _ptr(scene_clock), stream.h if stream else None),
               "pvx_batch_trace_closest")

    def trace_closest_counted(self, rays, ray_first, hits, node_visits, prim_tests, max_depth=None,
                              ray_kind=A.PVX_RAY_IMAGE, scene_clock=None, stream=None):
        first = self._first(ray_first)
        _check(lib().pvx_batch_trace_closest_counted(self.h, _ptr(rays), first.ctypes.data, _ptr(max_depth),
                                                     ray_kind, _ptr(hits), _ptr(node_visits), _ptr(prim_tests),
                                                     _ptr(scene_clock), stream.h if stream else None),
               "pvx_batch_trace_closest_counted")

    # -- numpy convenience ------------------------------------------------------------
    def find_intersection(self, rays_per_scene, ray_kind=A.PVX_RAY_IMAGE, max_depth_per_scene=None,
                          with_clock=False):
        """rays_per_scene: one (n_s, 6) array per scene.  Returns the list of
        hit arrays (and, with_clock, the n_scenes + 1 clock ticks)."""
        parts = [np.ascontiguousarray(r, np.float64).reshape(-1, 6) for r in rays_per_scene]
        if len(parts) != len(self.scenes):
            raise ValueError("one ray array per scene")
        counts = np.array([p.shape[0] for p in parts], np.uint64)
        first = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
        n = int(first[-1])
        rays = np.concatenate(parts, 0) if n else np.zeros((0, 6))
        d_rays = hip.DeviceBuffer.from_array(rays) if n else None
        d_hits = hip.DeviceBuffer(max(n, 1) * 16)
        d_md = None
        if max_depth_per_scene is not None:
            md = np.concatenate([np.asarray(m, np.float64).ravel() for m in max_depth_per_scene])
            d_md = hip.DeviceBuffer.from_array(md) if n else None
        d_clk = hip.DeviceBuffer(8 * (len(self.scenes) + 1)) if with_clock else None
        self.trace_closest(d_rays, first, d_hits, d_md, ray_kind, d_clk)
        hip.synchronize()
        out = d_hits.download(np.uint8, n * 16).view(HIT_DTYPE).copy() if n else np.zeros(0, HIT_DTYPE)
        res = [out[int(first[k]):int(first[k + 1])] for k in range(len(self.scenes))]
        if with_clock:
            return res, d_clk.download(np.uint64, len(self.scenes) + 1)
        return res


def clock_rate() -> float:
    """Ticks per second of the device wall clock (scene_clock units)."""
    hz = C.c_double()
    _check(lib().pvx_clock_rate(C.byref(hz)), "pvx_clock_rate")
    return hz.value


class Surrogate:
    """The irradiance MLP (pvx_surrogate_*): 6 inputs -> RGB."""

    N_PARAMS = 20867

    def __init__(self, params=None, seed: int = 1):
        h = C.c_void_p()
        fp = C.POINTER(C.c_float)
        if params is not None:
            params = np.ascontiguousarray(params, np.float32)
            assert params.size == self.N_PARAMS
            ptr = params.ctypes.data_as(fp)
        else:
            ptr = None
        _check(lib().pvx_surrogate_create(ptr, seed, C.byref(h)), "pvx_surrogate_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().pvx_surrogate_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.N_PARAMS, np.float32)
        _check(lib().pvx_surrogate_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))),
               "pvx_surrogate_get_params")
        return out

    def grad(self, x: np.ndarray, t: np.ndarray):
        """(loss, gradient) over the points (x: (n,6), t: (n,3))."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 6)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 3)
        dx, dt = hip.DeviceBuffer.from_array(x), hip.DeviceBuffer.from_array(t)
        loss = C.c_float()
        g = np.zeros(self.N_PARAMS, np.float32)
        _check(lib().pvx_surrogate_grad(self.h, dx.value, dt.value, x.shape[0], C.byref(loss),
                                        g.ctypes.data_as(C.POINTER(C.c_float)), None), "pvx_surrogate_grad")
        return float(loss.value), g

    def train(self, x, t, steps: int, batch: int = 4096, lr: float = 5e-3) -> float:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 6)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 3)
        dx, dt = hip.DeviceBuffer.from_array(x), hip.DeviceBuffer.from_array(t)
        loss = C.c_float()
        _check(lib().pvx_surrogate_train(self.h, dx.value, dt.value, x.shape[0], steps, batch, lr,
                                         C.byref(loss), None), "pvx_surrogate_train")
        hip.synchronize()
        return float(loss.value)

    def infer(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 6)
        n = x.shape[0]
        dx = hip.DeviceBuffer.from_array(x)
        dy = hip.DeviceBuffer(max(1, n) * 12)
        _check(lib().pvx_surrogate_infer(self.h, dx.value, n, dy.value, None), "pvx_surrogate_infer")
        hip.synchronize()
        return dy.download(np.float32, n * 3).reshape(n, 3)


def m1_offsets(aa_depth: int = 2):
    """Sample positions of non-adaptive supersampling at threshold 0 without
    jitter (TraceTask::SupersampleOnePixel, tracetask.cpp:861-890): the pixel
    centre, then an aa_depth x aa_depth grid (rows outer)."""
    step = 1.0 / aa_depth
    rng = 0.5 - step * 0.5
    grid = [rng * -1 + k * step for k in range(aa_depth)]
    return np.array([[0.5, 0.5]] + [[0.5 + xx, 0.5 + yy] for yy in grid for xx in grid])


def camera_rays(scene, width=None, height=None, jitter=(0.5, 0.5), out=None, stream=None, tile=None,
                pixel=None):
    """Primary rays for the scene's camera, generated on the GPU: row-major,
    or with tile=(w, h) in tile order (pvx_camera_rays_tiled; `pixel`, a
    device buffer of n uint32, receives each ray's pixel index)."""
    cam = scene.camera_struct(jitter=jitter, width=width, height=height)
    n = cam.width * cam.height
    if out is None:
        out = hip.DeviceBuffer(n * 48)
    if tile is None:
        _check(lib().pvx_camera_rays(C.byref(cam), out.value, stream.h if stream else None),
               "pvx_camera_rays")
    else:
        _check(lib().pvx_camera_rays_tiled(C.byref(cam), int(tile[0]), int(tile[1]), out.value,
                                           pixel.value if pixel is not None else None,
                                           stream.h if stream else None), "pvx_camera_rays_tiled")
    return out, n


def copy(dst, src, nbytes, stream=None):
    """Device copy kernel (byte-counter calibration workload)."""
    _check(lib().pvx_copy(_ptr(dst), _ptr(src), nbytes, stream.h if stream else None), "pvx_copy")


class RadiosityFunction:
    """RadiosityFunction::ComputeAmbient (core/lighting/radiosity.cpp:385-461)
    in batch form over the scene's device irradiance cache.

    Construction opens the cache (pvx_radiosity_begin: settings, camera
    position for the reuse distances); every `compute_ambient` call looks
    its points up, gathers where the cache cannot serve them (filing the new
    records, so later calls reuse them) and returns the ambient colour
    times `brightness`.  Only `mode` PLAIN is served per query."""

    def __init__(self, gscene: "GpuScene", **over):
        self.g = gscene
        cam = gscene.scene.camera_struct()
        rs = A.radiosity_settings(gscene.scene, mode=A.PVX_RADIOSITY_PLAIN, **over)
        _check(lib().pvx_radiosity_begin(gscene.h, C.byref(cam), C.byref(rs), None), "pvx_radiosity_begin")
        self.last_stats = None

    def compute_ambient(self, ipoint, raw_normal, layer_normal=None, weight=None, depth=0, final=True):
        """ipoint, raw_normal (, layer_normal): (n, 3); weight: (n,) or None
        (1.0); depth: the ticket's radiosity recursion depth; final: the
        final trace (isFinalTrace) rather than a pretrace.  Returns (n, 3)
        float32 ambient colours."""
        p = np.ascontiguousarray(ipoint, np.float64).reshape(-1, 3)
        n = p.shape[0]
        nr = np.ascontiguousarray(raw_normal, np.float64).reshape(n, 3)
        bufs = [hip.DeviceBuffer.from_array(p), hip.DeviceBuffer.from_array(nr)]
        ln = wt = None
        if layer_normal is not None:
            ln = hip.DeviceBuffer.from_array(np.ascontiguousarray(layer_normal, np.float64).reshape(n, 3))
        if weight is not None:
            wt = hip.DeviceBuffer.from_array(np.ascontiguousarray(weight, np.float64).reshape(n))
        out = hip.DeviceBuffer(max(1, n) * 12)
        st = A.PvxRadiosityStats()
        _check(lib().pvx_radiosity_compute_ambient(self.g.h, n, bufs[0].value, bufs[1].value,
                                                   ln.value if ln else None, wt.value if wt else None,
                                                   depth, 1 if final else 0, out.value, C.byref(st), None),
               "pvx_radiosity_compute_ambient")
        hip.synchronize()
        d = {}
        for k, _ in A.PvxRadiosityStats._fields_:
            v = getattr(st, k)
            d[k] = list(v) if hasattr(v, "__len__") else v
        self.last_stats = d
        return out.download(np.float32, n * 3).reshape(n, 3)
