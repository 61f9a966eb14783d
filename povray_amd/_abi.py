"""ctypes mirror of include/pvx.h (struct layouts and enums).

Used by the Python host side (scene upload, tests, bench) to call the C-ABI
library.  Layouts must match pvx.h byte for byte; the size checks at the end
of this module fail loudly if they drift.
"""
import ctypes as C

PVX_OK, PVX_EINVAL, PVX_ENOMEM, PVX_EDEVICE, PVX_EUNSUPPORTED, PVX_ESTATE = range(6)

PVX_SPHERE, PVX_BOX, PVX_TORUS, PVX_PLANE, PVX_MESH, PVX_CSG = 1, 2, 3, 4, 5, 6
PVX_TRIANGLE, PVX_SMOOTH_TRIANGLE = 7, 8
PVX_CONE = 9
PVX_DISC = 10
PVX_QUADRIC = 11
PVX_POLYGON = 12
PVX_LATHE, PVX_SOR = 13, 14
PVX_CONE_OPEN, PVX_CONE_CYLINDER = 1, 2
PVX_LATHE_LINEAR, PVX_LATHE_QUADRATIC, PVX_LATHE_CUBIC, PVX_LATHE_BEZIER, PVX_SOR_CUBIC = 0, 1, 2, 3, 4
PVX_REV_STURM, PVX_REV_OPEN = 1, 2

PVX_FLAG_NO_SHADOW = 1 << 0
PVX_FLAG_NO_IMAGE = 1 << 1
PVX_FLAG_NO_REFLECTION = 1 << 2
PVX_FLAG_NO_RADIOSITY = 1 << 3
PVX_FLAG_INVERTED = 1 << 4

PVX_RAY_PRIMARY = 0
PVX_RAY_IMAGE = 1
PVX_RAY_REFLECTION = 2
PVX_RAY_RADIOSITY = 4
PVX_RAY_SHADOW = 8

PVX_CSG_UNION, PVX_CSG_INTERSECTION, PVX_CSG_MERGE = 1, 2, 3

c_dp = C.POINTER(C.c_double)
c_fp = C.POINTER(C.c_float)
c_i32p = C.POINTER(C.c_int32)
c_u8p = C.POINTER(C.c_uint8)


class PvxObject(C.Structure):
    _fields_ = [("type", C.c_uint32), ("index", C.c_uint32),
                ("flags", C.c_uint32), ("pad", C.c_uint32)]


class PvxCsgNode(C.Structure):
    _fields_ = [("op", C.c_uint32), ("first_child", C.c_uint32),
                ("n_children", C.c_uint32), ("pad", C.c_uint32)]


class PvxMeshDesc(C.Structure):
    _fields_ = [("n_vertices", C.c_uint32), ("n_triangles", C.c_uint32),
                ("n_normals", C.c_uint32), ("pad", C.c_uint32),
                ("vertices", c_fp), ("indices", c_i32p), ("normals", c_fp)]


class PvxSceneDesc(C.Structure):
    _fields_ = [
        ("n_objects", C.c_uint32), ("objects", C.POINTER(PvxObject)),
        ("n_spheres", C.c_uint32), ("sphere_center", c_dp), ("sphere_radius", c_dp),
        ("n_planes", C.c_uint32), ("plane_normal", c_dp), ("plane_distance", c_dp),
        ("n_boxes", C.c_uint32), ("box_corner1", c_dp), ("box_corner2", c_dp),
        ("box_transform", c_i32p),
        ("n_tori", C.c_uint32), ("torus_major", c_dp), ("torus_minor", c_dp),
        ("torus_transform", c_i32p), ("torus_sturm", c_u8p),
        ("n_transforms", C.c_uint32), ("transforms", c_dp),
        ("n_meshes", C.c_uint32), ("meshes", C.POINTER(PvxMeshDesc)),
        ("n_csg", C.c_uint32), ("csg", C.POINTER(PvxCsgNode)),
        ("n_csg_children", C.c_uint32), ("csg_children", C.POINTER(PvxObject)),
        ("n_triangles", C.c_uint32), ("triangle_vertices", c_dp), ("triangle_normals", c_dp),
        ("n_cones", C.c_uint32), ("cone_apex", c_dp), ("cone_base", c_dp), ("cone_radius", c_dp),
        ("cone_flags", c_u8p), ("cone_transform", c_i32p),
        ("n_discs", C.c_uint32), ("disc_center", c_dp), ("disc_normal", c_dp), ("disc_radius", c_dp),
        ("disc_transform", c_i32p),
        ("n_quadrics", C.c_uint32), ("quadric_coef", c_dp), ("quadric_transform", c_i32p),
        ("n_polygons", C.c_uint32), ("polygon_first", C.POINTER(C.c_uint32)), ("polygon_points", c_dp),
        ("polygon_transform", c_i32p),
        ("mesh_transform", c_i32p),
        ("n_revolutions", C.c_uint32), ("revolution_first", C.POINTER(C.c_uint32)), ("revolution_points", c_dp),
        ("revolution_kind", c_u8p), ("revolution_flags", c_u8p), ("revolution_transform", c_i32p),
    ]


class PvxRay(C.Structure):
    _fields_ = [("o", C.c_double * 3), ("d", C.c_double * 3)]


class PvxHit(C.Structure):
    _fields_ = [("t", C.c_double), ("obj", C.c_int32), ("prim", C.c_int32)]


class PvxBuildOpts(C.Structure):
    _fields_ = [("max_leaf_size", C.c_uint32), ("sah_bins", C.c_uint32),
                ("accel", C.c_uint32), ("bsp_max_depth", C.c_uint32),
                ("builder", C.c_uint32), ("reserved", C.c_uint32 * 3)]


PVX_ACCEL_BVH, PVX_ACCEL_BSP = 0, 1
PVX_BUILDER_HOST_SAH, PVX_BUILDER_GPU = 0, 1

# traversal variants (pvx_scene_autotune)
PVX_TRAVERSAL_DEFAULT, PVX_TRAVERSAL_WIDE_WW, PVX_TRAVERSAL_WIDE_STEP, PVX_TRAVERSAL_BINARY = 0, 1, 2, 3
PVX_TRAVERSAL_PACKET, PVX_TRAVERSAL_FUSED, PVX_TRAVERSAL_REFILL, PVX_TRAVERSAL_LIGHT_BINARY = 4, 5, 6, 7
PVX_TRAVERSAL_BINARY_WW, PVX_TRAVERSAL_REFILL_BINARY = 8, 9


class PvxTuneResult(C.Structure):
    _fields_ = [("variant", C.c_uint32), ("grid_blocks", C.c_uint32),
                ("best_ms", C.c_double), ("default_ms", C.c_double),
                ("tile_w", C.c_uint32), ("tile_h", C.c_uint32)]


class PvxSceneStats(C.Structure):
    _fields_ = [("n_nodes", C.c_uint64), ("n_prims", C.c_uint64),
                ("node_bytes", C.c_uint64), ("tri_bytes", C.c_uint64),
                ("device_bytes", C.c_uint64), ("bvh_depth", C.c_uint32),
                ("n_infinite", C.c_uint32), ("build_ms", C.c_double),
                ("bvh_build_ms", C.c_double)]


class PvxCamera(C.Structure):
    _fields_ = [("location", C.c_double * 3), ("direction", C.c_double * 3),
                ("right", C.c_double * 3), ("up", C.c_double * 3),
                ("width", C.c_uint32), ("height", C.c_uint32),
                ("jitter_x", C.c_double), ("jitter_y", C.c_double)]


PVX_LIGHT_POINT, PVX_LIGHT_AREA = 1, 2


class PvxMaterial(C.Structure):
    _fields_ = [("pigment", C.c_float * 3), ("filter", C.c_float), ("transmit", C.c_float),
                ("ambient", C.c_float * 3), ("emission", C.c_float * 3),
                ("diffuse", C.c_float), ("brilliance", C.c_float),
                ("phong", C.c_float), ("phong_size", C.c_float),
                ("reflection_min", C.c_float * 3), ("reflection_max", C.c_float * 3),
                ("reflection_falloff", C.c_float), ("ior", C.c_float)]


class PvxLight(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("size1", C.c_uint32), ("size2", C.c_uint32),
                ("pad", C.c_uint32), ("position", C.c_double * 3), ("axis1", C.c_double * 3),
                ("axis2", C.c_double * 3), ("colour", C.c_float * 3), ("pad2", C.c_float)]


class PvxShadingDesc(C.Structure):
    _fields_ = [("n_materials", C.c_uint32), ("materials", C.POINTER(PvxMaterial)),
                ("n_lights", C.c_uint32), ("lights", C.POINTER(PvxLight)),
                ("ambient_light", C.c_float * 3), ("background", C.c_float * 3),
                ("max_trace_level", C.c_uint32), ("adc_bailout", C.c_float)]


class PvxRenderStats(C.Structure):
    _fields_ = [("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("bounces", C.c_uint32), ("pad", C.c_uint32), ("trace_ms", C.c_double),
                ("shadow_ms", C.c_double), ("total_ms", C.c_double)]


PVX_RADIOSITY_PLAIN, PVX_RADIOSITY_LEARNED = 0, 1


class PvxRadiositySettings(C.Structure):
    _fields_ = [("count", C.c_uint32), ("recursion_limit", C.c_uint32),
                ("nearest_count", C.c_uint32), ("mode", C.c_uint32),
                ("error_bound", C.c_float), ("minimum_reuse", C.c_float),
                ("maximum_reuse", C.c_float), ("pretrace_start", C.c_float),
                ("pretrace_end", C.c_float), ("brightness", C.c_float),
                ("adc_bailout", C.c_float), ("probe_threshold", C.c_float),
                ("train_samples", C.c_uint32), ("train_steps", C.c_uint32),
                ("probe_tile", C.c_uint32), ("probes_per_tile", C.c_uint32),
                ("train_batch", C.c_uint32)]


class PvxRadiosityStats(C.Structure):
    _fields_ = [("frame_ms", C.c_double), ("pretrace_ms", C.c_double), ("final_ms", C.c_double),
                ("train_ms", C.c_double), ("gathers", C.c_uint64 * 2),
                ("gather_rays", C.c_uint64), ("records", C.c_uint64 * 2),
                ("final_misses", C.c_uint64), ("surrogate_queries", C.c_uint64),
                ("fallbacks", C.c_uint64), ("probes", C.c_uint64), ("probe_error", C.c_float),
                ("train_loss", C.c_float)]


class PvxBatchRadiosityStats(C.Structure):
    _fields_ = [("batch_ms", C.c_double), ("gpu_frame_ms", C.c_double), ("pixels", C.c_uint64),
                ("gather_rays", C.c_uint64), ("surrogate_queries", C.c_uint64), ("fallbacks", C.c_uint64),
                ("final_misses", C.c_uint64), ("streams", C.c_uint32), ("pad", C.c_uint32)]


def radiosity_settings(scene, mode=PVX_RADIOSITY_PLAIN, **over):
    """pvx_radiosity_settings from scene.settings['radiosity'] (defaults of
    SceneRadiositySettings otherwise) with keyword overrides."""
    cfg = dict(count=35, recursion_limit=2, nearest_count=5, error_bound=1.8,
               minimum_reuse=0.015, maximum_reuse=0.2, pretrace_start=0.08, pretrace_end=0.04,
               brightness=1.0, adc_bailout=0.01, probe_threshold=0.02, train_samples=8192,
               train_steps=1500, probe_tile=32, probes_per_tile=2, train_batch=0)
    cfg.update(scene.settings.get("radiosity", {}))
    cfg.update(over)
    st = PvxRadiositySettings()
    for k, _ in PvxRadiositySettings._fields_:
        if k == "mode":
            st.mode = mode
        else:
            setattr(st, k, cfg[k])
    return st


assert C.sizeof(PvxObject) == 16
assert C.sizeof(PvxRay) == 48
assert C.sizeof(PvxHit) == 16
assert C.sizeof(PvxMeshDesc) == 40
assert C.sizeof(PvxCamera) == 4 * 24 + 8 + 16
