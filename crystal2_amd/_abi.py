"""ctypes mirror of include/crystal2.h (the C-ABI boundary)."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# CRYSTAL2_LIB selects another build of the library (for example a variant under test).
LIB_PATH = os.environ.get("CRYSTAL2_LIB") or os.path.join(HERE, "_lib", "libcrystal2.so")

C2_OK = 0
ERRORS = {-1: "C2_ERR_INVALID", -2: "C2_ERR_PARSE", -3: "C2_ERR_IO", -4: "C2_ERR_DEVICE",
          -5: "C2_ERR_NOMEM", -6: "C2_ERR_UNSUPPORTED"}

C2_SHAPE_TRIANGLE, C2_SHAPE_SPHERE, C2_SHAPE_DISK = 0, 1, 2
C2_MAT_DIFFUSE, C2_MAT_MIRROR, C2_MAT_GLASS, C2_MAT_MICRO, C2_MAT_PHONG, C2_MAT_PLASTIC, C2_MAT_ORENNAYAR = range(7)
C2_NDF_GGX, C2_NDF_BECKMANN = 0, 1
C2_LIGHT_AREA, C2_LIGHT_POINT = 0, 1
C2_ARGORDER_RTL, C2_ARGORDER_LTR = 0, 1
C2_FILM_F32, C2_FILM_F32_KAHAN, C2_FILM_F64 = 0, 1, 2
C2_INTEGRATOR_PATH, C2_INTEGRATOR_DIRECT = 0, 1
C2_ACCEL_BVH, C2_ACCEL_KDTREE, C2_ACCEL_BVH8 = 0, 1, 2


class c2_material(C.Structure):
    _fields_ = [("type", C.c_int32), ("ndf", C.c_int32), ("phong_k", C.c_int32), ("kd_texture", C.c_int32),
                ("color", C.c_float * 3), ("roughness", C.c_float * 2), ("ior", C.c_float),
                ("ks", C.c_float * 3), ("sigma", C.c_float)]


class c2_light(C.Structure):
    _fields_ = [("type", C.c_int32), ("entity", C.c_int32), ("radiance", C.c_float * 3),
                ("position", C.c_float * 3)]


class c2_camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("look_at", C.c_float * 3), ("up", C.c_float * 3),
                ("fov", C.c_float)]


class c2_texture(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgb", C.POINTER(C.c_uint8))]


class c2_scene_desc(C.Structure):
    _fields_ = [("num_entities", C.c_int32),
                ("entity_shape_kind", C.POINTER(C.c_int32)), ("entity_shape_index", C.POINTER(C.c_int32)),
                ("entity_material", C.POINTER(C.c_int32)), ("entity_light", C.POINTER(C.c_int32)),
                ("num_triangles", C.c_int32), ("triangles", C.POINTER(C.c_float)),
                ("num_spheres", C.c_int32), ("spheres", C.POINTER(C.c_float)),
                ("num_materials", C.c_int32), ("materials", C.POINTER(c2_material)),
                ("num_lights", C.c_int32), ("lights", C.POINTER(c2_light)),
                ("camera", c2_camera), ("num_textures", C.c_int32), ("textures", C.POINTER(c2_texture)),
                ("skybox", C.c_int32 * 6)]


class c2_config(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("threads", C.c_int32),
                ("max_depth", C.c_int32), ("accel", C.c_char * 16), ("integrator", C.c_char * 32),
                ("sampler", C.c_char * 16)]


class c2_accel_stats(C.Structure):
    _fields_ = [("num_nodes_bvh2", C.c_int64), ("num_nodes_bvh4", C.c_int64), ("max_depth_bvh4", C.c_int32),
                ("leaf_size", C.c_int32), ("build_seconds", C.c_double), ("device_bytes", C.c_int64)]


class c2_render_opts(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("max_depth", C.c_int32),
                ("arg_order", C.c_int32), ("film_mode", C.c_int32), ("paths_in_flight", C.c_int32),
                ("queue_blocks", C.c_int32), ("win_x0", C.c_int32), ("win_y0", C.c_int32), ("win_w", C.c_int32),
                ("win_h", C.c_int32), ("integrator", C.c_int32), ("accel", C.c_int32)]


class c2_nrc_opts(C.Structure):
    _fields_ = [("learning_rate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float),
                ("seed", C.c_int32), ("batch_capacity", C.c_int32), ("bbox_min", C.c_float * 3),
                ("bbox_max", C.c_float * 3)]


class c2_nrc_stats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("flop_forward_per_query", C.c_double), ("queries", C.c_int64),
                ("train_rows", C.c_int64)]


class c2_cache_times(C.Structure):
    _fields_ = [("infer_ms", C.c_double), ("train_ms", C.c_double), ("queries", C.c_int64),
                ("train_rows", C.c_int64), ("collect_ms", C.c_double)]


NRC_PARAMS = 5 * 64 * 64 + 5 * 64


class c2_pt_stats(C.Structure):
    _fields_ = [("samples", C.c_int64), ("extension_rays", C.c_int64), ("shadow_rays", C.c_int64),
                ("mis_rays", C.c_int64), ("kernel_launches", C.c_int64)]


class c2_batch_item(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32)]


C2_SCHED_WAVEFRONT, C2_SCHED_PERSISTENT = 0, 1


class c2_batch_tuning(C.Structure):
    _fields_ = [("paths_in_flight", C.c_int32), ("queue_blocks", C.c_int32), ("pass_ms", C.c_double),
                ("candidates", C.c_int32), ("schedule", C.c_int32), ("block_threads", C.c_int32),
                ("waves_per_simd", C.c_int32), ("sample0_ms_persistent", C.c_double), ("render_ms_wavefront", C.c_double),
                ("render_ms_persistent", C.c_double), ("bvh_width", C.c_int32), ("pass_ms_width8", C.c_double)]


class c2_batch_stats(C.Structure):
    _fields_ = [("samples", C.c_int64), ("samples_per_render", C.c_int64), ("extension_rays", C.c_int64),
                ("shadow_rays", C.c_int64), ("mis_rays", C.c_int64), ("passes", C.c_int32),
                ("paths_in_flight", C.c_int32), ("queue_blocks", C.c_int32), ("graph_nodes", C.c_int64),
                ("schedule", C.c_int32), ("block_threads", C.c_int32), ("waves_per_simd", C.c_int32)]


class c2_kernel_times(C.Structure):
    _fields_ = [("ms", C.c_double * 8), ("launches", C.c_int64 * 8)]


KERNEL_CLASSES = ["generate", "extend", "shade", "occlude", "mis", "resolve", "accumulate", "nee"]

P = C.c_void_p
_i32p = C.POINTER(C.c_int32)
_f32p = C.POINTER(C.c_float)

# (name, restype, argtypes) for every symbol include/crystal2.h declares.
SIGNATURES = [
    ("c2_last_error", C.c_char_p, []),
    ("c2_config_load", C.c_int, [C.c_char_p, C.POINTER(c2_config)]),
    ("c2_scene_load", C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(P)]),
    ("c2_scene_create", C.c_int, [C.POINTER(P)]),
    ("c2_scene_destroy", None, [P]),
    ("c2_scene_add_material", C.c_int, [P, C.POINTER(c2_material), _i32p]),
    ("c2_scene_add_triangles", C.c_int, [P, _f32p, C.c_int32, _f32p, C.c_int32, _f32p]),
    ("c2_scene_add_sphere", C.c_int, [P, _f32p, C.c_float, _f32p, C.c_int32, _f32p]),
    ("c2_scene_add_disk", C.c_int, [P, _f32p, C.c_float, _f32p, C.c_int32, _f32p]),
    ("c2_scene_add_point_light", C.c_int, [P, _f32p, _f32p]),
    ("c2_scene_add_texture", C.c_int, [P, C.c_int32, C.c_int32, C.POINTER(C.c_uint8), _i32p]),
    ("c2_scene_set_skybox", C.c_int, [P, _i32p]),
    ("c2_image_load", C.c_int, [C.c_char_p, _i32p, _i32p, C.POINTER(C.POINTER(C.c_uint8))]),
    ("c2_image_free", None, [C.POINTER(C.c_uint8)]),
    ("c2_scene_set_camera", C.c_int, [P, C.POINTER(c2_camera)]),
    ("c2_scene_get_desc", C.c_int, [P, C.POINTER(c2_scene_desc)]),
    ("c2_scene_save", C.c_int, [P, C.c_char_p]),
    ("c2_scene_load_binary", C.c_int, [C.c_char_p, C.POINTER(P)]),
    ("c2_film_write_hdr", C.c_int, [C.c_char_p, C.c_int32, C.c_int32, _f32p]),
    ("c2_film_write_pfm", C.c_int, [C.c_char_p, C.c_int32, C.c_int32, _f32p]),
    ("c2_accel_build", C.c_int, [C.POINTER(c2_scene_desc), C.POINTER(P)]),
    ("c2_accel_create", C.c_int, [C.POINTER(c2_scene_desc), C.c_int32, C.POINTER(P)]),
    ("c2_accel_destroy", None, [P]),
    ("c2_accel_get_stats", C.c_int, [P, C.POINTER(c2_accel_stats)]),
    ("c2_accel_intersect", C.c_int, [P, _f32p, C.c_int64, _i32p, _f32p]),
    ("c2_accel_intersect_test", C.c_int, [P, _f32p, _i32p, C.c_int64, C.POINTER(C.c_uint8)]),
    ("c2_accel_intersect_device", C.c_int, [P, P, C.c_int64, P, P, P]),
    ("c2_accel_intersect_noalloc", C.c_int, [P, _f32p, C.c_int64, _i32p, _f32p]),
    ("c2_accel_intersect_test_device", C.c_int, [P, P, P, C.c_int64, P, P]),
    ("c2_accel_count_visits", C.c_int, [P, _f32p, C.c_int64, C.c_int32, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]),
    ("c2_pt_create", C.c_int, [C.POINTER(c2_scene_desc), C.POINTER(c2_render_opts), C.POINTER(P)]),
    ("c2_pt_destroy", None, [P]),
    ("c2_pt_reset", C.c_int, [P]),
    ("c2_pt_render", C.c_int, [P, C.c_int32, C.c_int32, P]),
    ("c2_pt_get_film", C.c_int, [P, _f32p]),
    ("c2_pt_get_stats", C.c_int, [P, C.POINTER(c2_pt_stats)]),
    ("c2_pt_sync", C.c_int, [P]),
    ("c2_pt_set_profiling", C.c_int, [P, C.c_int32]),
    ("c2_pt_profile_kernels", C.c_int, [P, C.c_uint32]),
    ("c2_pt_get_kernel_times", C.c_int, [P, C.POINTER(c2_kernel_times)]),
    ("c2_pt_capture_rays", C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, _f32p, _i32p, C.c_int64,
                                     C.POINTER(C.c_int64)]),
    ("c2_nrc_create", C.c_int, [C.POINTER(c2_nrc_opts), C.POINTER(P)]),
    ("c2_nrc_destroy", None, [P]),
    ("c2_nrc_get_params", C.c_int, [P, _f32p]),
    ("c2_nrc_set_params", C.c_int, [P, _f32p]),
    ("c2_nrc_infer", C.c_int, [P, _f32p, C.c_int64, _f32p]),
    ("c2_nrc_train_step", C.c_int, [P, _f32p, _f32p, C.c_int64, C.c_int32, _f32p, _f32p]),
    ("c2_nrc_get_stats", C.c_int, [P, C.POINTER(c2_nrc_stats)]),
    ("c2_pt_set_cache", C.c_int, [P, P, C.c_int32, C.c_int32]),
    ("c2_pt_get_cache_times", C.c_int, [P, C.POINTER(c2_cache_times)]),
    ("c2_halton_index", C.c_int, [C.c_int32] * 5 + [_i32p]),
    ("c2_halton_value", C.c_int, [C.c_int32] * 4 + [_f32p]),
    ("c2_util_stream_copy", C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    ("c2_batch_create", C.c_int, [C.POINTER(c2_batch_item), C.c_int32, C.POINTER(c2_render_opts), C.POINTER(P)]),
    ("c2_batch_destroy", None, [P]),
    ("c2_batch_autotune", C.c_int, [P, C.POINTER(c2_batch_tuning)]),
    ("c2_batch_configure", C.c_int, [P, C.POINTER(c2_batch_tuning)]),
    ("c2_batch_render", C.c_int, [P, C.c_int32]),
    ("c2_batch_sync", C.c_int, [P]),
    ("c2_batch_get_film", C.c_int, [P, C.c_int32, _f32p]),
    ("c2_batch_get_stats", C.c_int, [P, C.POINTER(c2_batch_stats)]),
]

_lib = None


class Crystal2Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def load(path: str = LIB_PATH):
    """Load the product library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"crystal2 native library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int):
    if rc != C2_OK:
        msg = _lib.c2_last_error().decode(errors="replace") if _lib is not None else ""
        raise Crystal2Error(rc, msg)
