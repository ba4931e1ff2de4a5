"""ctypes mirror of include/prh.h and the loader of libprh.so.

The structures here are the C ABI's plain-data types; the functions are bound
with explicit prototypes.  The library is built in-tree by
``__graft_entry__.build()`` into ``polyred_amd/_lib/libprh.so``; loading fails
loudly when it is missing (there is no CPU fallback in this package).
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libprh.so")

PRH_OK = 0
PRH_E_INVALID = -1
PRH_E_UNSUPPORTED = -2
PRH_E_HIP = -3
PRH_E_NOMEM = -4
PRH_E_NODEVICE = -5

PRH_NOALLOC_SCRATCH_BYTES = 32 << 20  # prh_diff_backward_noalloc's per-GPU scratch array

PRH_LIGHT_POINT = 0
PRH_LIGHT_DIRECTIONAL = 1
PRH_CAMERA_PERSPECTIVE = 0
PRH_CAMERA_ORTHOGRAPHIC = 1

PRH_BATCH_GRAPH = 1
PRH_BATCH_STREAMS = 2

PRH_BUF_VIS = 0
PRH_BUF_DEPTH = 1
PRH_BUF_LINEAR = 2
PRH_BUF_SHADOW = 3
PRH_BUF_PREGAMMA = 4
PRH_BUF_SHADOW_COUNT = 5


class prh_texture(C.Structure):
    _fields_ = [
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("rgba", C.POINTER(C.c_uint8)),
        ("use_mipmap", C.c_int32),
    ]


class prh_material(C.Structure):
    _fields_ = [
        ("diffuse", C.c_uint8 * 4),
        ("specular", C.c_uint8 * 4),
        ("texture", C.c_uint8 * 4),
        ("shininess", C.c_float),
        ("flat_shading", C.c_int32),
        ("ambient_occlusion", C.c_int32),
        ("receive_shadow", C.c_int32),
        ("texture_id", C.c_int32),
    ]


class prh_light(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("position", C.c_float * 3),
        ("direction", C.c_float * 3),
        ("intensity", C.c_float),
        ("color", C.c_uint8 * 4),
        ("cast_shadow", C.c_int32),
    ]


_i32p = C.POINTER(C.c_int32)
_f32p = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)


class prh_scene_desc(C.Structure):
    _fields_ = [
        ("n_meshes", C.c_int32),
        ("mesh_vtx_off", _i32p),
        ("mesh_tri_off", _i32p),
        ("vpos", _f32p),
        ("vnor", _f32p),
        ("vcol", _u8p),
        ("vuv", _f32p),
        ("tri_v", _i32p),
        ("tri_mat", _i32p),
        ("n_objects", C.c_int32),
        ("obj_mesh", _i32p),
        ("obj_model", _f32p),
        ("obj_mat_base", _i32p),
        ("n_materials", C.c_int32),
        ("materials", C.POINTER(prh_material)),
        ("n_lights", C.c_int32),
        ("lights", C.POINTER(prh_light)),
        ("n_ambient", C.c_int32),
        ("ambient_intensity", _f32p),
        ("center", C.c_float * 3),
        ("n_textures", C.c_int32),
        ("textures", C.POINTER(prh_texture)),
    ]


class prh_camera(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("position", C.c_float * 3),
        ("target", C.c_float * 3),
        ("up", C.c_float * 3),
        ("fov", C.c_float),
        ("aspect", C.c_float),
        ("near_", C.c_float),
        ("far_", C.c_float),
        ("left", C.c_float),
        ("right", C.c_float),
        ("bottom", C.c_float),
        ("top", C.c_float),
    ]


class prh_options(C.Structure):
    _fields_ = [
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("msaa", C.c_int32),
        ("background", C.c_uint8 * 4),
        ("shadow_map", C.c_int32),
        ("gamma_correct", C.c_int32),
        ("shadow_size", C.c_int32),
        ("keep_linear", C.c_int32),
    ]


class prh_frame_stats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_float),
        ("ms_shadow", C.c_float),
        ("ms_vertex", C.c_float),
        ("ms_setup", C.c_float),
        ("ms_raster", C.c_float),
        ("ms_shade", C.c_float),
        ("ms_ao", C.c_float),
        ("n_prims", C.c_int32),
        ("n_clipped", C.c_int32),
        ("n_bin_pairs", C.c_int32),
        ("nan_rerun", C.c_int32),
        ("ms_diff", C.c_float),
    ]


class prh_diff_params(C.Structure):
    _fields_ = [(n, _f32p) for n in ("albedo", "normal", "kd", "ks", "shininess", "light_pos", "light_int")]


class prh_diff_grads(C.Structure):
    _fields_ = [(n, _f32p) for n in ("albedo", "normal", "kd", "ks", "shininess", "light_pos", "light_int")]


PRH_FIT_ALBEDO, PRH_FIT_KD, PRH_FIT_KS, PRH_FIT_SHININESS, PRH_FIT_LIGHT_POS, PRH_FIT_LIGHT_INT = 1, 4, 8, 16, 32, 64
PRH_FIT_STORED_GBUFFER = 256


class prh_tuning(C.Structure):
    _fields_ = [("raster_wg_per_cu", C.c_int32), ("item_prims", C.c_int32), ("shadow_direct_max", C.c_int32),
                ("main_direct_max", C.c_int32), ("overlap_shadows", C.c_int32), ("raster_tile", C.c_int32),
                ("ms", C.c_float)]


class prh_fit_options(C.Structure):
    _fields_ = [("steps", C.c_int32), ("groups", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float)]


class prh_fit_stats(C.Structure):
    _fields_ = [("visibility_bytes", C.c_double), ("ms_steps", C.c_float), ("ms_gbuffer", C.c_float),
                ("steps", C.c_int32), ("n_views", C.c_int32)]


# Every symbol include/prh.h declares, with its prototype.
PROTOTYPES = {
    "prh_last_error": (C.c_char_p, []),
    "prh_open": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    "prh_close": (C.c_int, [C.c_void_p]),
    "prh_scene_upload": (C.c_int, [C.c_void_p, C.POINTER(prh_scene_desc), C.POINTER(C.c_void_p)]),
    "prh_scene_free": (C.c_int, [C.c_void_p]),
    "prh_renderer_new": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(prh_camera), C.POINTER(prh_options),
                                   C.POINTER(C.c_void_p)]),
    "prh_renderer_free": (C.c_int, [C.c_void_p]),
    "prh_render": (C.c_int, [C.c_void_p]),
    "prh_read_image": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "prh_read_buffer": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t]),
    "prh_last_stats": (C.c_int, [C.c_void_p, C.POINTER(prh_frame_stats)]),
    "prh_version": (C.c_char_p, []),
    "prh_render_batch": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.c_int32, C.POINTER(C.c_float)]),
    "prh_diff_sizes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "prh_diff_forward": (C.c_int, [C.c_void_p, C.POINTER(prh_diff_params), C.c_void_p, C.c_size_t]),
    "prh_diff_backward": (C.c_int, [C.c_void_p, C.POINTER(prh_diff_params), C.c_void_p, C.c_size_t,
                                    C.POINTER(C.c_double), C.POINTER(prh_diff_grads)]),
    "prh_diff_backward_noalloc": (C.c_int, [C.c_void_p, C.POINTER(prh_diff_params), C.c_void_p, C.c_size_t,
                                            C.POINTER(C.c_double), C.POINTER(prh_diff_grads)]),
    "prh_fit_new": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "prh_fit_free": (C.c_int, [C.c_void_p]),
    "prh_fit_targets": (C.c_int, [C.c_void_p, C.POINTER(prh_diff_params)]),
    "prh_fit_set_target": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t]),
    "prh_fit_run": (C.c_int, [C.c_void_p, C.POINTER(prh_diff_params), C.POINTER(prh_fit_options),
                              C.POINTER(prh_diff_grads), C.c_void_p]),
    "prh_fit_last_stats": (C.c_int, [C.c_void_p, C.POINTER(prh_fit_stats)]),
    "prh_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(prh_tuning)]),
    "prh_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(prh_tuning)]),
    "prh_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(prh_tuning)]),
    "prh_autotune_batch": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(prh_tuning)]),
    "prh_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "prh_host_free": (C.c_int, [C.c_void_p]),
    "prh_surface_new": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "prh_surface_free": (C.c_int, [C.c_void_p]),
    "prh_surface_acquire": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "prh_surface_render": (C.c_int, [C.c_void_p, C.c_void_p]),
    "prh_surface_write": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "prh_surface_present": (C.c_int, [C.c_void_p]),
    "prh_surface_read": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "prh_surface_resize": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "prh_surface_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
}

_lib = None


class PrhError(RuntimeError):
    """A negative status from libprh, with prh_last_error()'s text."""

    def __init__(self, code: int, where: str, text: str):
        super().__init__(f"{where} failed with status {code}: {text}")
        self.code = code


def load() -> C.CDLL:
    """Load libprh.so (built in-tree); raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libprh is not built: {LIB_PATH} is missing "
                           "(run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int, where: str) -> None:
    if code != PRH_OK:
        text = load().prh_last_error()
        raise PrhError(code, where, text.decode() if text else "")
