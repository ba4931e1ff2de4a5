"""TEST INFRASTRUCTURE ONLY: ctypes loader of the C restatement of polyred's
CPU renderer (oracle/oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the parity checker or the timed CPU baseline.  The product
(polyred_amd, libprh) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from polyred_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class oracle_out(C.Structure):
    _fields_ = [
        ("vis_order", C.POINTER(C.c_uint32)),
        ("depth", C.POINTER(C.c_float)),
        ("linear", C.POINTER(C.c_float)),
        ("pregamma", C.POINTER(C.c_uint8)),
        ("image", C.POINTER(C.c_uint8)),
        ("shadow_maps", C.POINTER(C.c_float)),
        ("shade_mask", C.POINTER(C.c_uint8)),
        ("shade_err64", C.POINTER(C.c_double)),
        ("shadow_count", C.POINTER(C.c_uint8)),
        ("ao_total", C.POINTER(C.c_float)),
        ("light_matrices", C.POINTER(C.c_float)),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        f32p, u8p = C.POINTER(C.c_float), C.POINTER(C.c_uint8)
        L.oracle_set_workers.argtypes = [C.c_int]
        L.oracle_set_workers.restype = None
        L.oracle_render.argtypes = [C.POINTER(abi.prh_scene_desc), C.POINTER(abi.prh_camera),
                                    C.POINTER(abi.prh_options), C.POINTER(oracle_out)]
        L.oracle_barycoord.argtypes = [f32p] * 5
        L.oracle_viewport_matrix.argtypes = [C.c_float, C.c_float, f32p]
        L.oracle_view_matrix.argtypes = [f32p] * 4
        L.oracle_proj_matrix.argtypes = [C.POINTER(abi.prh_camera), f32p]
        L.oracle_mat4_inv.argtypes = [f32p, f32p]
        L.oracle_mat4_mulm.argtypes = [f32p, f32p, f32p]
        L.oracle_aabb_intersect.argtypes = [f32p] * 4
        L.oracle_aabb_contains.argtypes = [f32p, f32p, C.c_int, f32p]
        L.oracle_aabb_new.argtypes = [C.c_int, f32p, f32p, f32p]
        L.oracle_tri_is_valid.argtypes = [f32p] * 3
        L.oracle_interp_worldpos.argtypes = [f32p] * 5
        L.oracle_mat_at.argtypes = [C.c_int32, C.c_int64]
        L.oracle_srgb_from_linear.argtypes = [C.c_float]
        L.oracle_srgb_from_linear.restype = C.c_float
        L.oracle_linear_from_srgb.argtypes = [C.c_float]
        L.oracle_linear_from_srgb.restype = C.c_float
        L.oracle_gamma_u8.argtypes = [C.c_uint8]
        L.oracle_gamma_u8.restype = C.c_uint8
        L.oracle_shadow_quant.argtypes = [C.c_uint8, C.c_int]
        L.oracle_shadow_quant.restype = C.c_uint8
        L.oracle_fragment_shader.argtypes = [C.POINTER(abi.prh_material), f32p, f32p, f32p, f32p, C.c_int,
                                             C.POINTER(abi.prh_light), C.c_int, f32p, u8p, f32p]
        L.oracle_ao_dirs.argtypes = [f32p]
        L.oracle_ao_march.argtypes = [f32p, u8p, C.c_int, C.c_int, C.c_int, C.c_int, f32p, f32p, f32p]
        L.oracle_shadow_lookup.argtypes = [f32p, f32p, f32p, C.c_int, C.c_int, f32p, C.c_float, C.c_float,
                                           C.c_float]
        L.oracle_resize.argtypes = [u8p, C.c_int, C.c_int, u8p, C.c_int, C.c_int]
        L.oracle_texture_query.argtypes = [C.POINTER(abi.prh_texture), C.c_float, C.c_float, C.c_float, u8p]
        L.oracle_texture_level.argtypes = [C.POINTER(abi.prh_texture), C.c_int, C.POINTER(C.c_int),
                                           C.POINTER(C.c_int), u8p, C.c_size_t]
        f64p = C.POINTER(C.c_double)
        L.oracle_diff.argtypes = [C.POINTER(abi.prh_scene_desc), C.POINTER(abi.prh_camera),
                                  C.POINTER(abi.prh_options), C.POINTER(C.c_uint32), C.POINTER(oracle_diff_params),
                                  f64p, f64p, C.POINTER(oracle_diff_grads), f64p]
        _lib = L
    return _lib


def fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def f32(*v) -> np.ndarray:
    return np.ascontiguousarray(np.array(v, dtype=np.float32).reshape(-1))


class Frame:
    """Outputs of one oracle frame (raster-row layout unless noted)."""

    def __init__(self, scene, mask=None, shadows=True):
        m = getattr(scene, "msaa", 1)
        W, H = scene.width * m, scene.height * m  # the MSAA-sized framebuffer
        n = W * H
        self.vis_order = np.zeros(n, np.uint32)
        self.depth = np.zeros(n, np.float32)
        self.linear = np.zeros(3 * n, np.float32)
        self.pregamma = np.zeros(4 * n, np.uint8)
        self.image = np.zeros(4 * scene.width * scene.height, np.uint8)  # after Resize
        S = scene.shadow_size if scene.shadow_size > 0 else None
        nl = max(1, len(scene.lights))
        self.shadow_maps = np.zeros(nl * (S * S if S else n), np.float32) if scene.shadow_map else None
        self.err64 = C.c_double(0.0)
        self.mask = None if mask is None else np.ascontiguousarray(mask.astype(np.uint8).reshape(-1))
        self.shadow_count = np.zeros(n, np.uint8)
        self.ao_total = np.zeros(n, np.float32)
        self.light_matrices = np.zeros(32 * nl, np.float32)


def set_workers(n: int) -> None:
    """Threads of later oracle frames: 0 the parity order (the default), 1
    everything on one thread (Mode A), n > 1 the raster passes and the
    deferred pass on n threads with per-pixel locks (Mode B: tie order
    arbitrary, for timing only)."""
    lib().oracle_set_workers(int(n))


def render(scene, mask=None) -> Frame:
    """Render one frame of `scene` with the oracle."""
    L = lib()
    fr = Frame(scene, mask)
    o = oracle_out()
    o.vis_order = fr.vis_order.ctypes.data_as(C.POINTER(C.c_uint32))
    o.depth = fptr(fr.depth)
    o.linear = fptr(fr.linear)
    o.pregamma = fr.pregamma.ctypes.data_as(C.POINTER(C.c_uint8))
    o.image = fr.image.ctypes.data_as(C.POINTER(C.c_uint8))
    if fr.shadow_maps is not None:
        o.shadow_maps = fptr(fr.shadow_maps)
    if fr.mask is not None:
        o.shade_mask = fr.mask.ctypes.data_as(C.POINTER(C.c_uint8))
    o.shade_err64 = C.pointer(fr.err64)
    o.shadow_count = fr.shadow_count.ctypes.data_as(C.POINTER(C.c_uint8))
    o.ao_total = fptr(fr.ao_total)
    o.light_matrices = fptr(fr.light_matrices)
    desc = scene.desc()
    cam = scene.camera_struct()
    opts = scene.options_struct()
    rc = L.oracle_render(C.byref(desc), C.byref(cam), C.byref(opts), C.byref(o))
    if rc != 0:
        raise RuntimeError(f"oracle_render failed: {rc}")
    m = getattr(scene, "msaa", 1)
    W, H = scene.width * m, scene.height * m
    fr.vis_order = fr.vis_order.reshape(H, W)
    fr.depth = fr.depth.reshape(H, W)
    fr.linear = fr.linear.reshape(H, W, 3)
    fr.pregamma = fr.pregamma.reshape(H, W, 4)
    fr.image = fr.image.reshape(scene.height, scene.width, 4)
    fr.shadow_count = fr.shadow_count.reshape(H, W)
    fr.ao_total = fr.ao_total.reshape(H, W)
    fr.light_matrices = fr.light_matrices.reshape(-1, 2, 4, 4)
    return fr


def ao_march(depth: np.ndarray, ok: np.ndarray, px: int, py: int):
    """maxElevationAngle of the eight directions and AO's total and factor
    at raster pixel (px, py) of a fragment buffer given as (H, W) raster rows
    of depths and Ok flags: returns (angles[8], total, factor), float32."""
    depth = np.ascontiguousarray(depth, np.float32)
    okb = np.ascontiguousarray(ok, np.uint8)
    H, W = depth.shape
    ang = np.zeros(8, np.float32)
    tot, fac = C.c_float(), C.c_float()
    rc = lib().oracle_ao_march(fptr(depth), okb.ctypes.data_as(C.POINTER(C.c_uint8)), W, H, px, py, fptr(ang),
                               C.byref(tot), C.byref(fac))
    if rc != 0:
        raise RuntimeError(f"oracle_ao_march failed: {rc}")
    return ang, np.float32(tot.value), np.float32(fac.value)


def shadow_lookup(s2w, lview, lproj, smap: np.ndarray, x: float, y: float, depth: float) -> bool:
    """shadingVisibility's lookup of fragment (x, y, depth) in one shadow map
    ((SH, SW) float32, texel index x + y*SW): True if occluded."""
    m = [np.ascontiguousarray(np.asarray(a, np.float32).reshape(16)) for a in (s2w, lview, lproj)]
    smap = np.ascontiguousarray(smap, np.float32)
    SH, SW = smap.shape
    return bool(lib().oracle_shadow_lookup(fptr(m[0]), fptr(m[1]), fptr(m[2]), SW, SH, fptr(smap), x, y, depth))


class oracle_diff_params(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("albedo", "normal", "kd", "ks", "shininess", "light_pos",
                                                     "light_int")] + [("shadow_count", C.POINTER(C.c_uint8))]


class oracle_diff_grads(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("albedo", "normal", "kd", "ks", "shininess", "light_pos",
                                                     "light_int")]


def _dptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def diff(scene, vis_order: np.ndarray, params, target=None, grads=True, shadow_count=None):
    """The FP64 reference of the differentiable shading at the visibility
    `vis_order` ((H*m, W*m) uint32, as Frame.vis_order): returns (loss,
    linear (H*m, W*m, 3) float64, gradients as a dict of float64 arrays or
    None).  `params` is a polyred_amd.diff.DiffParams; `shadow_count` (as
    Frame.shadow_count) scales each pixel's shading by 2^-count."""
    names = ("albedo", "normal", "kd", "ks", "shininess", "light_pos", "light_int")
    p64 = {n: (None if getattr(params, n) is None else np.ascontiguousarray(getattr(params, n), np.float64))
           for n in names}
    ps = oracle_diff_params(**{n: _dptr(a) for n, a in p64.items()})
    cnt = None if shadow_count is None else np.ascontiguousarray(shadow_count, np.uint8)
    if cnt is not None:
        ps.shadow_count = cnt.ctypes.data_as(C.POINTER(C.c_uint8))
    vis = np.ascontiguousarray(vis_order, np.uint32)
    H, W = vis.shape
    lin = np.empty((H, W, 3), np.float64)
    tgt = None if target is None else np.ascontiguousarray(target, np.float64)
    g = None
    gs = None
    if grads and tgt is not None:
        g = {n: np.zeros_like(p64["albedo"] if n == "normal" and p64[n] is None else p64[n]) for n in names}
        gs = oracle_diff_grads(**{n: _dptr(a) for n, a in g.items()})
    loss = C.c_double()
    desc, cam, opts = scene.desc(), scene.camera_struct(), scene.options_struct()
    rc = lib().oracle_diff(C.byref(desc), C.byref(cam), C.byref(opts), vis.ctypes.data_as(C.POINTER(C.c_uint32)),
                           C.byref(ps), _dptr(tgt), _dptr(lin), C.byref(gs) if gs is not None else None,
                           C.byref(loss))
    if rc != 0:
        raise RuntimeError(f"oracle_diff failed: {rc}")
    return loss.value, lin, g


def resize(img: np.ndarray, width: int, height: int) -> np.ndarray:
    """imageutil.Resize of an (ih, iw, 4) uint8 image (the oracle's restatement)."""
    img = np.ascontiguousarray(img, np.uint8)
    ih, iw = img.shape[:2]
    out = np.zeros((height, width, 4), np.uint8)
    rc = lib().oracle_resize(img.ctypes.data_as(C.POINTER(C.c_uint8)), iw, ih,
                             out.ctypes.data_as(C.POINTER(C.c_uint8)), width, height)
    if rc != 0:
        raise RuntimeError(f"oracle_resize failed: {rc}")
    return out


def _texture(img: np.ndarray, use_mipmap: bool):
    img = np.ascontiguousarray(img, np.uint8)
    t = abi.prh_texture()
    t.height, t.width = img.shape[0], img.shape[1]
    t.rgba = img.ctypes.data_as(C.POINTER(C.c_uint8))
    t.use_mipmap = int(use_mipmap)
    return t, img


def texture_query(img: np.ndarray, lod: float, u: float, v: float, use_mipmap: bool = True) -> tuple:
    """Texture.Query(lod, u, v) of NewTexture(TextureImage(img),
    TextureIsoMipmap(use_mipmap)) for an (h, w, 4) uint8 image."""
    t, keep = _texture(img, use_mipmap)
    out = (C.c_uint8 * 4)()
    rc = lib().oracle_texture_query(C.byref(t), lod, u, v, out)
    if rc != 0:
        raise RuntimeError(f"oracle_texture_query failed: {rc}")
    return tuple(out)


def texture_levels(img: np.ndarray, use_mipmap: bool = True) -> list:
    """The mip levels NewTexture builds from an (h, w, 4) uint8 image."""
    t, keep = _texture(img, use_mipmap)
    w, h = C.c_int(), C.c_int()
    n = lib().oracle_texture_level(C.byref(t), 0, C.byref(w), C.byref(h), None, 0)
    if n <= 0:
        raise RuntimeError(f"oracle_texture_level failed: {n}")
    out = []
    for lv in range(n):
        lib().oracle_texture_level(C.byref(t), lv, C.byref(w), C.byref(h), None, 0)
        a = np.zeros((h.value, w.value, 4), np.uint8)
        lib().oracle_texture_level(C.byref(t), lv, C.byref(w), C.byref(h),
                                   a.ctypes.data_as(C.POINTER(C.c_uint8)), a.nbytes)
        out.append(a)
    return out
