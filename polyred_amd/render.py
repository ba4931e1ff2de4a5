"""Host-side mirror of polyred's render option API over libprh.

    ctx = render.Device(0)
    r = render.NewRenderer(render.Size(w, h), render.Camera(cam), render.Scene(scene),
                           render.ShadowMap(True), render.GammaCorrection(True),
                           render.HIP(ctx))
    img = r.Render()            # (H, W, 4) uint8, top row first

Option names, argument meaning and the pass-record follow render/options.go
(16-141) and render/raster.go (63-199).  ``HIP(dev)`` is the new option that
sits beside ``render.GPU(dev)``: every pass of the frame runs in libprh on the
given device.  There is no CPU fallback in this package -- a frame that libprh
cannot render raises ``abi.PrhError`` (the Go integration falls back to the
reference CPU passes instead, see the binding notes).
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import abi


_live_devices: "weakref.WeakSet" = weakref.WeakSet()


class Device:
    """An opened GPU (prh_open / prh_close): the context owns one HIP stream.
    Lifetimes nest as in the C ABI: closing a device first frees the scenes
    (and their renderers) still alive on it."""

    def __init__(self, ordinal: int = 0):
        self._lib = abi.load()
        h = C.c_void_p()
        abi.check(self._lib.prh_open(ordinal, C.byref(h)), "prh_open")
        self.handle = h
        self.ordinal = ordinal
        self._scenes: "weakref.WeakSet" = weakref.WeakSet()
        self._surfaces: "weakref.WeakSet" = weakref.WeakSet()
        _live_devices.add(self)

    def close(self) -> None:
        if self.handle:
            for sf in list(self._surfaces):
                sf.Release()
            for sc in list(self._scenes):
                sc.free()
            abi.check(self._lib.prh_close(self.handle), "prh_close")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GPUScene:
    """A scene uploaded to HBM (prh_scene_upload)."""

    def __init__(self, dev: Device, scene):
        self._lib = abi.load()
        self.dev = dev
        self.scene = scene
        self.handle = None
        self._renderers: "weakref.WeakSet" = weakref.WeakSet()
        self._fits: "weakref.WeakSet" = weakref.WeakSet()  # diff.Fit objects over its renderers
        desc = scene.desc()
        h = C.c_void_p()
        abi.check(self._lib.prh_scene_upload(dev.handle, C.byref(desc), C.byref(h)), "prh_scene_upload")
        self.handle = h
        dev._scenes.add(self)

    def free(self) -> None:
        if self.handle:
            for f in list(self._fits):
                f.free()
            for r in list(self._renderers):
                r.free()
            abi.check(self._lib.prh_scene_free(self.handle), "prh_scene_free")
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class _Option:  # render/options.go:16-33
    BatchSize: int = 32
    Workers: int = 0
    Width: int = 800
    Height: int = 600
    Background: tuple = (0, 0, 0, 0)
    MSAA: int = 1
    ShadowMap: bool = False
    ShadowSize: int = 0
    GammaCorrect: bool = False
    Camera: Optional[dict] = None
    Scene: object = None
    HIPDevice: Optional[Device] = None
    KeepLinear: bool = False
    extra: dict = field(default_factory=dict)


def Size(width: int, height: int):
    def f(o: _Option):
        o.Width, o.Height = width, height
    return f


def Camera(cam: dict):
    def f(o: _Option):
        o.Camera = cam
    return f


def Scene(s):
    def f(o: _Option):
        o.Scene = s
    return f


def Background(rgba):
    def f(o: _Option):
        o.Background = tuple(rgba)
    return f


def MSAA(n: int):
    def f(o: _Option):
        o.MSAA = n
    return f


def ShadowMap(enable: bool):
    def f(o: _Option):
        o.ShadowMap = enable
    return f


def ShadowSize(n: int):
    """Extension: shadow-map side in texels (0 = framebuffer-sized)."""
    def f(o: _Option):
        o.ShadowSize = n
    return f


def GammaCorrection(enable: bool):
    def f(o: _Option):
        o.GammaCorrect = enable
    return f


def BatchSize(n: int):  # accepted for API parity; the device path has no CPU batches
    def f(o: _Option):
        o.BatchSize = n
    return f


def Workers(n: int):  # accepted for API parity
    def f(o: _Option):
        o.Workers = n
    return f


def HIP(dev: Device):
    """Run every pass on `dev` through libprh (the option beside render.GPU)."""
    def f(o: _Option):
        o.HIPDevice = dev
    return f


def KeepLinear(enable: bool):
    """Also keep the pre-Round linear shading sums (parity/debug)."""
    def f(o: _Option):
        o.KeepLinear = enable
    return f


class Renderer:
    def __init__(self, cfg: _Option):
        self.cfg = cfg
        self._lib = abi.load()
        self.passGPU = {}
        if cfg.HIPDevice is None:
            raise ValueError("polyred_amd renders on the device only: pass render.HIP(dev)")
        if cfg.Scene is None or cfg.Camera is None:
            raise ValueError("NewRenderer needs Scene and Camera options")
        s = cfg.Scene
        s.width, s.height = cfg.Width, cfg.Height
        s.shadow_map, s.gamma, s.shadow_size = cfg.ShadowMap, cfg.GammaCorrect, cfg.ShadowSize
        s.background = cfg.Background
        s.camera = cfg.Camera
        s.msaa = cfg.MSAA
        self.handle = None
        shared = cfg.extra.get("gscene")
        self._owns_scene = shared is None
        self.gscene = shared if shared is not None else GPUScene(cfg.HIPDevice, s)
        cam = s.camera_struct()
        opts = s.options_struct(keep_linear=cfg.KeepLinear)
        opts.msaa = cfg.MSAA
        h = C.c_void_p()
        try:
            abi.check(self._lib.prh_renderer_new(cfg.HIPDevice.handle, self.gscene.handle, C.byref(cam),
                                                 C.byref(opts), C.byref(h)), "prh_renderer_new")
        except abi.PrhError:
            if self._owns_scene:
                self.gscene.free()
            raise
        self.handle = h
        self.gscene._renderers.add(self)
        self.W, self.H = cfg.Width, cfg.Height  # the output image
        self.msaa = max(1, cfg.MSAA)             # internal buffers are (H*msaa, W*msaa)
        self.n_shadow = sum(1 for l in s.lights if l.cast_shadow) if cfg.ShadowMap else 0
        self.S = (cfg.ShadowSize if cfg.ShadowSize > 0 else None)

    # (*Renderer).Render, render/raster.go:155-199
    def Render(self) -> np.ndarray:
        self.render_device()
        return self.image()

    def render_device(self) -> None:
        """One frame; the image stays in HBM."""
        abi.check(self._lib.prh_render(self.handle), "prh_render")
        for name in ("shadow", "forward", "deferred", "gamma"):
            self.passGPU[name] = True

    def passOnGPU(self, name: str) -> bool:  # render/raster.go:79
        return self.passGPU.get(name, False)

    def image(self, out: Optional[np.ndarray] = None) -> np.ndarray:
        """The frame's RGBA8 image (H, W, 4), top row first; into `out` when
        given (a pinned_array reads at the link's full rate)."""
        if out is None:
            out = np.empty((self.H, self.W, 4), np.uint8)
        if out.dtype != np.uint8 or out.shape != (self.H, self.W, 4) or not out.flags.c_contiguous:
            raise ValueError("image(out=...) needs a C-contiguous (H, W, 4) uint8 array")
        abi.check(self._lib.prh_read_image(self.handle, out.ctypes.data, out.nbytes), "prh_read_image")
        return out

    def read(self, which: int) -> np.ndarray:
        """An internal buffer, MSAA-sized (raster rows)."""
        W, H = self.W * self.msaa, self.H * self.msaa
        if which == abi.PRH_BUF_VIS:
            out = np.empty((H, W), np.uint64)
        elif which == abi.PRH_BUF_DEPTH:
            out = np.empty((H, W), np.float32)
        elif which == abi.PRH_BUF_LINEAR:
            out = np.empty((H, W, 3), np.float32)
        elif which == abi.PRH_BUF_PREGAMMA:
            out = np.empty((H, W, 4), np.uint8)
        elif which == abi.PRH_BUF_SHADOW_COUNT:
            out = np.empty((H, W), np.uint8)
        elif which == abi.PRH_BUF_SHADOW:
            sw, sh = (self.S, self.S) if self.S else (W, H)
            out = np.empty((max(1, self.n_shadow), sh, sw), np.float32)
            if self.n_shadow == 0:
                return out[:0]
        else:
            raise ValueError(which)
        abi.check(self._lib.prh_read_buffer(self.handle, which, out.ctypes.data, out.nbytes), "prh_read_buffer")
        return out

    # -------- differentiable shading (an extension: polyred has no adjoint)
    def ForwardLinear(self, params) -> np.ndarray:
        """Linear RGB (H*m, W*m, 3) of the shading with `params` (a
        diff.DiffParams) at the last frame's visibility; NaN where unshaded."""
        from . import diff
        return diff.forward(self._lib, self.handle, params, self.H * self.msaa, self.W * self.msaa)

    def Backward(self, params, target: np.ndarray):
        """(loss, gradients): loss = sum over shaded pixels of
        |ForwardLinear(params) - target|^2, gradients a diff.DiffParams."""
        from . import diff
        return diff.backward(self._lib, self.handle, params, target, self.H * self.msaa, self.W * self.msaa)

    # -------- autotuning (an extension: speed knobs, results unchanged)
    def tuning(self) -> abi.prh_tuning:
        t = abi.prh_tuning()
        abi.check(self._lib.prh_get_tuning(self.handle, C.byref(t)), "prh_get_tuning")
        return t

    def set_tuning(self, **knobs) -> None:
        """raster_wg_per_cu, item_prims, shadow_direct_max, main_direct_max,
        overlap_shadows, raster_tile."""
        t = self.tuning()
        for k, v in knobs.items():
            setattr(t, k, v)
        abi.check(self._lib.prh_set_tuning(self.handle, C.byref(t)), "prh_set_tuning")

    def Autotune(self, frames: int = 5) -> abi.prh_tuning:
        """Time the knobs' candidates (coordinate descent) and keep the fastest."""
        t = abi.prh_tuning()
        abi.check(self._lib.prh_autotune(self.handle, frames, C.byref(t)), "prh_autotune")
        for name in ("shadow", "forward", "deferred", "gamma"):
            self.passGPU[name] = True
        return t

    def stats(self) -> abi.prh_frame_stats:
        st = abi.prh_frame_stats()
        abi.check(self._lib.prh_last_stats(self.handle, C.byref(st)), "prh_last_stats")
        return st

    def free(self) -> None:
        if getattr(self, "handle", None):
            gs = getattr(self, "gscene", None)
            for f in list(gs._fits if gs is not None else ()):
                if self in f.views:
                    f.free()
            abi.check(self._lib.prh_renderer_free(self.handle), "prh_renderer_free")
            self.handle = None
        gs = getattr(self, "gscene", None)
        if gs is not None:
            self.gscene = None
            if getattr(self, "_owns_scene", True):
                gs.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@atexit.register
def _close_all() -> None:
    for d in list(_live_devices):
        try:
            d.close()
        except Exception:
            pass


def render_batch(renderers, graph: bool = False, streams: bool = False) -> float:
    """One frame of every renderer (all on one Device) in one batched
    submission (prh_render_batch): one launch per pass over every view, or
    with streams=True each view's own launches over the batch streams; with
    graph=True the batch is captured into a HIP graph once and replayed.
    Returns the batch's device time in ms.  Each renderer then holds the
    frame a Render() would have drawn."""
    rs = list(renderers)
    if not rs:
        return 0.0
    lib = rs[0]._lib
    arr = (C.c_void_p * len(rs))(*[r.handle for r in rs])
    ms = C.c_float()
    flags = (abi.PRH_BATCH_GRAPH if graph else 0) | (abi.PRH_BATCH_STREAMS if streams else 0)
    abi.check(lib.prh_render_batch(arr, len(rs), flags, C.byref(ms)), "prh_render_batch")
    for r in rs:
        for name in ("shadow", "forward", "deferred", "gamma"):
            r.passGPU[name] = True
    return ms.value


def autotune_batch(renderers, frames: int = 3):
    """prh_autotune_batch: the knobs of a wide batch's views timed by
    coordinate descent, per size class (views below / from 2^19 framebuffer
    pixels).  Every view keeps its class's winner; returns the two classes'
    prh_tuning (ms: the batch's median time)."""
    rs = list(renderers)
    lib = rs[0]._lib
    arr = (C.c_void_p * len(rs))(*[r.handle for r in rs])
    out = (abi.prh_tuning * 2)()
    abi.check(lib.prh_autotune_batch(arr, len(rs), frames, out), "prh_autotune_batch")
    for r in rs:
        for name in ("shadow", "forward", "deferred", "gamma"):
            r.passGPU[name] = True
    return out[0], out[1]


def NewRenderer(*opts) -> Renderer:  # render/raster.go:84-143
    cfg = _Option()
    for opt in opts:
        opt(cfg)
    return Renderer(cfg)


class Surface:
    """gpu.Surface's headless swapchain (gpu/surface.go:20-183) on the
    device: a ring of RGBA8 textures in HBM that renderers draw into, one per
    frame (prh_surface_*).

        surf = render.CreateSurface(dev, 1920, 1080)      # Frames 2
        i = surf.AcquireNextTexture()
        surf.Render(r)            # r's frame drawn into texture i
        surf.Present()
        pix = surf.Texture()      # the presented frame's pixels
    """

    def __init__(self, dev: Device, width: int, height: int, frames: int = 2):
        self._lib = abi.load()
        self.dev = dev
        self.handle = None
        h = C.c_void_p()
        abi.check(self._lib.prh_surface_new(dev.handle, width, height, frames, C.byref(h)), "prh_surface_new")
        self.handle = h
        dev._surfaces.add(self)

    def AcquireNextTexture(self) -> int:
        i = C.c_int32()
        abi.check(self._lib.prh_surface_acquire(self.handle, C.byref(i)), "prh_surface_acquire")
        return i.value

    def Render(self, r: "Renderer") -> None:
        abi.check(self._lib.prh_surface_render(self.handle, r.handle), "prh_surface_render")
        for name in ("shadow", "forward", "deferred", "gamma"):
            r.passGPU[name] = True

    def Present(self) -> None:
        abi.check(self._lib.prh_surface_present(self.handle), "prh_surface_present")

    def PresentImage(self, img: np.ndarray) -> None:
        """Acquire, upload a (h, w, 4) uint8 image, present (surface.go:133-137)."""
        self.AcquireNextTexture()
        a = np.ascontiguousarray(img, np.uint8)
        abi.check(self._lib.prh_surface_write(self.handle, a.ctypes.data, a.nbytes), "prh_surface_write")
        self.Present()

    def Texture(self) -> np.ndarray:
        """The most recently presented frame, (h, w, 4) uint8, top row first."""
        w, h = self.Size()
        out = np.empty((h, w, 4), np.uint8)
        abi.check(self._lib.prh_surface_read(self.handle, out.ctypes.data, out.nbytes), "prh_surface_read")
        return out

    def Resize(self, width: int, height: int) -> None:
        abi.check(self._lib.prh_surface_resize(self.handle, width, height), "prh_surface_resize")

    def Size(self) -> tuple:
        w, h = C.c_int32(), C.c_int32()
        abi.check(self._lib.prh_surface_size(self.handle, C.byref(w), C.byref(h)), "prh_surface_size")
        return w.value, h.value

    def Release(self) -> None:
        if self.handle:
            abi.check(self._lib.prh_surface_free(self.handle), "prh_surface_free")
            self.handle = None

    def __del__(self):
        try:
            self.Release()
        except Exception:
            pass


def CreateSurface(dev: Device, width: int, height: int, frames: int = 2) -> Surface:
    """Device.CreateSurface (gpu/surface.go:43-57)."""
    return Surface(dev, width, height, frames)


def pinned_array(shape, dtype=np.uint8) -> np.ndarray:
    """A numpy array in page-locked host memory (prh_host_alloc), freed with
    the array."""
    lib = abi.load()
    n = int(np.prod(shape)) * np.dtype(dtype).itemsize
    p = C.c_void_p()
    abi.check(lib.prh_host_alloc(n, C.byref(p)), "prh_host_alloc")
    buf = (C.c_uint8 * n).from_address(p.value)
    arr = np.frombuffer(buf, np.uint8).view(dtype).reshape(shape)
    weakref.finalize(arr, lib.prh_host_free, p.value)
    return arr


def views_of(scene, views, dev: Device):
    """Renderers of one scene uploaded once (prh_scene_upload), one per
    (camera dict, width, height) of `views`, with the scene's other
    settings.  Returns (GPUScene, renderers); freeing the GPUScene frees
    them."""
    gs = GPUScene(dev, scene)

    def share(o: _Option):
        o.extra["gscene"] = gs

    rs = []
    for cam, w, h in views:
        rs.append(NewRenderer(Size(w, h), Camera(cam), Scene(scene), Background(scene.background),
                              ShadowMap(scene.shadow_map), ShadowSize(scene.shadow_size),
                              GammaCorrection(scene.gamma), MSAA(getattr(scene, "msaa", 1)), HIP(dev), share))
    return gs, rs


def renderer_for(scene, dev: Device, keep_linear: bool = False) -> Renderer:
    """A renderer configured from a generated scene's own settings."""
    return NewRenderer(Size(scene.width, scene.height), Camera(scene.camera), Scene(scene),
                       Background(scene.background), ShadowMap(scene.shadow_map), ShadowSize(scene.shadow_size),
                       GammaCorrection(scene.gamma), MSAA(getattr(scene, "msaa", 1)), KeepLinear(keep_linear),
                       HIP(dev))
