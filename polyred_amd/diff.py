"""Differentiable shading over libprh: the forward and the adjoint of the
deferred Blinn-Phong shading at the fixed visibility of a rendered frame.

    r = render.renderer_for(scene, dev)
    r.Render()                                   # fixes the visibility
    p = diff.scene_params(scene)                 # the scene's own values
    lin = r.ForwardLinear(p)                     # (H, W, 3) linear RGB, NaN unshaded
    loss, g = r.Backward(p, target)              # L = sum (lin - target)^2, dL/dp

polyred has no derivative path: the shading is FragmentShader's sums before
Round (shader/blinn_cpu.go:24-106) with real-valued parameters -- a per-vertex
albedo interpolated perspective-correctly where the reference reads the
uniform texture colour, per-vertex normals, per-material Kd/Ks (0..255) and
shininess, per-light position and intensity.  The C ABI is
prh_diff_sizes / prh_diff_forward / prh_diff_backward (include/prh.h).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, replace
from typing import Optional

import numpy as np

from . import abi

F32 = np.float32


@dataclass
class DiffParams:
    albedo: np.ndarray               # (n_vertices, 3) per transformed vertex
    normal: Optional[np.ndarray]     # (n_vertices, 3), or None: the scene's normals
    kd: np.ndarray                   # (n_materials, 3)
    ks: np.ndarray                   # (n_materials, 3)
    shininess: np.ndarray            # (n_materials,)
    light_pos: np.ndarray            # (n_lights, 3)
    light_int: np.ndarray            # (n_lights,)

    NAMES = ("albedo", "normal", "kd", "ks", "shininess", "light_pos", "light_int")

    def arrays(self):
        return {n: getattr(self, n) for n in self.NAMES if getattr(self, n) is not None}

    def astype(self, dt) -> "DiffParams":
        return DiffParams(**{n: (None if getattr(self, n) is None else np.ascontiguousarray(getattr(self, n), dt))
                             for n in self.NAMES})

    def copy(self) -> "DiffParams":
        return replace(self, **{n: getattr(self, n).copy() for n in self.NAMES if getattr(self, n) is not None})


def scene_params(scene) -> DiffParams:
    """The parameters a scene renders with: albedo = each transformed vertex's
    mesh vertex colour, the scene's normals, the materials' Kd/Ks/shininess,
    the lights' positions and intensities."""
    albedo = np.concatenate([scene.meshes[m].col[:, :3] for m in scene.obj_mesh]).astype(F32)
    mats = scene.materials
    return DiffParams(
        albedo=np.ascontiguousarray(albedo),
        normal=None,
        kd=np.array([m.diffuse[:3] for m in mats], F32).reshape(-1, 3),
        ks=np.array([m.specular[:3] for m in mats], F32).reshape(-1, 3),
        shininess=np.array([m.shininess for m in mats], F32),
        light_pos=np.array([l.position for l in scene.lights], F32).reshape(-1, 3),
        light_int=np.array([l.intensity for l in scene.lights], F32),
    )


def perturb(p: DiffParams, rng: np.random.Generator, sigma: float = 0.1) -> DiffParams:
    """Gaussian noise with a standard deviation of `sigma` times each value."""
    out = p.copy()
    for n, a in out.arrays().items():
        a *= (1 + sigma * rng.standard_normal(a.shape)).astype(a.dtype)
    return out


def _f32p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


class _Bound:
    """Keeps the float32 arrays a prh_diff_params / prh_diff_grads points at alive."""

    def __init__(self, p: DiffParams, cls):
        self.p = p.astype(F32)
        self.s = cls(**{n: _f32p(getattr(self.p, n)) for n in DiffParams.NAMES})


def sizes(lib, handle) -> tuple:
    nv, nm, nl = C.c_int32(), C.c_int32(), C.c_int32()
    abi.check(lib.prh_diff_sizes(handle, C.byref(nv), C.byref(nm), C.byref(nl)), "prh_diff_sizes")
    return nv.value, nm.value, nl.value


def check_shapes(p: DiffParams, nv: int, nm: int, nl: int) -> None:
    want = dict(albedo=(nv, 3), normal=(nv, 3), kd=(nm, 3), ks=(nm, 3), shininess=(nm,), light_pos=(nl, 3),
                light_int=(nl,))
    for n, a in p.arrays().items():
        if tuple(a.shape) != want[n]:
            raise ValueError(f"DiffParams.{n} has shape {a.shape}, the scene needs {want[n]}")


def forward(lib, handle, p: DiffParams, H: int, W: int) -> np.ndarray:
    check_shapes(p, *sizes(lib, handle))
    b = _Bound(p, abi.prh_diff_params)
    out = np.empty((H, W, 3), F32)
    abi.check(lib.prh_diff_forward(handle, C.byref(b.s), out.ctypes.data, out.nbytes), "prh_diff_forward")
    return out


def backward(lib, handle, p: DiffParams, target: np.ndarray, H: int, W: int, noalloc: bool = False):
    """(loss, gradients) through prh_diff_backward, or with `noalloc` through
    prh_diff_backward_noalloc (no device allocation: the working memory is
    the library's fixed per-GPU scratch array)."""
    nv, nm, nl = sizes(lib, handle)
    check_shapes(p, nv, nm, nl)
    target = np.ascontiguousarray(target, F32)
    if target.shape != (H, W, 3):
        raise ValueError(f"target has shape {target.shape}, the frame is {(H, W, 3)}")
    b = _Bound(p, abi.prh_diff_params)
    g = DiffParams(albedo=np.zeros((nv, 3), F32), normal=np.zeros((nv, 3), F32), kd=np.zeros((nm, 3), F32),
                   ks=np.zeros((nm, 3), F32), shininess=np.zeros(nm, F32), light_pos=np.zeros((nl, 3), F32),
                   light_int=np.zeros(nl, F32))
    gs = abi.prh_diff_grads(**{n: _f32p(getattr(g, n)) for n in DiffParams.NAMES})
    loss = C.c_double()
    fn = "prh_diff_backward_noalloc" if noalloc else "prh_diff_backward"
    abi.check(getattr(lib, fn)(handle, C.byref(b.s), target.ctypes.data, target.nbytes, C.byref(loss), C.byref(gs)),
              fn)
    return loss.value, g


__all__ = ["DiffParams", "scene_params", "perturb", "forward", "backward", "sizes", "Fit", "rmse", "gauge_free_error"]


# ------------------------------------------------ batched inverse rendering
GROUPS = {"albedo": abi.PRH_FIT_ALBEDO, "kd": abi.PRH_FIT_KD, "ks": abi.PRH_FIT_KS,
          "shininess": abi.PRH_FIT_SHININESS, "light_pos": abi.PRH_FIT_LIGHT_POS, "light_int": abi.PRH_FIT_LIGHT_INT}


class Fit:
    """Gradient descent on the shading parameters over many views of one
    scene (prh_fit_*): each view's last frame fixes its visibility and its
    shadow visibility; one step is the adjoint of every view in one launch
    plus an Adam update on the device, the steps one replayed HIP graph.

        gs, views = render.views_of(scene, cams, dev)
        for r in views: r.render_device()
        fit = diff.Fit(views)
        fit.targets(truth)                       # or fit.set_target(i, image)
        p, hist = fit.run(start, steps=200, lr=1e-2, groups=("kd", "ks", ...))

    polyred has no derivative path; the shading is prh_diff_*'s with each
    pixel scaled by 2^-count, count the lights whose shadow test reported
    true (render/shadow.go:230-283) as the reference scales the colour
    (render/raster.go:339-347)."""

    def __init__(self, renderers, stored_gbuffer: bool = False):
        self.views = list(renderers)
        if not self.views:
            raise ValueError("Fit needs at least one view")
        self._lib = self.views[0]._lib
        self.sizes = sizes(self._lib, self.views[0].handle)
        arr = (C.c_void_p * len(self.views))(*[r.handle for r in self.views])
        h = C.c_void_p()
        abi.check(self._lib.prh_fit_new(arr, len(self.views), abi.PRH_FIT_STORED_GBUFFER if stored_gbuffer else 0,
                                        C.byref(h)), "prh_fit_new")
        self.handle = h
        gs = self.views[0].gscene
        if gs is not None:
            gs._fits.add(self)  # freed before the scene's renderers

    def targets(self, truth: DiffParams) -> None:
        """Every view's target := its forward shading at `truth`, on the device."""
        check_shapes(truth, *self.sizes)
        b = _Bound(truth, abi.prh_diff_params)
        abi.check(self._lib.prh_fit_targets(self.handle, C.byref(b.s)), "prh_fit_targets")

    def set_target(self, i: int, target: np.ndarray) -> None:
        r = self.views[i]
        t = np.ascontiguousarray(target, F32)
        if t.shape != (r.H * r.msaa, r.W * r.msaa, 3):
            raise ValueError(f"target has shape {t.shape}")
        abi.check(self._lib.prh_fit_set_target(self.handle, i, t.ctypes.data, t.nbytes), "prh_fit_set_target")

    def run(self, start: DiffParams, steps: int = 200, lr: float = 1e-2, groups=("kd", "ks", "shininess",
            "light_pos", "light_int"), beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        """Adam from `start`; returns (final parameters, loss before each step)."""
        nv, nm, nl = self.sizes
        check_shapes(start, nv, nm, nl)
        b = _Bound(start, abi.prh_diff_params)
        o = abi.prh_fit_options(steps=steps, groups=sum(GROUPS[g] for g in groups), lr=lr, beta1=beta1,
                                beta2=beta2, eps=eps)
        out = DiffParams(albedo=np.zeros((nv, 3), F32), normal=None, kd=np.zeros((nm, 3), F32),
                         ks=np.zeros((nm, 3), F32), shininess=np.zeros(nm, F32), light_pos=np.zeros((nl, 3), F32),
                         light_int=np.zeros(nl, F32))
        gs = abi.prh_diff_grads(**{n: (_f32p(getattr(out, n)) if getattr(out, n) is not None else None)
                                   for n in DiffParams.NAMES})
        hist = np.zeros(max(1, steps), np.float64)
        abi.check(self._lib.prh_fit_run(self.handle, C.byref(b.s), C.byref(o), C.byref(gs), hist.ctypes.data),
                  "prh_fit_run")
        return out, hist[:steps]

    def stats(self) -> abi.prh_fit_stats:
        st = abi.prh_fit_stats()
        abi.check(self._lib.prh_fit_last_stats(self.handle, C.byref(st)), "prh_fit_last_stats")
        return st

    def free(self) -> None:
        if getattr(self, "handle", None):
            abi.check(self._lib.prh_fit_free(self.handle), "prh_fit_free")
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def rmse(a: DiffParams, b: DiffParams, names=("kd", "ks", "shininess", "light_pos", "light_int")) -> dict:
    """Root-mean-square error of `a` against `b` per parameter group."""
    return {n: float(np.sqrt(np.mean((np.asarray(getattr(a, n), np.float64) -
                                      np.asarray(getattr(b, n), np.float64)) ** 2))) for n in names}


def gauge_free_error(a: DiffParams, b: DiffParams) -> dict:
    """Errors of `a` against `b` in what the images determine.  The shading
    is unchanged when every light intensity is multiplied by s and every Kd
    and Ks divided by s, so the products Kd[m] I[l] and Ks[m] I[l] are
    compared (relative RMSE) rather than the factors; shininess and light
    positions as they are (RMSE)."""
    def rel(x, y):
        x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
        return float(np.sqrt(np.mean(((x - y) / np.where(y != 0, y, 1)) ** 2)))
    ia, ib = np.asarray(a.light_int, np.float64), np.asarray(b.light_int, np.float64)
    out = {"kd_x_intensity": rel(np.multiply.outer(a.kd, ia), np.multiply.outer(b.kd, ib)),
           "ks_x_intensity": rel(np.multiply.outer(a.ks, ia), np.multiply.outer(b.ks, ib))}
    out.update(rmse(a, b, ("shininess", "light_pos")))
    return out
