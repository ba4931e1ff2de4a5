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


def backward(lib, handle, p: DiffParams, target: np.ndarray, H: int, W: int):
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
    abi.check(lib.prh_diff_backward(handle, C.byref(b.s), target.ctypes.data, target.nbytes, C.byref(loss),
                                    C.byref(gs)), "prh_diff_backward")
    return loss.value, g


__all__ = ["DiffParams", "scene_params", "perturb", "forward", "backward", "sizes"]
