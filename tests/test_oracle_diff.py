"""The FP64 reference of the differentiable shading (oracle_diff).

polyred has no derivative path, so nothing in the reference pins this
extension directly.  The FP64 reference is pinned by:
  * central finite differences of its own FP64 loss, for every kind of
    parameter (albedo, normals, Kd, Ks, shininess, point-light position and
    intensity, directional-light intensity) -- a wrong sign, a dropped term
    or a transposed index in the adjoint fails them;
  * its forward: with every vertex's albedo equal to the material's texture
    colour it is the reference's own linear shading sums (FragmentShader
    before Round, the oracle's restatement of shader/blinn_cpu.go:24-106).
"""
import numpy as np
import pytest

import oracle
from polyred_amd import abi, diff, scenes


def small_scene():
    s = scenes.c4_adjoint(width=48, height=40, n_spheres=10)
    s.materials[1].flat_shading = True  # the face normal: no normal gradient
    s.lights.append(scenes.Light(kind=abi.PRH_LIGHT_DIRECTIONAL, direction=scenes.unit3((0.2, -1.0, -0.4)),
                                 intensity=0.7, color=(250, 240, 230, 255)))
    return s


def params64(s):
    p = diff.scene_params(s).astype(np.float64)
    # explicit per-vertex normals (any values define the function): the mesh
    # normals of each object
    p.normal = np.concatenate([s.meshes[m].nor[:, :3] for m in s.obj_mesh]).astype(np.float64)
    return p


def test_gradients_match_central_differences(orc):
    s = small_scene()
    fr = orc.render(s)
    p0 = params64(s)
    rng = np.random.default_rng(7)
    truth = diff.perturb(p0, rng, 0.1)
    _, target, _ = oracle.diff(s, fr.vis_order, truth)
    target = np.nan_to_num(target)
    loss, _, g = oracle.diff(s, fr.vis_order, p0, target)
    assert loss > 0

    def L(p):
        return oracle.diff(s, fr.vis_order, p, target, grads=False)[0]

    checked = 0
    for name in diff.DiffParams.NAMES:
        gg = g[name].reshape(-1)
        flat = getattr(p0, name).reshape(-1)
        # the entries with the largest gradients, plus two at random
        idx = list(np.argsort(-np.abs(gg))[:4]) + list(rng.integers(0, gg.size, 2))
        scale = np.abs(gg).max()
        assert scale > 0, name
        for i in idx:
            h = 1e-6 * max(1.0, abs(flat[i]))
            pp, pm = p0.copy(), p0.copy()
            getattr(pp, name).reshape(-1)[i] += h
            getattr(pm, name).reshape(-1)[i] -= h
            fd = (L(pp) - L(pm)) / (2 * h)
            assert abs(fd - gg[i]) <= 1e-5 * scale + 1e-4 * abs(gg[i]), (name, i, fd, gg[i])
            checked += 1
    assert checked >= 40


def test_forward_is_the_reference_shading_at_texture_albedo(orc):
    """Albedo = texture colour at every vertex: the linear sums equal the
    reference shading's (which reads the uniform texture colour)."""
    s = scenes.c1_plain(width=64, height=48)
    tex = (200, 180, 160, 255)
    for m in s.materials:
        m.texture = tex
    fr = orc.render(s)
    p = diff.scene_params(s).astype(np.float64)
    p.albedo[:] = tex[:3]
    _, lin, _ = oracle.diff(s, fr.vis_order, p)
    ref = fr.linear
    # covered pixels (uncovered ones copy pixel (0,0)'s shading in the
    # reference, render/raster.go:325-328; the differentiable pass skips them)
    shaded = (fr.vis_order != 0xFFFFFFFF) & ~np.isnan(ref).any(-1)
    assert shaded.sum() > 500
    assert np.array_equal(shaded, ~np.isnan(lin).any(-1))
    err = np.abs(lin[shaded] - ref[shaded])
    assert err.max() <= 1e-4 * max(1.0, np.abs(ref[shaded]).max()), err.max()


def test_loss_is_zero_at_the_target(orc):
    s = small_scene()
    fr = orc.render(s)
    p = params64(s)
    _, lin, _ = oracle.diff(s, fr.vis_order, p)
    loss, _, g = oracle.diff(s, fr.vis_order, p, np.nan_to_num(lin))
    assert loss == 0.0
    assert all(np.all(a == 0) for a in g.values())
