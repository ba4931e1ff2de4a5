"""libprh's differentiable shading on the MI355X against the FP64 reference
(oracle_diff), through the C ABI.

The bar (float32 kernel against float64): the linear shading within 1e-5 of
the largest value, the loss within 1e-5 relative, and every gradient group
(albedo, normals, Kd, Ks, shininess, light positions, light intensities)
within 1e-3 relative in L2 norm -- with each material's and light's entries
also within 1e-3 of their group's largest entry.
"""
import numpy as np
import pytest

import oracle
from polyred_amd import abi, diff, render, scenes

pytestmark = pytest.mark.gpu


def _setup(device, scene, seed=11, sigma=0.1):
    r = render.renderer_for(scene, device)
    r.render_device()
    vis = r.read(abi.PRH_BUF_VIS)
    order = (vis & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    order[vis == np.uint64(0xFFFFFFFFFFFFFFFF)] = 0xFFFFFFFF
    p0 = diff.scene_params(scene)
    truth = diff.perturb(p0, np.random.default_rng(seed), sigma)
    return r, order, p0, truth


def _check_grads(g, gref, names=diff.DiffParams.NAMES):
    for n in names:
        a = np.asarray(getattr(g, n), np.float64).reshape(-1)
        b = np.asarray(gref[n], np.float64).reshape(-1)
        nb = np.linalg.norm(b)
        assert nb > 0, n
        rel = np.linalg.norm(a - b) / nb
        assert rel <= 1e-3, (n, rel)
        if n not in ("albedo", "normal"):
            assert np.abs(a - b).max() <= 1e-3 * np.abs(b).max(), (n, np.abs(a - b).max(), np.abs(b).max())


@pytest.mark.parametrize("which", ["small", "c4"])
def test_forward_matches_fp64(device, orc, which):
    s = scenes.c4_adjoint(width=96, height=80, n_spheres=20) if which == "small" else scenes.c4_adjoint()
    r, order, p0, truth = _setup(device, s)
    lin = r.ForwardLinear(truth)
    _, ref, _ = oracle.diff(s, order, truth)
    shaded = ~np.isnan(ref).any(-1)
    assert np.array_equal(shaded, ~np.isnan(lin).any(-1))
    err = np.abs(lin[shaded].astype(np.float64) - ref[shaded])
    assert err.max() <= 1e-5 * np.abs(ref[shaded]).max(), err.max()
    r.free()


@pytest.mark.parametrize("which", ["small", "c4"])
def test_backward_matches_fp64(device, orc, which):
    """The C4 configuration: 512x512, ~100k triangles, 4 point lights, no
    shadows; target rendered from seed parameters, gradients at a 10 %
    Gaussian perturbation of them."""
    s = scenes.c4_adjoint(width=96, height=80, n_spheres=20) if which == "small" else scenes.c4_adjoint()
    r, order, p0, truth = _setup(device, s)
    target = np.nan_to_num(r.ForwardLinear(truth))
    loss, g = r.Backward(p0, target)
    ref_loss, _, gref = oracle.diff(s, order, p0, target.astype(np.float64))
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss, (loss, ref_loss)
    _check_grads(g, gref)
    st = r.stats()
    assert st.ms_diff > 0
    r.free()


def test_explicit_normals(device, orc):
    s = scenes.c4_adjoint(width=64, height=64, n_spheres=12)
    r, order, p0, truth = _setup(device, s)
    p = p0.copy()
    p.normal = np.concatenate([s.meshes[m].nor[:, :3] for m in s.obj_mesh]).astype(np.float32)
    target = np.nan_to_num(r.ForwardLinear(truth))
    loss, g = r.Backward(p, target)
    ref_loss, _, gref = oracle.diff(s, order, p, target.astype(np.float64))
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss
    _check_grads(g, gref)
    r.free()


def test_needs_a_frame_and_at_most_8_lights(device):
    s = scenes.c4_adjoint(width=32, height=32, n_spheres=4)
    r = render.renderer_for(s, device)
    with pytest.raises(abi.PrhError) as e:
        r.ForwardLinear(diff.scene_params(s))
    assert e.value.code == abi.PRH_E_INVALID
    r.free()
    s.lights = s.lights * 3  # 12 lights
    r = render.renderer_for(s, device)
    r.render_device()
    with pytest.raises(abi.PrhError) as e:
        r.ForwardLinear(diff.scene_params(s))
    assert e.value.code == abi.PRH_E_UNSUPPORTED
    r.free()
