"""Batched inverse rendering (prh_fit_*) and the fixed shadow visibility of the
differentiable shading on the MI355X, through the C ABI.

  * Shadow counts (the lights whose shadow test reported true per pixel,
    render/shadow.go:230-283) equal the oracle's, exactly.
  * The adjoint with shadows against the FP64 reference (oracle_diff with
    the oracle's counts): loss within 1e-5, gradients within 1e-3 relative.
  * One fit step from `start` is one Adam step (Kingma & Ba 2015) on the
    sum over views of the single-view gradients: theta_1 = -lr sign(g) for
    p = start (1 + theta); two steps follow a float64 Adam on the host fed
    the single-view gradients at each step's parameters.
  * The stored G-buffer gives the rebuilt path's losses and parameters.
  * 200 steps recover materials and lights from a 30 %-perturbed start, in
    what the images determine (light intensities and Kd/Ks trade a common
    scale: their products are compared).
"""
import numpy as np
import pytest

import oracle
from polyred_amd import abi, diff, render, scenes

from .test_gpu_diff import _check_grads

pytestmark = pytest.mark.gpu

GROUPS = ("kd", "ks", "shininess", "light_pos", "light_int")


def _views(device, n_views=3, sides=(64, 112), n_spheres=16):
    s, views = scenes.c5_inverse(n_views=n_views, n_spheres=n_spheres, min_side=sides[0], max_side=sides[1],
                                 shadow_size=256)
    gs, rs = render.views_of(s, views, device)
    for r in rs:
        r.render_device()
    return s, views, gs, rs


def _order(r):
    vis = r.read(abi.PRH_BUF_VIS)
    order = (vis & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    order[vis == np.uint64(0xFFFFFFFFFFFFFFFF)] = 0xFFFFFFFF
    return order


def _perturbed(p, seed, sigma):
    rng = np.random.default_rng(seed)
    q = p.copy()
    for n in GROUPS:
        a = getattr(q, n)
        a *= (1 + sigma * rng.standard_normal(a.shape)).astype(a.dtype)
    return q


def _summed_backward(rs, p, targets):
    loss, G = 0.0, None
    for r, t in zip(rs, targets):
        l, g = r.Backward(p, t)
        loss += l
        G = g if G is None else diff.DiffParams(**{n: getattr(G, n) + getattr(g, n) for n in diff.DiffParams.NAMES})
    return loss, G


def test_shadow_counts_match_oracle(device, orc):
    s, views, gs, rs = _views(device)
    total = 0
    for (cam, w, h), r in zip(views, rs):
        s.camera, s.width, s.height = cam, w, h
        fr = orc.render(s)
        got = r.read(abi.PRH_BUF_SHADOW_COUNT)
        assert np.array_equal(got, fr.shadow_count), (w, h, int((got != fr.shadow_count).sum()))
        total += int((got > 0).sum())
    assert total > 100
    gs.free()


def test_shadowed_backward_matches_fp64(device, orc):
    s, views, gs, rs = _views(device, n_views=2)
    p0 = diff.scene_params(s)
    truth = _perturbed(p0, 3, 0.1)
    for (cam, w, h), r in zip(views, rs):
        s.camera, s.width, s.height = cam, w, h
        fr = orc.render(s)
        target = np.nan_to_num(r.ForwardLinear(truth))
        loss, g = r.Backward(p0, target)
        ref_loss, _, gref = oracle.diff(s, _order(r), p0, target.astype(np.float64), shadow_count=fr.shadow_count)
        assert abs(loss - ref_loss) <= 1e-5 * ref_loss, (loss, ref_loss)
        _check_grads(g, gref, names=GROUPS)
    gs.free()


def test_one_step_is_adam(device):
    s, views, gs, rs = _views(device)
    p0 = diff.scene_params(s)
    truth = _perturbed(p0, 5, 0.3)
    fit = diff.Fit(rs)
    fit.targets(truth)
    targets = [np.nan_to_num(r.ForwardLinear(truth)) for r in rs]
    loss, G = _summed_backward(rs, p0, targets)
    lr = 1e-3
    p1, hist = fit.run(p0, steps=1, lr=lr, groups=GROUPS)
    assert hist.shape == (1,) and abs(hist[0] - loss) <= 1e-5 * loss, (hist[0], loss)
    for n in GROUPS:
        base = getattr(p0, n).astype(np.float64).reshape(-1)
        g = getattr(G, n).astype(np.float64).reshape(-1) * base
        theta = np.where(base != 0, getattr(p1, n).reshape(-1) / np.where(base != 0, base, 1) - 1, 0.0)
        big = np.abs(g) > 1e-3 * np.abs(g).max()
        assert big.any(), n
        assert np.abs(theta[big] + lr * np.sign(g[big])).max() <= 1e-3 * lr, n
    # the untouched groups keep their values
    assert np.array_equal(p1.albedo, p0.albedo)
    fit.free()
    gs.free()


def _adam64(theta, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    theta = theta - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return theta, m, v


def test_two_steps_follow_host_adam(device):
    s, views, gs, rs = _views(device)
    p0 = diff.scene_params(s)
    truth = _perturbed(p0, 6, 0.3)
    fit = diff.Fit(rs)
    fit.targets(truth)
    targets = [np.nan_to_num(r.ForwardLinear(truth)) for r in rs]
    lr = 1e-2
    p1, _ = fit.run(p0, steps=1, lr=lr, groups=GROUPS)
    p2, hist = fit.run(p0, steps=2, lr=lr, groups=GROUPS)
    _, G0 = _summed_backward(rs, p0, targets)
    l1, G1 = _summed_backward(rs, p1, targets)
    assert abs(hist[1] - l1) <= 1e-4 * l1, (hist[1], l1)
    for n in GROUPS:
        base = getattr(p0, n).astype(np.float64).reshape(-1)
        th, m, v = np.zeros_like(base), np.zeros_like(base), np.zeros_like(base)
        g0 = getattr(G0, n).astype(np.float64).reshape(-1) * base
        g1 = getattr(G1, n).astype(np.float64).reshape(-1) * base
        th, m, v = _adam64(th, m, v, g0, 1, lr)
        th, m, v = _adam64(th, m, v, g1, 2, lr)
        want = base * (1 + th)
        got = getattr(p2, n).astype(np.float64).reshape(-1)
        big = (np.abs(g0) > 1e-2 * np.abs(g0).max()) & (np.abs(g1) > 1e-2 * np.abs(g1).max())
        assert big.any(), n
        assert np.abs(got - want)[big].max() <= 1e-3 * lr * np.abs(base[big]).max() + 1e-6, n
    fit.free()
    gs.free()


def test_stored_gbuffer_matches_rebuild(device):
    s, views, gs, rs = _views(device)
    p0 = diff.scene_params(s)
    truth = _perturbed(p0, 7, 0.3)
    a, b = diff.Fit(rs), diff.Fit(rs, stored_gbuffer=True)
    a.targets(truth)
    b.targets(truth)
    pa, ha = a.run(p0, steps=20, lr=1e-2, groups=GROUPS)
    pb, hb = b.run(p0, steps=20, lr=1e-2, groups=GROUPS)
    assert np.all(np.abs(ha - hb) <= 1e-4 * ha[0]), (ha, hb)
    for n in GROUPS:
        x, y = getattr(pa, n).astype(np.float64), getattr(pb, n).astype(np.float64)
        assert np.abs(x - y).max() <= 1e-4 * max(1.0, np.abs(x).max()), n
    sa, sb = a.stats(), b.stats()
    assert sa.steps == sb.steps == 20 and sa.n_views == len(rs)
    assert sb.visibility_bytes > 5 * sa.visibility_bytes and sb.ms_gbuffer > 0
    with pytest.raises(abi.PrhError) as e:
        b.run(p0, steps=1, groups=("albedo", "kd"))
    assert e.value.code == abi.PRH_E_UNSUPPORTED
    a.free()
    b.free()
    gs.free()


def test_fit_recovers_materials_and_lights(device):
    """Targets from the scene's own parameters, start 30 % off, 200 steps."""
    s, views, gs, rs = _views(device, n_views=6, sides=(96, 160), n_spheres=24)
    truth = diff.scene_params(s)
    start = _perturbed(truth, 9, 0.3)
    fit = diff.Fit(rs)
    fit.targets(truth)
    p, hist = fit.run(start, steps=200, lr=1e-2, groups=GROUPS)
    assert np.all(np.isfinite(hist)) and hist[-1] < 1e-2 * hist[0], (hist[0], hist[-1])
    # the loss falls 100-fold within 200 steps; the parameters follow more
    # slowly (light positions, intensities and Ks/shininess trade against
    # each other): every gauge-free error falls
    r0, r1 = diff.gauge_free_error(start, truth), diff.gauge_free_error(p, truth)
    for n in ("kd_x_intensity", "ks_x_intensity", "light_pos", "shininess"):
        assert r1[n] < 0.85 * r0[n], (n, r0, r1)
    fit.free()
    gs.free()


def test_rendering_again_invalidates(device):
    s, views, gs, rs = _views(device, n_views=2)
    fit = diff.Fit(rs)
    rs[1].render_device()
    with pytest.raises(abi.PrhError) as e:
        fit.targets(diff.scene_params(s))
    assert e.value.code == abi.PRH_E_INVALID
    fit.free()
    gs.free()
