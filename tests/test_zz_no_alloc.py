"""prh_diff_backward_noalloc (the adjoint with no device allocation: its
working memory is the library's fixed per-GPU scratch array) against
prh_diff_backward on the same renderers, through the C ABI.

The two run the same kernels; only where the parameters, gradients, target,
loss and shadow counts live differs.  Gradients are summed with atomics, so
their order (and the last bits of a sum) may differ between any two runs:
the bar is the loss within 1e-12 relative and every gradient group within
1e-4 of its largest entry (the adjoint's own bar against FP64 is 1e-3).
Calls follow each other on one GPU with shrinking frames, so contents a
larger call left in the array must not leak into a smaller one.
"""
import numpy as np
import pytest

from polyred_amd import abi, diff, render, scenes

pytestmark = pytest.mark.gpu


def _close(a, b):
    (la, ga), (lb, gb) = a, b
    assert abs(la - lb) <= 1e-12 * abs(lb), (la, lb)
    for n in diff.DiffParams.NAMES:
        x = np.asarray(getattr(ga, n), np.float64)
        y = np.asarray(getattr(gb, n), np.float64)
        assert x.shape == y.shape, n
        scale = np.abs(y).max() if y.size else 0.0
        assert np.abs(x - y).max(initial=0.0) <= 1e-4 * scale, (n, np.abs(x - y).max(), scale)


def _target(r, p, seed):
    q = diff.perturb(p, np.random.default_rng(seed), 0.1)
    return np.nan_to_num(r.ForwardLinear(q))


def test_noalloc_matches_backward(device):
    # two views of one shadowed scene (shadow counts in the scratch array),
    # the second smaller, then a smaller unshadowed frame
    s, cams = scenes.c5_inverse(n_views=2, n_spheres=16, min_side=64, max_side=160, shadow_size=256)
    cams = [(cams[0][0], 160, 120), (cams[1][0], 96, 64)]
    gs, rs = render.views_of(s, cams, device)
    s4 = scenes.c4_adjoint(width=64, height=48, n_spheres=12)
    r4 = render.renderer_for(s4, device)
    runs = []
    for r, sc, seed in ((rs[0], s, 1), (rs[1], s, 2), (r4, s4, 3)):
        r.render_device()
        p = diff.scene_params(sc)
        runs.append((r, p, _target(r, p, seed)))
    ref = [diff.backward(r._lib, r.handle, p, t, r.H * r.msaa, r.W * r.msaa) for r, p, t in runs]
    got = [diff.backward(r._lib, r.handle, p, t, r.H * r.msaa, r.W * r.msaa, noalloc=True) for r, p, t in runs]
    for a, b in zip(got, ref):
        _close(a, b)
    assert ref[0][0] > 0 and ref[1][0] > 0
    # the first frame again, after the smaller ones
    r, p, t = runs[0]
    _close(diff.backward(r._lib, r.handle, p, t, r.H * r.msaa, r.W * r.msaa, noalloc=True), ref[0])
    r4.free()
    for r in rs:
        r.free()
    gs.free()


def test_noalloc_rejects_a_frame_beyond_the_scratch(device):
    # 2048 x 1536: the 36 MiB target alone exceeds the 32 MiB array
    s = scenes.c4_adjoint(width=2048, height=1536, n_spheres=4)
    r = render.renderer_for(s, device)
    r.render_device()
    p = diff.scene_params(s)
    t = np.zeros((1536, 2048, 3), np.float32)
    with pytest.raises(abi.PrhError) as e:
        diff.backward(r._lib, r.handle, p, t, 1536, 2048, noalloc=True)
    assert e.value.code == abi.PRH_E_INVALID
    assert "scratch" in str(e.value)
    # the renderer and the allocating entry point are unaffected
    loss, _ = diff.backward(r._lib, r.handle, p, t, 1536, 2048)
    assert loss >= 0
    r.free()
