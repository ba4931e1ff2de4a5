"""The batched workloads at their full size, gated on correctness.

* 64-view batched forward (the batch the bench times: 256^2 to 1920x1080,
  10k-500k triangles, 1-8 shadow-casting lights with 1024^2 maps): every view
  of the wide batch, eager and replayed from its HIP graph, is bit-identical
  to that view's own prh_render; three views' single renders (the one with
  the most shadow maps, the largest and the one with the most triangles)
  equal the oracle, their visibility, depth and shadow maps over every pixel
  and their shading on a seeded sample of 4096 pixels.
* Batched inverse rendering (32 views of the 250k-triangle scene, 4
  shadow-casting lights): one fit step's loss is the sum of the 32 views'
  prh_diff_backward losses, and its update, inverted, gives the sum of their
  gradients.
"""
import hashlib

import numpy as np
import pytest

from polyred_amd import abi, diff, render, scenes

from .test_gpu_parity import assert_frame_equal, gpu_frame

pytestmark = pytest.mark.gpu


def _digest(r):
    h = hashlib.sha256()
    for a in (r.image(), r.read(abi.PRH_BUF_VIS), r.read(abi.PRH_BUF_SHADOW)):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_c3_full_batch_matches_single_renders_and_oracle(device, orc):
    views = scenes.c3_views()
    assert len(views) == 64
    rs = [render.renderer_for(s, device) for s in views]
    render.render_batch(rs)
    want = [_digest(r) for r in rs]
    render.render_batch(rs, graph=True)
    render.render_batch(rs, graph=True)
    assert [_digest(r) for r in rs] == want, "graph replay differs from the eager batch"
    for i, r in enumerate(rs):
        r.render_device()
        assert _digest(r) == want[i], f"view {i}: batch differs from its own prh_render"
    for r in rs:
        r.free()
    most_lights = max(range(64), key=lambda i: (len(views[i].lights), views[i].n_pixels))
    largest = max(range(64), key=lambda i: views[i].n_pixels)
    most_tris = max(range(64), key=lambda i: views[i].n_triangles)
    rng = np.random.default_rng(64)
    for i in sorted({most_lights, largest, most_tris}):
        s = views[i]
        mask = np.zeros(s.height * s.width, np.uint8)
        mask[rng.choice(mask.size, 4096, replace=False)] = 1
        g = gpu_frame(device, s)
        fr = orc.render(s, mask=mask)
        assert_frame_equal(g, fr, mask=mask, scene=s)


def test_c5_full_fit_step_sums_the_views(device):
    s, cams = scenes.c5_inverse()
    assert len(cams) == 32
    gs, rs = render.views_of(s, cams, device)
    render.render_batch(rs)
    p0 = diff.scene_params(s)
    truth = p0.copy()
    rng = np.random.default_rng(55)
    groups = ("kd", "ks", "shininess", "light_pos", "light_int")
    for n in groups:
        a = getattr(truth, n)
        a *= (1 + 0.3 * rng.standard_normal(a.shape)).astype(a.dtype)
    fit = diff.Fit(rs)
    fit.targets(truth)
    loss, G = 0.0, None
    for r in rs:
        l, g = r.Backward(p0, np.nan_to_num(r.ForwardLinear(truth)))
        loss += l
        G = g if G is None else diff.DiffParams(**{n: getattr(G, n) + getattr(g, n) for n in diff.DiffParams.NAMES})
    lr = 1e-2
    for n in groups:
        # Adam's first step is theta = -lr * g / (|g| + eps) for g the
        # gradient with respect to theta (p = base * (1 + theta)): with eps
        # ten times the largest |g| the step stays small and linear in g,
        # and g = -sign(theta) |theta| eps / (lr - |theta|)
        base = getattr(p0, n).astype(np.float64).reshape(-1)
        want = getattr(G, n).astype(np.float64).reshape(-1) * base
        eps = 10.0 * np.abs(want).max()
        p1, hist = fit.run(p0, steps=1, lr=lr, groups=(n,), eps=float(eps))
        assert abs(hist[0] - loss) <= 1e-5 * loss, (n, hist[0], loss)
        th = np.where(base != 0, getattr(p1, n).astype(np.float64).reshape(-1) / np.where(base != 0, base, 1) - 1, 0)
        got = -np.sign(th) * np.abs(th) * eps / (lr - np.abs(th))
        big = np.abs(want) > 1e-2 * np.abs(want).max()
        assert big.sum() > 0, n
        assert np.abs(got - want)[big].max() <= 2e-3 * np.abs(want).max(), (n, np.abs(got - want)[big].max())
    fit.free()
    gs.free()
