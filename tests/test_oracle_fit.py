"""Fixed shadow visibility for the differentiable shading, on the oracle (CPU).

The inverse-rendering path holds each frame's shadow visibility fixed: a
pixel's shading is scaled by 2^-count, count the lights whose shadow test
reported true there (render/shadow.go:230-283), as the reference halves the
shaded colour per such light (render/raster.go:339-347).

  * The oracle's per-pixel count is pinned by the frame itself: at every
    covered pixel (every material receives shadows, no AO) the shadowed
    colour is the unshadowed colour put through the reference's halving
    `count` times (the quantisation of render/raster.go:339-347, whose
    restatement the known-answer tests pin) -- a wrong count changes the
    colour of any pixel that is not black.
  * The adjoint with the scale is pinned by the exact identity
    lin(count) = 2^-count lin(no shadows) and by central differences.
"""
import numpy as np

import oracle
from polyred_amd import diff, scenes


def shadow_view(i=0):
    s, views = scenes.c5_inverse(n_views=2, n_spheres=14, min_side=72, max_side=96, shadow_size=256)
    s.camera, s.width, s.height = views[i]
    return s


def test_shadow_count_is_the_frames_halving(orc):
    s = shadow_view()
    fr = orc.render(s)
    s.shadow_map = False
    fr0 = orc.render(s)
    cov = fr.vis_order != 0xFFFFFFFF
    assert np.array_equal(cov, fr0.vis_order != 0xFFFFFFFF)
    cnt = fr.shadow_count
    assert cnt.shape == cov.shape and cnt.dtype == np.uint8
    assert np.all(cnt[~cov] == 0)
    assert cnt.max() <= len(s.lights)
    assert len(np.unique(cnt[cov])) >= 3  # lit, partly and more shadowed pixels
    L = orc.lib()
    tab = np.array([[L.oracle_shadow_quant(c, k) for k in range(len(s.lights) + 1)] for c in range(256)], np.uint8)
    c0 = fr0.pregamma[..., :3][cov].astype(np.int64)
    k = np.repeat(cnt[cov].astype(np.int64)[:, None], 3, 1)
    assert np.array_equal(fr.pregamma[..., :3][cov], tab[c0, k])
    # the check has teeth: a count off by one changes many colours
    k1 = np.minimum(k + 1, len(s.lights))
    assert (tab[c0, k1] != fr.pregamma[..., :3][cov]).mean() > 0.2


def test_shadow_scale_identity(orc):
    s = shadow_view(1)
    fr = orc.render(s)
    p = diff.scene_params(s).astype(np.float64)
    _, lin0, _ = oracle.diff(s, fr.vis_order, p)
    _, lin1, _ = oracle.diff(s, fr.vis_order, p, shadow_count=fr.shadow_count)
    shaded = ~np.isnan(lin0).any(-1)
    assert np.array_equal(shaded, ~np.isnan(lin1).any(-1))
    w = np.ldexp(1.0, -fr.shadow_count.astype(np.int64))[..., None]
    assert np.array_equal(lin1[shaded], (w * lin0)[shaded])
    assert (fr.shadow_count[shaded] > 0).any()


def test_shadowed_gradients_match_central_differences(orc):
    s = shadow_view()
    fr = orc.render(s)
    p0 = diff.scene_params(s).astype(np.float64)
    rng = np.random.default_rng(8)
    truth = diff.perturb(p0, rng, 0.2)
    truth.albedo = p0.albedo.copy()
    cnt = fr.shadow_count
    _, target, _ = oracle.diff(s, fr.vis_order, truth, shadow_count=cnt)
    target = np.nan_to_num(target)
    loss, _, g = oracle.diff(s, fr.vis_order, p0, target, shadow_count=cnt)
    assert loss > 0

    def L(p):
        return oracle.diff(s, fr.vis_order, p, target, grads=False, shadow_count=cnt)[0]

    checked = 0
    for name in ("kd", "ks", "shininess", "light_pos", "light_int"):
        gg = g[name].reshape(-1)
        flat = getattr(p0, name).reshape(-1)
        scale = np.abs(gg).max()
        assert scale > 0, name
        for i in list(np.argsort(-np.abs(gg))[:4]) + list(rng.integers(0, gg.size, 2)):
            h = 1e-6 * max(1.0, abs(flat[i]))
            pp, pm = p0.copy(), p0.copy()
            getattr(pp, name).reshape(-1)[i] += h
            getattr(pm, name).reshape(-1)[i] -= h
            fd = (L(pp) - L(pm)) / (2 * h)
            assert abs(fd - gg[i]) <= 1e-5 * scale + 1e-4 * abs(gg[i]), (name, i, fd, gg[i])
            checked += 1
    assert checked >= 25
