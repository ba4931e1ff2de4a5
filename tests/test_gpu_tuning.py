"""The autotuners (prh_autotune, prh_autotune_batch) and their knobs on the
MI355X: raster workgroups per CU, primitives per raster work item (bin
capacity), the small-triangle thresholds of the two setups and the raster
tile side (8, 16 or 32).  Every knob only changes how
the order-independent merges are scheduled, so every setting must give a
frame bit-identical to the default one (which the parity tests hold to the
oracle): visibility keys, depth, shadow maps and the image."""
import numpy as np
import pytest

from polyred_amd import abi, render, scenes

pytestmark = pytest.mark.gpu


def scene():
    return scenes.c2_scale(width=320, height=180, n_spheres=120, grid=20, shadow_size=512, ao=False)


def frame(r):
    r.render_device()
    return [r.image(), r.read(abi.PRH_BUF_VIS), r.read(abi.PRH_BUF_DEPTH), r.read(abi.PRH_BUF_SHADOW)]


def same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))


def test_knobs_do_not_change_the_frame(device):
    s = scene()
    r = render.renderer_for(s, device)
    t = r.tuning()
    assert (t.raster_wg_per_cu, t.item_prims, t.shadow_direct_max, t.main_direct_max) == (8, 2048, 64, 48)
    assert t.raster_tile == 16
    assert t.overlap_shadows == 0  # 320x180: below 2^19 pixels
    want = frame(r)
    for knobs in (dict(item_prims=64), dict(item_prims=8192, raster_wg_per_cu=1),
                  dict(shadow_direct_max=0, main_direct_max=0), dict(shadow_direct_max=256, main_direct_max=256),
                  dict(raster_wg_per_cu=16, item_prims=256), dict(overlap_shadows=1),
                  dict(overlap_shadows=0, item_prims=512), dict(raster_tile=8), dict(raster_tile=32, item_prims=64),
                  dict(raster_tile=8, shadow_direct_max=0, main_direct_max=0), dict(raster_tile=32, overlap_shadows=1),
                  dict(raster_tile=16)):
        r.set_tuning(**knobs)
        same(frame(r), want)
    for bad in (dict(item_prims=1000), dict(item_prims=32), dict(raster_wg_per_cu=0), dict(main_direct_max=300),
                dict(overlap_shadows=2), dict(raster_tile=4), dict(raster_tile=24), dict(raster_tile=64)):
        with pytest.raises(abi.PrhError) as e:
            r.set_tuning(**bad)
        assert e.value.code == abi.PRH_E_INVALID
    r.free()


def test_autotune_keeps_the_frame(device):
    s = scene()
    r = render.renderer_for(s, device)
    want = frame(r)
    t = r.Autotune(frames=2)
    assert t.item_prims in (256, 512, 1024, 2048, 4096) and t.raster_wg_per_cu in (1, 2, 4, 8)
    assert t.shadow_direct_max in (0, 16, 32, 64, 128, 256) and t.main_direct_max in (0, 24, 48, 96, 192)
    assert t.raster_tile in (8, 16, 32)
    assert t.ms > 0
    got = r.tuning()
    assert (got.item_prims, got.raster_wg_per_cu) == (t.item_prims, t.raster_wg_per_cu)
    same(frame(r), want)
    # a tuned renderer in a wide batch still renders its own frame
    render.render_batch([r])
    same([r.image(), r.read(abi.PRH_BUF_VIS), r.read(abi.PRH_BUF_DEPTH), r.read(abi.PRH_BUF_SHADOW)], want)
    r.free()


def test_autotune_batch_keeps_every_view(device):
    """prh_autotune_batch on a batch with views in both size classes (below
    and from 2^19 pixels): every view's frame stays bit-identical to its own
    prh_render with the default knobs, and each class gets a candidate."""
    views = scenes.c3_views(seed=17, n_views=5, min_tris=5_000, max_tris=40_000, min_px=96 * 96,
                            max_px=1024 * 640, shadow_size=256)
    views.append(scenes.c2_scale(width=1024, height=600, n_spheres=60, grid=20, shadow_size=256))
    rs = [render.renderer_for(s, device) for s in views]
    assert {r.W * r.H >= (1 << 19) for r in rs} == {False, True}
    want = [frame(r) for r in rs]
    small, large = render.autotune_batch(rs, frames=1)
    for t in (small, large):
        assert t.item_prims in (256, 512, 1024, 2048, 4096) and t.raster_wg_per_cu in (2, 4, 8, 16)
        assert t.shadow_direct_max in (0, 16, 32, 64, 128, 256) and t.main_direct_max in (0, 24, 48, 96, 192)
        assert t.raster_tile in (8, 16, 32)
        assert t.ms > 0
    for r, w in zip(rs, want):  # the batch the tuner left
        same([r.image(), r.read(abi.PRH_BUF_VIS), r.read(abi.PRH_BUF_DEPTH), r.read(abi.PRH_BUF_SHADOW)], w)
    render.render_batch(rs, graph=True)
    for r, w in zip(rs, want):
        same([r.image(), r.read(abi.PRH_BUF_VIS), r.read(abi.PRH_BUF_DEPTH), r.read(abi.PRH_BUF_SHADOW)], w)
    for r in rs:
        r.free()
