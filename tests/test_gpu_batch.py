"""prh_render_batch: many views, each its own scene, in one batched
submission -- one launch per pass over every view (the default), or each
view's launches over several streams -- eager or captured into a HIP graph
and replayed.  Every view must be bit-identical to a separate prh_render of
the same scene (which the other GPU tests hold to the oracle), and a graph
replay to the eager batch, replay after replay with eager work in between."""
import numpy as np
import pytest

from polyred_amd import abi, render, scenes

from . import golden_cases

pytestmark = pytest.mark.gpu


def small_views(n=6, seed=5):
    return scenes.c3_views(seed=seed, n_views=n, min_tris=2_000, max_tris=30_000, min_px=64 * 64,
                           max_px=320 * 240, shadow_size=256)


def outputs(r):
    return r.image(), r.read(abi.PRH_BUF_VIS), r.read(abi.PRH_BUF_SHADOW)


def same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))


@pytest.mark.parametrize("streams", [False, True])
def test_batch_matches_single_renders(device, streams):
    views = small_views()
    single = [render.renderer_for(s, device) for s in views]
    for r in single:
        r.render_device()
    want = [outputs(r) for r in single]
    batch = [render.renderer_for(s, device) for s in views]
    ms = render.render_batch(batch, streams=streams)
    assert ms > 0
    for r, w in zip(batch, want):
        same(outputs(r), w)
    for r in single + batch:
        r.free()


def test_mixed_batch_matches_single_renders(device):
    """Views that skip passes (no shadows), add some (AO, MSAA resize) or
    hold textures, in one wide batch, eager and as a graph."""
    tex = scenes.textured(seed=8, width=96, height=72, msaa=2)
    views = [scenes.c1_plain(width=80, height=64), golden_cases.c1_shadow_ao(), tex] + small_views(n=2, seed=21)
    single = [render.renderer_for(s, device) for s in views]
    for r in single:
        r.render_device()
    want = [(r.image(), r.read(abi.PRH_BUF_VIS)) for r in single]
    batch = [render.renderer_for(s, device) for s in views]
    for graph in (False, True, True):
        render.render_batch(batch, graph=graph)
        for r, w in zip(batch, want):
            same((r.image(), r.read(abi.PRH_BUF_VIS)), w)
    for r in single + batch:
        r.free()


@pytest.mark.parametrize("streams", [False, True])
def test_graph_replays_match_eager(device, streams):
    views = small_views(n=5, seed=9)
    batch = [render.renderer_for(s, device) for s in views]
    render.render_batch(batch, streams=streams)
    want = [outputs(r) for r in batch]
    other = render.renderer_for(views[0], device)
    for k in range(4):
        render.render_batch(batch, graph=True, streams=streams)
        for r, w in zip(batch, want):
            same(outputs(r), w)
        other.render_device()  # eager work between the replays
        same(outputs(other), want[0])
    for r in batch + [other]:
        r.free()


def test_batch_view_matches_oracle(device, orc):
    views = small_views(n=3, seed=13)
    batch = [render.renderer_for(s, device) for s in views]
    render.render_batch(batch)
    fr = orc.render(views[1])
    img, vis, _ = outputs(batch[1])
    order = (vis & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    order[vis == np.uint64(0xFFFFFFFFFFFFFFFF)] = 0xFFFFFFFF
    assert np.array_equal(order, fr.vis_order)
    assert np.array_equal(img, fr.image)
    for r in batch:
        r.free()


def test_batch_rejects_mixed_contexts(device):
    other = render.Device(0)
    s = small_views(n=1)[0]
    a, b = render.renderer_for(s, device), render.renderer_for(s, other)
    with pytest.raises(abi.PrhError) as e:
        render.render_batch([a, b])
    assert e.value.code == abi.PRH_E_INVALID
    a.free(), b.free()
    other.close()
