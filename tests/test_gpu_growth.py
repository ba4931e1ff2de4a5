"""Work buffers that outgrow their first estimate: the frame reports the
overflow through its counters, the renderer grows the buffers and draws the
frame again (prh_render), and a view that overflows inside a batch is drawn
again on its own, the next batch being recorded with the grown buffers.  The
frames stay bit-exact against the oracle throughout."""
import numpy as np
import pytest

from polyred_amd import abi, render, scenes

from . import cases

pytestmark = pytest.mark.gpu


def order_of(r):
    vis = r.read(abi.PRH_BUF_VIS)
    order = (vis & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    order[vis == np.uint64(0xFFFFFFFFFFFFFFFF)] = 0xFFFFFFFF
    return order


def test_single_render_grows(device, orc):
    s = cases.stacked_layers()
    fr = orc.render(s)
    r = render.renderer_for(s, device)
    img = r.Render()
    st = r.stats()
    n_tris, n_tiles = 2 * 100, ((160 + 15) // 16) * ((120 + 15) // 16)
    assert st.n_bin_pairs > 4 * n_tris + 4 * n_tiles + 1024  # past the first estimate
    assert np.array_equal(order_of(r), fr.vis_order)
    assert np.array_equal(img, fr.image)
    assert np.array_equal(r.read(abi.PRH_BUF_SHADOW)[0], fr.shadow_maps.reshape(-1, 120, 160)[0])
    img2 = r.Render()  # the grown buffers hold the next frame directly
    assert np.array_equal(img2, fr.image)
    r.free()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_view_grows(device, orc, graph):
    big = cases.stacked_layers(n_layers=80)
    views = [scenes.c1_plain(width=64, height=48), big, scenes.c1_plain(seed=3, width=80, height=60)]
    want = [orc.render(s) for s in views]
    rs = [render.renderer_for(s, device) for s in views]
    for k in range(3):  # the first batch overflows in view 1; later ones fit
        render.render_batch(rs, graph=graph)
        for r, fr in zip(rs, want):
            assert np.array_equal(order_of(r), fr.vis_order), k
            assert np.array_equal(r.image(), fr.image), k
    for r in rs:
        r.free()
