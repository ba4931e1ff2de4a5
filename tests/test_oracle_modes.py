"""The CPU baseline's thread modes of the oracle (oracle.set_workers): Mode A
(one thread) is the parity order itself; Mode B (several threads, triangles
drawn concurrently under per-pixel locks) gives the same frame wherever no
two triangles tie in depth, which holds for a generic seeded scene."""
import numpy as np

from polyred_amd import scenes


def test_modes_render_the_same_frame(orc):
    s = scenes.c2_scale(width=240, height=136, n_spheres=200, grid=30, shadow_size=256)
    try:
        frames = {}
        for w in (0, 1, 4):
            orc.set_workers(w)
            frames[w] = orc.render(s)
    finally:
        orc.set_workers(0)
    a = frames[0]
    for w in (1, 4):
        b = frames[w]
        assert np.array_equal(a.vis_order, b.vis_order), w
        assert np.array_equal(a.depth.view(np.uint32), b.depth.view(np.uint32)), w
        assert np.array_equal(a.shadow_maps.view(np.uint32), b.shadow_maps.view(np.uint32)), w
        assert np.array_equal(a.image, b.image), w
