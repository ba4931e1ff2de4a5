"""The headless present seam (gpu.Surface, gpu/surface.go:20-183) on the
device, through the C ABI, following the reference's own conformance test
(gpu/surface_gl_linux_test.go:21-120): three frames on a two-texture ring,
each presented frame read back; double-buffer rotation; Present without an
acquire is an error; Resize reallocates and resets the ring.  A frame drawn
into a surface texture equals the renderer's own image, bit for bit."""
import numpy as np
import pytest

from polyred_amd import abi, render, scenes

pytestmark = pytest.mark.gpu


def test_headless_present_ring(device):
    W, H = 96, 64
    surf = render.CreateSurface(device, W, H, frames=2)
    assert surf.Size() == (W, H)
    acquired, renderers = [], []
    for frame in range(3):
        s = scenes.c1_plain(seed=frame + 1, width=W, height=H)
        r = render.renderer_for(s, device)
        want = r.Render()  # the renderer's own image
        renderers.append(r)
        acquired.append(surf.AcquireNextTexture())
        surf.Render(r)
        surf.Present()
        assert np.array_equal(surf.Texture(), want), frame
        assert r.passOnGPU("deferred")
    # double buffering: frames 0 and 1 on distinct textures, frame 2 reuses 0
    assert acquired == [0, 1, 0]
    with pytest.raises(abi.PrhError) as e:
        surf.Present()
    assert e.value.code == abi.PRH_E_INVALID
    surf.Resize(16, 16)
    assert surf.Size() == (16, 16)
    assert surf.AcquireNextTexture() == 0
    with pytest.raises(abi.PrhError):  # the renderer is 96x64
        surf.Render(renderers[0])
    surf.Release()
    for r in renderers:
        r.free()


def test_present_image_and_msaa_frame(device):
    rng = np.random.default_rng(3)
    surf = render.CreateSurface(device, 40, 30, frames=3)
    img = rng.integers(0, 256, (30, 40, 4)).astype(np.uint8)
    surf.PresentImage(img)
    assert np.array_equal(surf.Texture(), img)
    # an MSAA renderer: its resized image lands in the texture
    s = scenes.c1_plain(width=40, height=30)
    s.msaa = 2
    r = render.renderer_for(s, device)
    want = r.Render()
    assert surf.AcquireNextTexture() == 1
    surf.Render(r)
    surf.Present()
    assert np.array_equal(surf.Texture(), want)
    with pytest.raises(abi.PrhError) as e:  # nothing acquired
        surf.Render(r)
    assert e.value.code == abi.PRH_E_INVALID
    surf.Release()
    r.free()


def test_bad_surfaces_are_rejected(device):
    for w, h in ((0, 10), (10, -1)):
        with pytest.raises(abi.PrhError) as e:
            render.CreateSurface(device, w, h)
        assert e.value.code == abi.PRH_E_INVALID


def test_pinned_readback(device):
    """prh_host_alloc memory as the destination of prh_read_image."""
    s = scenes.c1_plain(width=72, height=40)
    r = render.renderer_for(s, device)
    want = r.Render()
    host = render.pinned_array((40, 72, 4))
    assert r.image(host) is host and np.array_equal(host, want)
    with pytest.raises(ValueError):
        r.image(np.empty((40, 72, 3), np.uint8))
    r.free()
