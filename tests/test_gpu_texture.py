"""Textured materials on the MI355X against the oracle, through the C ABI:
mip chains built on the device (buffer.NewTexture, buffer/texture.go:35-68),
Texture.Query per fragment (buffer/texture.go:83-195) with the level of detail
FragmentShader derives from the fragment's UV derivatives
(shader/blinn_cpu.go:28-35, render/raster.go:511-535).  Bit-exact, as every
other frame output."""
import numpy as np
import pytest

from polyred_amd import abi, render, scenes

from .test_gpu_parity import assert_frame_equal, gpu_frame
from .test_oracle_texture import _index_texture, _quad_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("perspective", [True, False])
@pytest.mark.parametrize("use_mipmap", [True, False])
def test_textured_scene_bit_exact(device, orc, perspective, use_mipmap):
    s = scenes.textured(seed=6, width=200, height=150, perspective=perspective, use_mipmap=use_mipmap)
    g, fr = gpu_frame(device, s), orc.render(s)
    assert_frame_equal(g, fr, scene=s)


def test_textured_msaa_and_shadows(device, orc):
    s = scenes.textured(seed=8, width=96, height=72, msaa=2)
    s.shadow_map, s.shadow_size = True, 256
    s.lights[0].cast_shadow = True
    for m in s.materials:
        m.receive_shadow = True
    g, fr = gpu_frame(device, s), orc.render(s)
    assert_frame_equal(g, fr, scene=s)


@pytest.mark.parametrize("size", [(1, 1), (2, 2), (1000, 3), (3, 700), (333, 77)])
def test_texture_shapes(device, orc, size):
    """1x1 (one level), a 2x2, and strips whose short side's later levels
    have zero rows or columns (their reads are transparent black)."""
    s = scenes.textured(seed=9, width=120, height=90, tex_sizes=(size, (5, 9)))
    g, fr = gpu_frame(device, s), orc.render(s)
    assert_frame_equal(g, fr, scene=s)


@pytest.mark.parametrize("perspective", [False, True])
def test_unlit_index_texture(device, orc, perspective):
    s = _quad_scene(_index_texture(), False, perspective)
    g, fr = gpu_frame(device, s), orc.render(s)
    assert_frame_equal(g, fr, scene=s)


def test_large_textured_scene(device, orc):
    """A 1024x1024 texture on the ground, 640x480, against the oracle over
    every pixel."""
    s = scenes.textured(seed=10, width=640, height=480, tex_sizes=((1024, 1024), (300, 200)))
    g, fr = gpu_frame(device, s), orc.render(s)
    assert_frame_equal(g, fr, scene=s)


def test_bad_textures_are_rejected(device):
    s = scenes.textured(width=32, height=24)
    s.materials[0].texture_id = 5
    with pytest.raises(abi.PrhError) as e:
        render.renderer_for(s, device)
    assert e.value.code == abi.PRH_E_INVALID
    s = scenes.textured(width=32, height=24)
    s.textures[0].rgba = np.zeros((0, 4, 4), np.uint8)
    with pytest.raises(abi.PrhError) as e:
        render.renderer_for(s, device)
    assert e.value.code == abi.PRH_E_INVALID
