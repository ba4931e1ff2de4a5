"""libprh on the MI355X against the oracle, through the C ABI.

Bit-exact is the bar for every output that is integer or a float the
reference stores: visibility (coverage + primitive order), fragment depth
bits, shadow-map bits, the pre-Round linear shading sums (stricter than the
1e-4 tolerance the workload allows), the deferred colour and the final sRGB
image (stricter than +-1 LSB).  The full 1080p workload is compared over every
pixel for visibility, depth and shadow maps and on a seeded sample of 4096
pixels for shading.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from polyred_amd import abi, render, scenes

from . import cases, golden_cases

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMPTY = np.uint64(0xFFFFFFFFFFFFFFFF)


def gpu_frame(device, scene, frames=1, **knobs):
    r = render.renderer_for(scene, device, keep_linear=True)
    if knobs:
        r.set_tuning(**knobs)
    for _ in range(frames):
        r.render_device()
    out = dict(image=r.image(), vis=r.read(abi.PRH_BUF_VIS), depth=r.read(abi.PRH_BUF_DEPTH),
               linear=r.read(abi.PRH_BUF_LINEAR), pregamma=r.read(abi.PRH_BUF_PREGAMMA), stats=r.stats())
    if scene.shadow_map:
        out["shadow"] = r.read(abi.PRH_BUF_SHADOW)
    r.free()
    order = (out["vis"] & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    order[out["vis"] == EMPTY] = 0xFFFFFFFF
    out["order"] = order
    return out


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float32:
        return np.array_equal(a.view(np.uint32), b.view(np.uint32)) or bool(
            np.all((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))))
    return np.array_equal(a, b)


def assert_frame_equal(g, fr, mask=None, scene=None):
    n_vis = int((g["order"] != fr.vis_order).sum())
    assert n_vis == 0, f"{n_vis} visibility mismatches"
    assert same_bits(g["depth"], fr.depth), "depth bits differ"
    if scene is not None and scene.shadow_map:
        sm = g["shadow"].reshape(-1)
        assert same_bits(sm, fr.shadow_maps[: sm.size]), "shadow maps differ"
    sel = np.ones(fr.vis_order.shape, bool) if mask is None else mask.reshape(fr.vis_order.shape).astype(bool)
    sel.reshape(-1)[0] = True
    lin_g, lin_o = g["linear"][sel], fr.linear[sel]
    assert same_bits(lin_g, lin_o), f"linear sums differ at {int((~((lin_g == lin_o) | (np.isnan(lin_g) & np.isnan(lin_o)))).any(-1).sum())} px"
    d = np.abs(g["pregamma"][sel].astype(int) - fr.pregamma[sel].astype(int))
    assert d.max() == 0, f"deferred colour differs at {int((d > 0).any(-1).sum())} px (max {d.max()})"
    if g["image"].shape[:2] == sel.shape:
        img_sel = sel[::-1]
        d = np.abs(g["image"][img_sel].astype(int) - fr.image[img_sel].astype(int))
    else:  # MSAA: the image is the resized frame, compared whole
        d = np.abs(g["image"].astype(int) - fr.image.astype(int))
    assert d.max() == 0, f"image differs at {int((d > 0).any(-1).sum())} px"


def canon(a: np.ndarray) -> np.ndarray:
    """float32 bits with every NaN as 0x7FC00000 (NaN sign/payload differ
    between x86 SSE and the GPU and carry no meaning)."""
    b = np.ascontiguousarray(a, np.float32).view(np.uint32).copy()
    b[np.isnan(a)] = 0x7FC00000
    return b


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "oracle_hashes.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", sorted(golden_cases.all_cases()))
def test_golden_cases_bit_exact(device, orc, golden, name):
    scene = golden_cases.all_cases()[name]
    g = gpu_frame(device, scene)
    fr = orc.render(scene)
    assert_frame_equal(g, fr, scene=scene)
    h = golden[name]
    assert digest(g["order"]) == h["vis_order"]
    assert digest(canon(g["depth"])) == h["depth"]
    assert digest(g["image"]) == h["image"]
    if name.startswith("nan_"):
        assert g["stats"].nan_rerun == 1


@pytest.mark.parametrize("seed", range(8))
def test_tiny_random_scenes(device, orc, seed):
    scene = scenes.tiny(1000 + seed, width=53, height=41, n_tris=24, perspective=seed % 3 != 0)
    assert_frame_equal(gpu_frame(device, scene), orc.render(scene), scene=scene)


def test_c2_reduced_full_compare(device, orc):
    """The C2 scene (1000 spheres, 5 shadow-casting lights, AO) at 480x270
    with 512^2 shadow maps: every output of every pixel bit-exact."""
    scene = scenes.c2_scale(width=480, height=270, n_spheres=1000, grid=50, shadow_size=512)
    g = gpu_frame(device, scene)
    assert_frame_equal(g, orc.render(scene), scene=scene)
    assert g["stats"].n_clipped > 0


def test_c2_full_size(device, orc):
    """The full C2 workload (1920x1080, 1.04M triangles, 5 x 2048^2 shadow
    maps, AO): visibility, depth and shadow maps over every pixel/texel,
    shading on a seeded sample of pixels."""
    scene = scenes.c2_scale()
    rng = np.random.default_rng(11)
    mask = np.zeros(scene.width * scene.height, np.uint8)
    mask[rng.choice(mask.size, 4096, replace=False)] = 1
    g = gpu_frame(device, scene)
    fr = orc.render(scene, mask=mask)
    assert_frame_equal(g, fr, mask=mask, scene=scene)


def test_frames_repeat_and_shadow_maps_persist(device):
    """Shadow maps are zeroed once and max-accumulated (render/shadow.go:87-88,
    221-228): a static scene renders the same frame again and again."""
    scene = golden_cases.c1_shadow_ao()
    r = render.renderer_for(scene, device)
    r.render_device()
    a, sa = r.image(), r.read(abi.PRH_BUF_SHADOW).copy()
    for _ in range(3):
        r.render_device()
    assert np.array_equal(a, r.image()) and np.array_equal(sa, r.read(abi.PRH_BUF_SHADOW))
    r.free()


def test_background_when_pixel_zero_uncovered(device, orc):
    s = scenes.c1_plain(width=64, height=64)
    s.background = (10, 20, 30, 40)
    s.camera = dict(position=(0.0, 0.5, 4.0), target=(0.0, 1.5, 0.0), fov=45.0, near=0.1, far=100.0)
    g = gpu_frame(device, s)
    fr = orc.render(s)
    assert fr.vis_order[0, 0] == 0xFFFFFFFF
    assert_frame_equal(g, fr)


@pytest.mark.parametrize("m", [2, 3])
def test_msaa_bit_exact(device, orc, m):
    """MSAA m: every pass at (W*m, H*m), the viewport-box quirk
    (render/raster.go:416-439), then Resize after gamma: all buffers and the
    resized image bit-exact."""
    s = golden_cases.c1_shadow_ao() if m == 2 else scenes.tiny(77, width=45, height=31, n_tris=24)
    s.width, s.height = (96, 72) if m == 2 else (s.width, s.height)
    s.msaa = m
    g = gpu_frame(device, s)
    fr = orc.render(s)
    assert g["image"].shape == (s.height, s.width, 4) and g["order"].shape == (s.height * m, s.width * m)
    assert_frame_equal(g, fr, scene=s)


def test_msaa_out_of_range_is_invalid(device):
    s = scenes.c1_plain(width=32, height=32)
    with pytest.raises(abi.PrhError) as e:
        render.NewRenderer(render.Size(32, 32), render.Camera(s.camera), render.Scene(s), render.MSAA(17),
                           render.HIP(device))
    assert e.value.code == abi.PRH_E_INVALID


def test_bad_scene_is_rejected(device):
    s = scenes.c1_plain(width=32, height=32)
    s.meshes[0].tri = s.meshes[0].tri.copy()
    s.meshes[0].tri[0, 0] = -1  # numpy accepts it, the library must not
    with pytest.raises(abi.PrhError) as e:
        render.GPUScene(device, s)
    assert e.value.code == abi.PRH_E_INVALID


def test_lifetimes_nest():
    """A context outlives its scenes and a scene its renderers: the C ABI
    refuses to close/free a parent with live children."""
    import ctypes as C

    lib = abi.load()
    ctx, sc, rd = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert lib.prh_open(0, C.byref(ctx)) == abi.PRH_OK
    s = scenes.c1_plain(width=32, height=32)
    desc = s.desc()
    assert lib.prh_scene_upload(ctx, C.byref(desc), C.byref(sc)) == abi.PRH_OK
    cam, opts = s.camera_struct(), s.options_struct()
    assert lib.prh_renderer_new(ctx, sc, C.byref(cam), C.byref(opts), C.byref(rd)) == abi.PRH_OK
    assert lib.prh_close(ctx) == abi.PRH_E_INVALID
    assert lib.prh_scene_free(sc) == abi.PRH_E_INVALID
    assert lib.prh_render(rd) == abi.PRH_OK
    assert lib.prh_renderer_free(rd) == abi.PRH_OK
    assert lib.prh_scene_free(sc) == abi.PRH_OK
    assert lib.prh_close(ctx) == abi.PRH_OK


@pytest.mark.parametrize("overlap", [0, 1])
def test_shadow_passes_beside_main_passes(device, orc, overlap):
    """The shadow passes on the second stream and, with AO, the shading split
    around k_ao (its first half beside it): every output bit-exact, at a size
    whose default keeps one stream."""
    s = golden_cases.c1_shadow_ao()
    g = gpu_frame(device, s, frames=2, overlap_shadows=overlap)
    assert_frame_equal(g, orc.render(s), scene=s)
