"""The oracle's textures (buffer.NewTexture, Texture.Query and their use in
FragmentShader and the raster) against polyred's own Query known-answer
tests (tests/golden/reference_kats.json: buffer/texture_test.go:25-155) and
against what the geometry fixes.  CPU only.

The ground.png and pic.jpg cases load their images from the reference's test
data when it is present (skipped otherwise).  The pic.jpg case holds to +-1
per channel: its expected colour comes from Go's image/jpeg decoder, whose
rounding differs from libjpeg's (PIL) by one in places; Query itself is
exact on the ground.png cases.
"""
import os

import numpy as np
import pytest

import oracle
from polyred_amd import abi, scenes

REF = "/root/reference"
F32 = np.float32


def _load(rel):
    path = os.path.join(REF, rel)
    if not os.path.exists(path):
        pytest.skip(f"{rel} not present")
    Image = pytest.importorskip("PIL.Image")
    return np.array(Image.open(path).convert("RGBA"), np.uint8)


def test_query_kats_1x1_and_2x2(orc, kats):
    k = kats["texture_query"]
    white = np.full((1, 1, 4), 255, np.uint8)  # NewTexture() with no image
    for u, v, lod, want in k["default_1x1"]["cases"]:
        assert oracle.texture_query(white, lod, u, v) == tuple(want)
    c = k["checker_2x2"]
    img = np.array(c["rgba"], np.uint8).reshape(c["height"], c["width"], 4)
    for u, v, lod, want in c["cases"]:
        assert oracle.texture_query(img, lod, u, v, c["use_mipmap"]) == tuple(want), (u, v, lod)


def test_query_kats_ground_png(orc, kats):
    g = kats["texture_query"]["ground_png"]
    img = _load(g["file"])
    assert img.shape == (1024, 1024, 4)
    for u, v, lod, want in g["cases"]:
        assert oracle.texture_query(img, lod, u, v, g["use_mipmap"]) == tuple(want), lod


def test_query_kat_pic_jpg(orc, kats):
    g = kats["texture_query"]["pic_jpg"]
    img = _load(g["file"])
    for u, v, lod, want in g["cases"]:
        got = np.array(oracle.texture_query(img, lod, u, v, g["use_mipmap"]), int)
        assert np.abs(got - np.array(want)).max() <= 1, (got, want)


@pytest.mark.parametrize("w,h", [(2, 2), (256, 256), (96, 40), (17, 5), (64, 1)])
def test_mip_levels(orc, w, h):
    """int(Log2(max(w, h))) + 1 levels of sizes (w >> i, h >> i) while both
    are positive; a constant image stays constant at every level (each
    level is a normalised filter of the image)."""
    img = np.empty((h, w, 4), np.uint8)
    img[:] = (37, 200, 91, 180)
    lv = oracle.texture_levels(img)
    assert len(lv) == int(np.floor(np.log2(max(w, h)))) + 1
    for i, a in enumerate(lv):
        if (w >> i) > 0 and (h >> i) > 0:
            assert a.shape[:2] == (h >> i, w >> i), i
        assert a.size == 0 or np.all(a == img[0, 0]), i
    assert len(oracle.texture_levels(img[:1, :1])) == 1


def test_query_wrap_and_outside(orc):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (16, 32, 4)).astype(np.uint8)
    for mip in (True, False):
        for u in (0.25, 0.375, 0.8125):
            for v in (0.125, 0.5):
                base = oracle.texture_query(img, 0.0, u, v, mip)
                for k in (1, 2, 5):  # Modf keeps the fraction of u + k
                    assert oracle.texture_query(img, 0.0, u + k, v + k, mip) == base
        # an integer coordinate other than 0 wraps to 1, not 0
        assert oracle.texture_query(img, 0.0, 2.0, 0.5, mip) == oracle.texture_query(img, 0.0, 1.0, 0.5, mip)
        # a negative fraction f maps to 1 - f > 1: outside the image
        assert oracle.texture_query(img, 0.0, -0.5, 0.5, mip) == (0, 0, 0, 0)
        # Modf(+-Inf) = +-Inf, NaN: a NaN coordinate reads outside as well
        assert oracle.texture_query(img, 0.0, float("inf"), 0.5, mip) == (0, 0, 0, 0)
    # nearest texel without mip-mapping: floor(u * (w - 1)), floor(v * (h - 1))
    for u, v in ((0.1, 0.9), (0.5, 0.5), (0.99, 0.01), (1.0, 1.0)):
        x, y = int(np.floor(F32(u) * F32(31))), int(np.floor(F32(v) * F32(15)))
        assert oracle.texture_query(img, 3.0, u, v, False) == tuple(img[y, x])


# --------------------------------------------------- textures in the frame
def _quad_scene(tex, use_mipmap, perspective, W=64, H=64):
    """The square (-1,-1,0)..(1,1,0), UV (0,0)..(1,1), unlit (no lights: the
    fragment colour is the texture colour), no gamma.  The perspective view
    keeps the square inside the viewport: a clipped piece takes UVs from
    screen-space barycentrics (render/clipping.go:88-91), not the surface."""
    P = np.array([[-1, -1, 0, 1], [1, -1, 0, 1], [-1, 1, 0, 1], [1, 1, 0, 1]], F32)
    N = np.tile(np.array([0, 0, 1, 0], F32), (4, 1))
    uv = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], F32)
    tri = np.array([[0, 1, 3], [0, 3, 2]], np.int32)
    mesh = scenes.Mesh(P, N, np.full((4, 4), 255, np.uint8), uv, tri, np.zeros(2, np.int32))
    mat = scenes.Material(texture_id=0)
    if perspective:
        cam = dict(position=(0.25, -1.5, 2.1), target=(0.0, 0.1, 0.0), up=(0.0, 0.0, 1.0), fov=60.0, near=0.1,
                   far=100.0)
    else:
        cam = dict(kind=abi.PRH_CAMERA_ORTHOGRAPHIC, position=(0.0, 0.0, 3.0), near=1.0, far=-10.0,
                   frustum=(-1.0, 1.0, -1.0, 1.0))
    return scenes.Scene(meshes=[mesh], obj_mesh=np.array([0]), obj_model=np.eye(4, dtype=F32).reshape(1, 16),
                        obj_mat_base=np.array([0]), materials=[mat], lights=[], ambient=[], camera=cam, width=W,
                        height=H, gamma=False, textures=[scenes.Texture(tex, use_mipmap)])


def _index_texture(n=64):
    """Texel (x, y) holds R = 4x, G = 4y: the colour names the texel."""
    y, x = np.mgrid[0:n, 0:n]
    return np.stack([4 * x, 4 * y, np.zeros_like(x), np.full_like(x, 255)], -1).astype(np.uint8)


def _screen_matrix(orc, s):
    """viewport . projection . view (row-major, column vectors), from the
    oracle's matrix entry points, which polyred's own tests pin."""
    import ctypes as C
    cam = s.camera_struct()
    f = lambda a: np.ascontiguousarray(a, F32).ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    out = [np.zeros(16, F32) for _ in range(3)]
    L = orc.lib()
    L.oracle_view_matrix(f(np.array(cam.position[:], F32)), f(np.array(cam.target[:], F32)),
                         f(np.array(cam.up[:], F32)), f(out[0]))
    L.oracle_proj_matrix(C.byref(cam), f(out[1]))
    L.oracle_viewport_matrix(F32(s.width), F32(s.height), f(out[2]))
    V, Pm, VP = (o.reshape(4, 4).astype(np.float64) for o in out)
    return VP @ Pm @ V


def _uv_at_pixels(M, W, H):
    """The UV of the quad's point seen at each pixel centre, in float64: the
    plane homography (s, t, 1) -> screen (x w, y w, w)."""
    B = np.array([[2, 0, -1], [0, 2, -1], [0, 0, 0], [0, 0, 1]], np.float64)  # (s, t, 1) -> world
    A = M[[0, 1, 3]] @ B
    r, c = np.mgrid[0:H, 0:W]
    scr = np.stack([c + 0.5, (H - 1 - r) + 0.5, np.ones_like(c, float)], -1)  # image row 0 is the top
    st = scr @ np.linalg.inv(A).T
    return st[..., 0] / st[..., 2], st[..., 1] / st[..., 2]


@pytest.mark.parametrize("perspective", [False, True])
def test_unlit_uv_is_the_surface_point_under_the_pixel(orc, perspective):
    """Nearest sampling of an index texture: every covered pixel shows the
    texel of the quad point under its centre (Query(u, 1 - v), floor(u *
    (w - 1))), to one texel at texel borders.  With perspective the UV must be
    perspective-correct: screen-linear UV is several texels off."""
    s = _quad_scene(_index_texture(), False, perspective)
    fr = orc.render(s)
    cov = fr.image[..., 3] == 255  # the texture is opaque, the background transparent
    assert cov.sum() > (4000 if not perspective else 1200)
    u, v = _uv_at_pixels(_screen_matrix(orc, s), s.width, s.height)
    ex, ey = np.floor(u * 63), np.floor((1 - v) * 63)
    tx, ty = fr.image[..., 0] // 4, fr.image[..., 1] // 4
    dx, dy = (tx - ex)[cov], (ty - ey)[cov]
    assert np.abs(dx).max() <= 1 and np.abs(dy).max() <= 1
    assert ((dx == 0) & (dy == 0)).mean() >= 0.9
    if perspective:  # the test has teeth: screen-affine UV is far off
        M = _screen_matrix(orc, s)
        corners = np.array([[-1, -1, 0, 1], [1, -1, 0, 1], [-1, 1, 0, 1]], np.float64) @ M.T
        sc = corners[:, :2] / corners[:, 3:]
        T = np.linalg.solve(np.c_[sc, np.ones(3)], np.array([[0, 0], [1, 0], [0, 1]], np.float64))
        r, c = np.mgrid[0:s.width, 0:s.height]
        aff = np.stack([c + 0.5, (s.height - 1 - r) + 0.5, np.ones_like(c, float)], -1) @ T
        assert np.abs(np.floor(aff[..., 0] * 63) - ex)[cov].max() >= 3


def test_level_of_detail_follows_minification(orc):
    """A 1-texel checker: shown 1:1 (lod 0) the pixels keep its extremes;
    256 texels over 64 pixels give lod log2(4) = 2, a filtered level, all
    near grey; without mip-mapping the nearest texel is black or white."""
    n = 256
    y, x = np.mgrid[0:n, 0:n]
    chk = ((x + y) % 2 * 255).astype(np.uint8)
    img = np.stack([chk, chk, chk, np.full_like(chk, 255)], -1)
    fine = orc.render(_quad_scene(img[:64, :64], True, False)).image[..., 0].astype(int)
    assert fine.min() <= 40 and fine.max() >= 215
    coarse = orc.render(_quad_scene(img, True, False)).image[..., 0].astype(int)
    assert coarse.min() >= 100 and coarse.max() <= 155
    near = orc.render(_quad_scene(img, False, False)).image[..., 0]
    assert set(np.unique(near)) <= {0, 255}


def test_constant_texture_is_the_uniform_material(orc):
    """A texture whose every texel is the material colour renders exactly
    as the uniform 1x1 texture of that colour (lighting, ambient sum, alpha)."""
    s = scenes.textured(width=96, height=72)
    plain = scenes.textured(width=96, height=72)
    for sc in (s, plain):
        for m in sc.materials:
            m.texture = (201, 150, 90, 230)
    for t in s.textures:
        t.rgba[:] = (201, 150, 90, 230)
    # UVs whose Query lands inside the image (a negative fraction reads
    # outside it: transparent black, unlike the uniform texture)
    s.meshes[2].uv = np.random.default_rng(1).uniform(0.02, 0.98, s.meshes[2].uv.shape).astype(F32)
    for m in plain.materials:
        m.texture_id = -1
    plain.textures = []
    a, b = orc.render(s), orc.render(plain)
    assert np.array_equal(a.vis_order, b.vis_order)
    assert np.array_equal(a.image, b.image)
    assert np.array_equal(np.nan_to_num(a.linear, nan=-1), np.nan_to_num(b.linear, nan=-1))
