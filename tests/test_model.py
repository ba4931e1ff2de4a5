"""The OBJ/MTL loader (polyred_amd.model: model/load.go, model/obj/obj.go)
and the reference's own textured scene.  CPU only.

The end-to-end pin is render/raster_test.go's TestRender: the reference
renders the textured bunny at 1920x1080 with MSAA 2 and keeps the result as
internal/testdata/render.png.  That image was made on the reference's GPU
path (NewRenderer acquires a device by default, render/raster.go:105), which
its own tests hold to the CPU path within 1 LSB (shader/blinn_cpu.go:20-23)
and whose forward raster differs along triangle edges.  The oracle's frame of
the same scene, loaded through this loader, must match it to +-1 everywhere
except a few pixels on triangle edges.  The data files (bunny.obj, bunny.mtl,
bunny.png) are fixtures under tests/golden/testdata; render.png is read from
the reference's test data when present.
"""
import os

import numpy as np
import pytest

from polyred_amd import model, scenes

from . import cases

F32 = np.float32
BUNNY = os.path.join(cases.TESTDATA, "bunny.obj")
RENDER_PNG = "/root/reference/internal/testdata/render.png"


def _count(path, prefix):
    with open(path) as f:
        return sum(1 for line in f if line.startswith(prefix))


def test_parse_bunny():
    f = model.parse_obj(BUNNY)
    assert len(f.vertices) == 3 * _count(BUNNY, "v ")
    assert len(f.uvs) == 2 * _count(BUNNY, "vt ")
    assert len(f.normals) == 3 * _count(BUNNY, "vn ")
    assert [o.name for o in f.objs] == ["bunny_bunny.002"]
    assert len(f.objs[0].faces) == _count(BUNNY, "f ") == 1316
    m = f.materials["Default_OBJ"]
    assert (m.illum, m.shininess, m.map_kd) == (2, 250.0, "bunny.png")
    assert m.diffuse == (204, 204, 204, 255) and m.specular == (128, 128, 128, 255)  # FromValue(0.8), (0.5)


def test_bunny_geometry_is_the_file():
    """Each triangle owns three vertices taken from the file's v/vt/vn
    lines (read here independently with numpy), white vertex colours."""
    g = model.load_obj(BUNNY)
    (geo,) = g.geometries
    m = geo.mesh
    assert m.tri.shape == (1316, 3) and m.pos.shape == (3 * 1316, 4)
    assert np.array_equal(m.tri.reshape(-1), np.arange(3 * 1316))
    assert np.all(m.col == 255) and np.all(m.mat == 0) and np.all(m.pos[:, 3] == 1) and np.all(m.nor[:, 3] == 0)
    rows = {"v ": [], "vt ": [], "vn ": [], "f ": []}
    with open(BUNNY) as fh:
        for line in fh:
            for k in rows:
                if line.startswith(k):
                    rows[k].append(line[len(k):].split())
    V = np.array(rows["v "], F32)
    T = np.array(rows["vt "], F32)
    N = np.array(rows["vn "], F32)
    idx = np.array([[c.split("/") for c in r] for r in rows["f "]], np.int64).reshape(-1, 3) - 1
    assert np.array_equal(m.pos[:, :3], V[idx[:, 0]])
    assert np.array_equal(m.uv, T[idx[:, 1]])
    assert np.array_equal(m.nor[:, :3], N[idx[:, 2]])
    assert geo.texture_images[0].shape == (1024, 1024, 4)


def test_obj_forms(tmp_path):
    """Relative indices, the `v x y z w` divisor, faces before any object or
    material (material.Default(): Kd 0.7, Ks 0.5, Ns 30, Blue), a usemtl
    name no library defines, and a zero normal (replaced by the face normal,
    into V1 whichever vertex had it: model/load.go:176-184)."""
    (tmp_path / "t.obj").write_text(
        "v 0 0 0\nv 2 0 0 2\nv 0 1 0\nvn 0 0 1\nvn 0 0 0\n"
        "f 1//1 2//1 3//1\n"
        "f -3//-2 -2//-1 -1//-2\n"
        "usemtl nothing\nf 1//1 2//1 3//1\n")
    g = model.load_obj(str(tmp_path / "t.obj"))
    (geo,) = g.geometries
    m = geo.mesh
    assert np.array_equal(m.pos[1], [1, 0, 0, 1])  # 2 / w with w = 2
    assert np.array_equal(m.mat, [0, 0, 1])
    assert len(geo.materials) == 2
    # faces before any usemtl name "polyred_default", which no file defines
    d = geo.materials[0]
    assert (d.diffuse, d.specular, d.shininess, d.texture) == (
        (179, 179, 179, 255), (128, 128, 128, 255), 30.0, (0, 0, 255, 255))
    # usemtl creates the material it names (obj.go:392-413): all zero, Blue
    z = geo.materials[1]
    assert (z.diffuse, z.specular, z.shininess, z.texture) == ((0, 0, 0, 0), (0, 0, 0, 0), 0.0, (0, 0, 255, 255))
    # second triangle: its second vertex had the zero normal; V1 gets the face normal
    assert np.array_equal(m.nor[3], [0, 0, 1, 0]) and np.array_equal(m.nor[4], [0, 0, 0, 0])


def test_quads_and_polygons(tmp_path):
    """QuadMesh splits (1,2,3), (1,3,4) and keeps the material; a mesh mixing
    triangles and quads lists triangles, then quads, then polygons (fans),
    all with material -1; a triangle-only or quad-only mesh drops polygons;
    polygons alone leave no mesh (model/load.go:120-141,
    geometry/primitive/quad.go:55-63, polygon.go:21-25,71-78)."""
    head = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 1.5 0\n"
    (tmp_path / "q.obj").write_text(head + "f 1 2 3 4\nf 1 2 3 4 5\n")
    m = model.load_obj(str(tmp_path / "q.obj")).geometries[0].mesh
    assert np.array_equal(m.tri, [[0, 1, 2], [0, 2, 3]]) and np.array_equal(m.mat, [0, 0])
    assert np.array_equal(m.pos[:, :2], [[0, 0], [1, 0], [1, 1], [0, 1]])
    (tmp_path / "h.obj").write_text(head + "f 1 2 3 4 5\nf 1 2 3 4\nf 2 3 4\n")
    m = model.load_obj(str(tmp_path / "h.obj")).geometries[0].mesh
    assert m.pos.shape[0] == 3 + 4 + 5
    assert np.array_equal(m.pos[:3, :2], [[1, 0], [1, 1], [0, 1]])  # the triangle first
    assert np.array_equal(m.tri, [[0, 1, 2], [3, 4, 5], [3, 5, 6], [7, 8, 9], [7, 9, 10], [7, 10, 11]])
    assert np.all(m.mat == -1)
    (tmp_path / "p.obj").write_text(head + "f 1 2 3 4 5\n")
    with pytest.raises(model.ObjError):
        model.load_obj(str(tmp_path / "p.obj"))


def test_parse_f32_rounds_once():
    assert model.parse_f32("0.1") == F32(0.1) and model.parse_f32("-2.5e-3") == F32(-2.5e-3)
    # 1 + 2^-24 is half way between 1 and the next float32: ties to even
    assert model.parse_f32("1.000000059604644775390625") == F32(1)
    # just above it, though float64 rounds the decimal onto the midpoint
    s = "1.0000000596046448"
    assert float(s) == 1 + 2.0 ** -24 and F32(float(s)) == F32(1)
    assert model.parse_f32(s) == np.nextafter(F32(1), F32(2))


def test_transform_context():
    """Rotate, Scale, Translate then the model matrix internal . R(q), against
    the float64 closed form T . S . R_y(-pi/6)."""
    g = model.Group([])
    a = float(F32(-np.pi / 6))
    g.rotate((0, 1, 0), a)
    g.scale(4, 4, 4)
    g.translate(0.1, 0, -0.2)
    c, s = np.cos(a), np.sin(a)
    want = np.array([[4 * c, 0, 4 * s, 0.1], [0, 4, 0, 0], [-4 * s, 0, 4 * c, -0.2], [0, 0, 0, 1]])
    got = g.model_matrix().astype(np.float64)
    assert np.abs(got - want).max() <= 4e-6
    # two quarter turns about x compose to a half turn
    h = model.Group([])
    h.rotate((1, 0, 0), float(F32(np.pi / 2)))
    h.rotate((1, 0, 0), float(F32(np.pi / 2)))
    assert np.abs(h.model_matrix().astype(np.float64) - np.diag([1, -1, -1, 1])).max() <= 1e-6
    assert np.array_equal(model.Group([]).model_matrix(), np.eye(4, dtype=F32))


def test_srgb_to_linear_table(orc):
    """LoadImage's GammaCorrect table against the oracle's restatement of
    FromsRGB2Linear (the 1025-entry LUT, color/srgb.go:33-84) and against
    the exact sRGB curve to +-1."""
    t = model._srgb_to_linear_u8()
    L = orc.lib()
    for c in range(256):
        v = L.oracle_linear_from_srgb(F32(c) / F32(255))
        assert t[c] == int(F32(F32(F32(v) * F32(255)) + F32(0.5))), c
        x = c / 255
        exact = x / 12.92 if x <= 0.04045 else ((x + 0.055) / 1.055) ** 2.4
        assert abs(int(t[c]) - exact * 255) <= 1, c
    assert t[0] == 0 and t[255] == 255


def test_alpha_is_premultiplied(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    px = np.array([[[200, 100, 50, 255], [200, 100, 50, 0], [200, 100, 50, 128], [255, 255, 255, 1]]], np.uint8)
    Image.fromarray(px, "RGBA").save(tmp_path / "a.png")
    img = model.load_image(str(tmp_path / "a.png"))
    assert np.array_equal(img[0, 0], px[0, 0]) and np.array_equal(img[0, 1], [0, 0, 0, 0])
    want = px[0, 2, :3].astype(float) * 128 / 255
    assert np.all(np.abs(img[0, 2, :3] - want) < 1) and img[0, 2, 3] == 128
    assert np.array_equal(img[0, 3], [1, 1, 1, 1])


def test_reference_render_png(orc):
    """TestRender's frame (render/raster_test.go:55-90) against the
    reference's own render.png."""
    if not os.path.exists(RENDER_PNG):
        pytest.skip("render.png not present")
    Image = pytest.importorskip("PIL.Image")
    ref = np.array(Image.open(RENDER_PNG).convert("RGBA"), np.uint8)
    s = cases.reference_bunny()
    fr = orc.render(s)
    img = fr.image
    assert img.shape == ref.shape == (1080, 1920, 4)
    d = np.abs(img.astype(int) - ref.astype(int)).max(-1)
    assert np.array_equal(img[..., 3], ref[..., 3])
    big = np.argwhere(d > 1)
    assert len(big) <= 40, len(big)
    top = fr.vis_order[::-1]  # the internal frame, rows top first
    for y, x in big:  # each on a triangle edge of the 2x2-supersampled frame
        blk = top[max(0, 2 * y - 2):2 * y + 4, max(0, 2 * x - 2):2 * x + 4]
        assert len(np.unique(blk)) > 1, (y, x)
    assert (d > 0).mean() <= 0.05
    bunny = (top.reshape(1080, 2, 1920, 2) != 0xFFFFFFFF).any((1, 3))
    assert bunny.sum() > 200_000
    near = bunny.copy()  # the resize filter reaches 2 output pixels past the silhouette
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            near |= np.roll(np.roll(bunny, dy, 0), dx, 1)
    assert np.all(d[~near] == 0)  # the background is exact
