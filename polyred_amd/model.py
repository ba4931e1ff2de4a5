"""model.Load for Wavefront OBJ files (model/load.go:31-150, model/obj/obj.go)
feeding the scene contract's SoA arrays directly.

`load_obj(path)` parses the OBJ and its MTL library as the reference's parser
does (obj.go:83-619: relative indices, the `w` divisor, materials created on
first `usemtl`/`newmtl`, faces before any `o`/`g` opening an unnamed object)
and builds what `loadObj` builds: one geometry per OBJ object, whose
triangles each own three fresh vertices (mesh.NewTriangleMesh), with
geometry-local material ids in order of first use, white vertex colours, and
`material.Default()` for faces whose material is unknown.  Blinn-Phong
materials take Kd, Ks and Ns; a `map_Kd` image is loaded with the sRGB ->
linear conversion LoadImage applies under GammaCorrect(true)
(internal/imageutil/image.go:48-106) and mip-mapped (TextureIsoMipmap(true)),
otherwise the material's texture is the uniform colour Blue.

`Group` mirrors scene.Group's TransformContext (math/context.go): Rotate
accumulates a quaternion, Scale and Translate left-multiply the internal
matrix, and the model matrix is internal . R(q), all in float32 with Go's
rounding (Dot and Cross through float32(fma64)).

Quads and polygons become triangles as the reference's meshes split them
(QuadMesh: (1,2,3), (1,3,4); PolygonMesh: a fan), with its quirks: a mesh of
triangles or of quads alone drops the object's polygons, and a mixed mesh's
triangles lose their material (-1: vertex colours).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from fractions import Fraction
from typing import Dict, List, Optional

import numpy as np

from . import scenes

F32 = np.float32
INV_INDEX = 0xFFFFFFFF  # obj.go:75 invINDEX


def parse_f32(s: str) -> F32:
    """strconv.ParseFloat(s, 32): the float32 nearest the decimal, rounded
    once (Python's float() rounds to float64 first; the rare decimal whose
    float64 lands on a float32 midpoint is settled exactly)."""
    d = float(s)
    f = F32(d)
    if not math.isfinite(d) or d == 0.0:
        return f
    bits = np.array([d], np.float64).view(np.uint64)[0]
    if (int(bits) & 0x1FFFFFFF) != 0x10000000:  # not exactly half way between float32s
        return f
    lo = F32(np.nextafter(f, F32(-np.inf) if f > d else F32(np.inf)))
    a, b = sorted((f, lo), key=float)
    exact, mid = Fraction(s.strip()), (Fraction(float(a)) + Fraction(float(b))) / 2
    if exact < mid:
        return a
    if exact > mid:
        return b
    return a if (int(np.array([a], F32).view(np.uint32)[0]) & 1) == 0 else b  # ties to even


def from_value(r, g, b, a=1.0) -> tuple:
    """color.FromValue (color/color.go:33-44)."""
    return tuple(scenes.from_value(float(F32(x))) for x in (r, g, b, a))


# ------------------------------------------------------------------ parser
@dataclass
class ObjMaterial:
    name: str
    illum: int = 0
    shininess: float = 0.0
    diffuse: tuple = (0, 0, 0, 0)
    specular: tuple = (0, 0, 0, 0)
    map_kd: str = ""


@dataclass
class ObjFace:
    vertices: List[int]
    uvs: List[int]
    normals: List[int]
    material: str


@dataclass
class ObjObject:
    name: str
    faces: List[ObjFace] = field(default_factory=list)
    materials: List[str] = field(default_factory=list)


@dataclass
class ObjFile:
    objs: List[ObjObject] = field(default_factory=list)
    matlib: str = ""
    materials: Dict[str, Optional[ObjMaterial]] = field(default_factory=dict)
    vertices: List[float] = field(default_factory=list)
    normals: List[float] = field(default_factory=list)
    uvs: List[float] = field(default_factory=list)
    mtl_dir: str = ""
    warnings: List[str] = field(default_factory=list)


class ObjError(ValueError):
    pass


def _index(tok: str, count: int, what: str, line: int) -> int:
    v = int(tok)
    if not -2**31 <= v < 2**31:
        raise ObjError(f"{what} index out of range in line:{line}")
    if v > 0:
        return v - 1
    if v < 0:
        return count - 1 + v + 1
    raise ObjError(f"face {what} index value equal to 0 in line:{line}")


def parse_obj(path: str) -> ObjFile:
    """obj.Load (model/obj/obj.go:83-150)."""
    f = ObjFile()
    cur_obj: Optional[ObjObject] = None
    cur_mat: Optional[ObjMaterial] = None

    def new_obj(name):
        nonlocal cur_obj
        f.objs.append(ObjObject(name))
        cur_obj = f.objs[-1]

    with open(path, "r", encoding="utf-8", errors="replace") as fh:
        lines = fh.read().split("\n")
    for ln, raw in enumerate(lines, 1):
        fields = raw.strip("\r\n\t ").split()
        if not fields or fields[0].startswith("#"):
            continue
        t, a = fields[0], fields[1:]
        if t == "mtllib":
            if not a:
                raise ObjError(f"Material library (mtllib) with no fields in line:{ln}")
            f.matlib = a[0]
        elif t in ("o", "g"):
            if not a:
                raise ObjError(f"obj line (o) with no fields in line:{ln}")
            new_obj(a[0])
        elif t == "v":
            if len(a) < 3:
                raise ObjError(f"vertex has less than 3 vertices in 'v' line: {a}")
            w = F32(1)
            if len(a) == 4:
                w = F32(1) / parse_f32(a[3])
            for tok in a[:3]:
                x = parse_f32(tok)
                f.vertices.append(x if w == 1 else F32(x * w))
        elif t == "vn":
            if len(a) < 3:
                raise ObjError(f"normal has less than 3 vertices in 'vn' line: {a}")
            f.normals += [parse_f32(tok) for tok in a[:3]]
        elif t == "vt":
            if len(a) < 2:
                raise ObjError("Less than 2 texture coords. in 'vt' line")
            f.uvs += [parse_f32(tok) for tok in a[:2]]
        elif t == "f":
            if cur_obj is None:
                new_obj(f"unnamed{ln}")
            if not cur_obj.materials and cur_mat is not None:
                cur_obj.materials.append(cur_mat.name)
            if len(a) < 3:
                raise ObjError(f"face line with less 3 fields in line:{ln}")
            face = ObjFace([], [], [], cur_mat.name if cur_mat is not None else "polyred_default")
            for tok in a:
                parts = tok.split("/")
                face.vertices.append(_index(parts[0], len(f.vertices) // 3, "vertex", ln))
                if len(parts) > 1 and parts[1]:
                    face.uvs.append(_index(parts[1], len(f.uvs) // 2, "uv", ln))
                else:
                    face.uvs.append(INV_INDEX)
                face.normals.append(_index(parts[2], len(f.normals) // 3, "normal", ln)
                                    if len(parts) >= 3 else INV_INDEX)
            cur_obj.faces.append(face)
        elif t == "usemtl":
            if not a:
                raise ObjError(f"Usemtl with no fields in line:{ln}")
            if cur_obj is None:
                new_obj(f"unnamed{ln}")
            m = f.materials.get(a[0])
            if m is None:
                m = f.materials[a[0]] = ObjMaterial(a[0])
            cur_obj.materials.append(a[0])
            cur_mat = m
        elif t == "s":
            if not a:
                raise ObjError(f"'s' with no fields in line:{ln}")
        else:
            f.warnings.append(f"obj({ln}): unsupported field: {t}")

    if f.matlib:
        f.mtl_dir = os.path.dirname(path)
        _parse_mtl(f, os.path.join(f.mtl_dir, f.matlib))
    return f


def _parse_mtl(f: ObjFile, path: str) -> None:
    """parseMtlLine (obj.go:446-610); the reference panics when the library
    named by mtllib cannot be opened, so does this (OSError)."""
    cur: Optional[ObjMaterial] = None
    with open(path, "r", encoding="utf-8", errors="replace") as fh:
        lines = fh.read().split("\n")
    for ln, raw in enumerate(lines, 1):
        fields = raw.strip("\r\n\t ").split()
        if not fields or fields[0].startswith("#"):
            continue
        t, a = fields[0], fields[1:]
        if t == "newmtl":
            if not a:
                raise ObjError(f"newmtl with no fields in line:{ln}")
            cur = f.materials.get(a[0])
            if cur is None:
                cur = f.materials[a[0]] = ObjMaterial(a[0])
        elif t in ("Ka", "Kd", "Ke", "Ks"):
            if len(a) < 3:
                raise ObjError(f"'{t}' with less than 3 fields in line:{ln}")
            rgb = [parse_f32(x) for x in a[:3]]
            if t == "Kd":
                cur.diffuse = from_value(*rgb)
            elif t == "Ks":
                cur.specular = from_value(*rgb)
        elif t in ("d", "Ni"):
            if not a:
                raise ObjError(f"'{t}' with no fields in line:{ln}")
            parse_f32(a[0])
        elif t == "Ns":
            if not a:
                raise ObjError(f"'Ns' with no fields in line:{ln}")
            cur.shininess = float(parse_f32(a[0]))
        elif t == "illum":
            if not a:
                raise ObjError(f"'illum' with no fields in line:{ln}")
            cur.illum = int(a[0])
        elif t == "map_Kd":
            if not a:
                raise ObjError(f"No fields in line:{ln}")
            cur.map_kd = a[0]
        else:
            f.warnings.append(f"mtl({ln}): field not supported: {t}")


# ------------------------------------------------------------ textures
_SRGB_LUT = None


def _srgb_to_linear_u8() -> np.ndarray:
    """The 256-entry table LoadImage's GammaCorrect(true) applies to R, G, B:
    uint8(FromsRGB2Linear(float32(c)/0xff)*0xff + 0.5) with the 1025-entry
    float64 LUT and its float32 lerp (color/srgb.go:33-84)."""
    global _SRGB_LUT
    if _SRGB_LUT is None:
        lut = np.empty(1025, np.float64)
        for i in range(1024):
            v = i / 1024
            lut[i] = v / 12.92 if v <= 0.04045 else ((v + 0.055) / 1.055) ** 2.4
        lut[1024] = lut[1023]
        out = np.empty(256, np.uint8)
        for c in range(256):
            v = F32(c) / F32(255)
            if v <= 0:
                r = F32(0)
            elif v >= 1:
                r = F32(1)
            else:
                i = F32(v * F32(1024))
                fl = int(i) & 1023
                v0, v1 = F32(lut[fl]), F32(lut[fl + 1])
                i = F32(i - F32(fl))
                r = F32(F32(v0 * F32(F32(1) - i)) + F32(v1 * i))
            out[c] = int(F32(F32(r * F32(255)) + F32(0.5)))
        _SRGB_LUT = out
    return _SRGB_LUT


def load_image(path: str, gamma_correct: bool = False) -> np.ndarray:
    """imageutil.LoadImage: an (h, w, 4) uint8 image.RGBA, optionally with
    sRGB -> linear applied to R, G, B.  Decoding uses PIL, which gives
    straight alpha; image.RGBA is premultiplied, and LoadImage's draw.Draw
    from a decoded NRGBA converts as image/draw's drawNRGBASrc does:
    c' = ((c * a * 0x101) / 0xff) >> 8 (the identity where a = 255)."""
    from PIL import Image
    img = np.array(Image.open(path).convert("RGBA"), np.uint8)
    a = img[..., 3:4].astype(np.uint32) * 0x101
    img[..., :3] = ((img[..., :3].astype(np.uint32) * a // 0xFF) >> 8).astype(np.uint8)
    if gamma_correct:
        img[..., :3] = _srgb_to_linear_u8()[img[..., :3]]
    return img


# --------------------------------------------------------------- groups
def _fma(a, b, c):
    return F32(np.float64(a) * np.float64(b) + np.float64(c))


def _mulm(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Mat4.MulM: plain float32 sums left to right (math/mat4.go:201-231)."""
    out = np.zeros((4, 4), F32)
    for i in range(4):
        for j in range(4):
            s = F32(a[i, 0] * b[0, j])
            for k in range(1, 4):
                s = F32(s + F32(a[i, k] * b[k, j]))
            out[i, j] = s
    return out


@dataclass
class Geometry:
    mesh: scenes.Mesh
    materials: List[scenes.Material]
    texture_images: List[Optional[np.ndarray]]  # per material: its map_Kd image or None


class Group:
    """scene.Group with its TransformContext (math/context.go)."""

    def __init__(self, geometries: List[Geometry]):
        self.geometries = geometries
        self.rotation = (F32(1), F32(0), F32(0), F32(0))
        self.internal = np.eye(4, dtype=F32)

    def rotate(self, axis, angle: float) -> None:
        x, y, z = (F32(v) for v in axis)
        n = F32(F32(1.0) / F32(math.sqrt(float(_fma(x, x, _fma(y, y, F32(z * z)))))))
        ux, uy, uz = F32(x * n), F32(y * n), F32(z * n)
        half = F32(F32(angle) * F32(0.5))
        c, s = F32(math.cos(float(half))), F32(math.sin(float(half)))
        qa, qv = c, (F32(s * ux), F32(s * uy), F32(s * uz))
        pa, pv = self.rotation[0], self.rotation[1:]
        # Quaternion.Mul (math/quaternion.go:19-23)
        dot = _fma(qv[0], pv[0], _fma(qv[1], pv[1], F32(qv[2] * pv[2])))
        aa = F32(F32(qa * pa) - dot)
        cross = (_fma(qv[1], pv[2], -F32(qv[2] * pv[1])), _fma(qv[2], pv[0], -F32(qv[0] * pv[2])),
                 _fma(qv[0], pv[1], -F32(qv[1] * pv[0])))
        vv = tuple(F32(F32(F32(pv[k] * qa) + F32(qv[k] * pa)) + cross[k]) for k in range(3))
        self.rotation = (aa, *vv)

    def scale(self, sx, sy, sz) -> None:
        self.internal = _mulm(np.diag(np.array([sx, sy, sz, 1], F32)), self.internal)

    def translate(self, tx, ty, tz) -> None:
        t = np.eye(4, dtype=F32)
        t[:3, 3] = np.array([tx, ty, tz], F32)
        self.internal = _mulm(t, self.internal)

    def model_matrix(self) -> np.ndarray:
        """internal . ToRoMat(rotation) (math/context.go:36-42,
        math/quaternion.go:25-37)."""
        w, x, y, z = self.rotation
        two = F32(2)

        def p(a, b):
            return F32(F32(two * a) * b)
        r = np.array([
            [F32(F32(F32(1) - p(y, y)) - p(z, z)), F32(p(x, y) - p(z, w)), F32(p(x, z) + p(y, w)), 0],
            [F32(p(x, y) + p(z, w)), F32(F32(F32(1) - p(x, x)) - p(z, z)), F32(p(y, z) - p(x, w)), 0],
            [F32(p(x, z) - p(y, w)), F32(p(y, z) + p(x, w)), F32(F32(F32(1) - p(x, x)) - p(y, y)), 0],
            [0, 0, 0, 1]], F32)
        return _mulm(self.internal, r)


def _face_normal(p1, p2, p3) -> np.ndarray:
    """Triangle.Normal (geometry/primitive/triangle.go:131-139)."""
    a = (p1 - p2).astype(F32)  # v2v1
    b = (p3 - p2).astype(F32)  # v2v3
    c = np.array([_fma(b[1], a[2], -F32(b[2] * a[1])), _fma(b[2], a[0], -F32(b[0] * a[2])),
                  _fma(b[0], a[1], -F32(b[1] * a[0])), 0], F32)
    ln = F32(math.sqrt(float(_fma(c[0], c[0], _fma(c[1], c[1], _fma(c[2], c[2], F32(c[3] * c[3])))))))
    n = F32(F32(1.0) / ln)
    return (c * n).astype(F32)


def _is_zero(v) -> bool:
    return all(abs(F32(c)) <= F32(1e-7) for c in v[:3])  # ApproxEq(c, 0, Epsilon)


def load_obj(path: str) -> Group:
    """model.MustLoad for an .obj path (model/load.go:23-150)."""
    fi = parse_obj(path)
    mats: Dict[str, tuple] = {}
    for name, m in fi.materials.items():
        if m is None:
            continue
        if m.illum not in (0, 2):
            raise ObjError("unsupported illumination model")
        mat = scenes.Material(diffuse=m.diffuse, specular=m.specular, texture=(0, 0, 255, 255),
                              shininess=float(m.shininess))
        img = None
        if m.map_kd:
            img = load_image(os.path.join(os.path.normpath(fi.mtl_dir), os.path.normpath(m.map_kd)), True)
        mats[name] = (mat, img)
    default = (scenes.Material(diffuse=from_value(0.7, 0.7, 0.7), specular=from_value(0.5, 0.5, 0.5),
                               texture=(0, 0, 255, 255), shininess=30.0), None)
    V = np.array(fi.vertices, F32).reshape(-1, 3)
    N = np.array(fi.normals, F32).reshape(-1, 3)
    UV = np.array(fi.uvs, F32).reshape(-1, 2)
    geoms = []
    for ob in fi.objs:
        local: Dict[int, int] = {}
        gm: List[tuple] = []
        tris, quads, polys = [], [], []
        for face in ob.faces:
            m = mats.get(face.material) or default
            if id(m) not in local:
                local[id(m)] = len(gm)
                gm.append(m)
            k = len(face.vertices)
            (tris if k == 3 else quads if k == 4 else polys).append((face, local[id(m)]))
        # the mesh loadObj makes (model/load.go:120-141): triangles only, quads
        # only (polygons dropped in both), or a PolygonMesh of triangles, then
        # quads, then polygons whose primitives carry material -1
        # (primitive.NewPolygon, geometry/primitive/polygon.go:21-25)
        if tris and not quads:
            faces, hybrid = tris, False
        elif quads and not tris:
            faces, hybrid = quads, False
        elif tris and quads:
            faces, hybrid = tris + quads + polys, True
        else:
            raise ObjError(f"object {ob.name!r} has no triangle or quad faces: the reference builds no mesh for it")
        pos, nor, uv, tri, mid = [], [], [], [], []
        for face, lm in faces:
            b, k = len(pos), len(face.vertices)
            p = [np.array([*V[i], 1], F32) for i in face.vertices]
            if len(N):
                n = [np.array([*N[i], 0], F32) for i in face.normals]
                if k == 3:  # newTrianglePrimitive replaces a zero normal by the
                    # face normal -- always into V1 (model/load.go:176-184)
                    for j in range(3):
                        if _is_zero(n[j]):
                            n[0] = _face_normal(*p)
            else:
                n = [np.zeros(4, F32)] * k
            pos += p
            nor += n
            uv += [np.array(UV[i], F32) for i in face.uvs] if len(UV) else [np.zeros(2, F32)] * k
            # Quad.Triangles is (1,2,3), (1,3,4); Polygon.Triangles the fan from
            # vertex 1 (geometry/primitive/quad.go:55-63, polygon.go:71-78)
            for j in range(k - 2):
                tri.append((b, b + j + 1, b + j + 2))
                mid.append(-1 if hybrid else lm)
        mesh = scenes.Mesh(np.array(pos, F32), np.array(nor, F32), np.full((len(pos), 4), 255, np.uint8),
                           np.array(uv, F32), np.array(tri, np.int32), np.array(mid, np.int32))
        geoms.append(Geometry(mesh, [m for m, _ in gm], [img for _, img in gm]))
    return Group(geoms)


def scene_from_groups(groups: List[Group], lights, ambient, camera: dict, width: int, height: int,
                      **opts) -> scenes.Scene:
    """A scene of loaded groups in IterObjects order: one object per
    geometry with its group's model matrix, the geometries' materials in
    the flat table, and their map_Kd images as textures (an image shared by
    materials is uploaded once)."""
    meshes, models, bases, materials, textures = [], [], [], [], []
    tex_of: Dict[int, int] = {}
    for g in groups:
        mm = g.model_matrix().reshape(16)
        for geo in g.geometries:
            bases.append(len(materials))
            for mat, img in zip(geo.materials, geo.texture_images):
                mat = scenes.Material(**vars(mat))
                if img is not None:
                    if id(img) not in tex_of:
                        tex_of[id(img)] = len(textures)
                        textures.append(scenes.Texture(img, True))
                    mat.texture_id = tex_of[id(img)]
                materials.append(mat)
            meshes.append(geo.mesh)
            models.append(mm)
    return scenes.Scene(meshes=meshes, obj_mesh=np.arange(len(meshes)), obj_model=np.stack(models),
                        obj_mat_base=np.array(bases), materials=materials, lights=list(lights),
                        ambient=list(ambient), camera=camera, width=width, height=height, textures=textures,
                        **opts)
