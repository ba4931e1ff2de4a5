"""Seeded synthetic scenes in polyred's scene contract.

A scene here is what the Go side would marshal from a polyred scene graph:
meshes, objects (mesh + composed model matrix + material base) in
``scene.IterObjects`` order, the flat material table, the light sources in
``Scene.Lights()`` order, ambient intensities, and ``Scene.Center()``.  The
generators reproduce the workload configurations the renderer is benchmarked
on; every draw comes from ``numpy.random.Generator(PCG64(seed))``.

Topology note: the UV sphere is the standard 32 longitude x 16 latitude
sphere with single-vertex poles -- 482 vertices, 960 triangles (2 x 32 cap
fans + 14 x 64 band triangles).  A 32 x 16 quad grid split into triangles is
1024 triangles, 960 of which are non-degenerate; no closed 32 x 16 topology
has 992.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import abi

F32 = np.float32


def _f32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=F32))


def _i32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def from_value(v: float) -> int:
    """color.FromValue for one channel: uint8(Round(float32(v)*255))
    (color/color.go:33-44)."""
    x = F32(F32(v) * F32(255))
    return int(math.floor(float(x) + 0.5)) if x >= 0 else int(math.ceil(float(x) - 0.5))


@dataclass
class Mesh:
    pos: np.ndarray  # (V,4) f32
    nor: np.ndarray  # (V,4) f32
    col: np.ndarray  # (V,4) u8
    uv: np.ndarray   # (V,2) f32
    tri: np.ndarray  # (T,3) i32, local vertex ids
    mat: np.ndarray  # (T,) i32, geometry-local material id

    def aabb(self):
        used = self.pos[np.unique(self.tri.reshape(-1))][:, :3]
        return used.min(axis=0).astype(F32), used.max(axis=0).astype(F32)


@dataclass
class Material:
    diffuse: tuple = (128, 128, 128, 255)
    specular: tuple = (128, 128, 128, 255)
    texture: tuple = (255, 255, 255, 255)
    shininess: float = 1.0
    flat_shading: bool = False
    ambient_occlusion: bool = False
    receive_shadow: bool = False
    texture_id: int = -1        # index into Scene.textures; -1: the uniform `texture` colour


@dataclass
class Texture:
    """buffer.Texture: an (h, w, 4) uint8 image, row 0 first, and whether
    Query uses its mip levels (TextureIsoMipmap; NewTexture's default)."""
    rgba: np.ndarray
    use_mipmap: bool = True


@dataclass
class Light:
    kind: int = abi.PRH_LIGHT_POINT
    position: tuple = (1.0, 1.0, 1.0)
    direction: tuple = (0.0, -1.0, 0.0)
    intensity: float = 1.0
    color: tuple = (255, 255, 255, 255)
    cast_shadow: bool = False


def unit3(v) -> tuple:
    """math.Vec3.Unit in float32 (math/vec3.go:89-93), as NewDirectional
    normalises its direction (light/directional.go:45)."""
    x, y, z = (F32(c) for c in v)
    # a float32*float32 product is exact in float64, so a*b+c in float64 is
    # the double FMA of math.FMA; the outer float32() is the Go conversion.
    inner = F32(np.float64(y) * np.float64(y) + np.float64(F32(z * z)))
    d = F32(np.float64(x) * np.float64(x) + np.float64(inner))
    n = F32(F32(1.0) / F32(math.sqrt(float(d))))
    return (float(F32(x * n)), float(F32(y * n)), float(F32(z * n)))


@dataclass
class Scene:
    meshes: List[Mesh]
    obj_mesh: np.ndarray
    obj_model: np.ndarray      # (O,16) f32
    obj_mat_base: np.ndarray
    materials: List[Material]
    lights: List[Light]
    ambient: List[float]
    camera: dict
    width: int
    height: int
    shadow_map: bool = False
    gamma: bool = True
    shadow_size: int = 0
    background: tuple = (0, 0, 0, 0)
    msaa: int = 1
    textures: list = field(default_factory=list)  # Texture images the materials name
    name: str = ""
    _keep: list = field(default_factory=list, repr=False)

    # ----------------------------------------------------------------- sizes
    @property
    def n_triangles(self) -> int:
        per_mesh = [m.tri.shape[0] for m in self.meshes]
        return int(sum(per_mesh[m] for m in self.obj_mesh))

    @property
    def n_pixels(self) -> int:
        return self.width * self.height

    # -------------------------------------------------------- scene centre
    def center(self) -> np.ndarray:
        """Scene.Center() (scene/scene.go:49-52): the union of the objects'
        untransformed AABBs (geometry: mesh AABB; point light: its position;
        directional and ambient lights: the origin), then (min+max)*0.5."""
        mins, maxs = [], []
        for m in self.obj_mesh:
            a, b = self.meshes[m].aabb()
            mins.append(a)
            maxs.append(b)
        for l in self.lights:
            p = np.array(l.position if l.kind == abi.PRH_LIGHT_POINT else (0, 0, 0), dtype=F32)
            mins.append(p)
            maxs.append(p)
        if self.ambient:
            mins.append(np.zeros(3, F32))
            maxs.append(np.zeros(3, F32))
        mn = np.min(np.stack(mins), axis=0).astype(F32)
        mx = np.max(np.stack(maxs), axis=0).astype(F32)
        return ((mn + mx).astype(F32) * F32(0.5)).astype(F32)

    # ------------------------------------------------------------ C structs
    def desc(self) -> abi.prh_scene_desc:
        keep = []

        def ptr(a, ct):
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ct))

        vtx_off = [0]
        tri_off = [0]
        for m in self.meshes:
            vtx_off.append(vtx_off[-1] + m.pos.shape[0])
            tri_off.append(tri_off[-1] + m.tri.shape[0])
        vpos = _f32(np.concatenate([m.pos for m in self.meshes]))
        vnor = _f32(np.concatenate([m.nor for m in self.meshes]))
        vcol = np.ascontiguousarray(np.concatenate([m.col for m in self.meshes]).astype(np.uint8))
        vuv = _f32(np.concatenate([m.uv for m in self.meshes]))
        tri_v = _i32(np.concatenate([m.tri for m in self.meshes]))
        tri_mat = _i32(np.concatenate([m.mat for m in self.meshes]))
        mats = (abi.prh_material * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i].diffuse[:] = m.diffuse
            mats[i].specular[:] = m.specular
            mats[i].texture[:] = m.texture
            mats[i].shininess = m.shininess
            mats[i].flat_shading = int(m.flat_shading)
            mats[i].ambient_occlusion = int(m.ambient_occlusion)
            mats[i].receive_shadow = int(m.receive_shadow)
            mats[i].texture_id = int(m.texture_id)
        texs = (abi.prh_texture * max(1, len(self.textures)))()
        for i, t in enumerate(self.textures):
            px = np.ascontiguousarray(t.rgba, np.uint8)
            texs[i].height, texs[i].width = px.shape[0], px.shape[1]
            texs[i].rgba = ptr(px, C.c_uint8)
            texs[i].use_mipmap = int(t.use_mipmap)
        lights = (abi.prh_light * max(1, len(self.lights)))()
        for i, l in enumerate(self.lights):
            lights[i].kind = l.kind
            lights[i].position[:] = l.position
            lights[i].direction[:] = l.direction
            lights[i].intensity = l.intensity
            lights[i].color[:] = l.color
            lights[i].cast_shadow = int(l.cast_shadow)
        amb = _f32(self.ambient if self.ambient else [0.0])
        keep += [mats, lights, texs]
        d = abi.prh_scene_desc()
        d.n_meshes = len(self.meshes)
        d.mesh_vtx_off = ptr(_i32(vtx_off), C.c_int32)
        d.mesh_tri_off = ptr(_i32(tri_off), C.c_int32)
        d.vpos = ptr(vpos, C.c_float)
        d.vnor = ptr(vnor, C.c_float)
        d.vcol = ptr(vcol, C.c_uint8)
        d.vuv = ptr(vuv, C.c_float)
        d.tri_v = ptr(tri_v, C.c_int32)
        d.tri_mat = ptr(tri_mat, C.c_int32)
        d.n_objects = len(self.obj_mesh)
        d.obj_mesh = ptr(_i32(self.obj_mesh), C.c_int32)
        d.obj_model = ptr(_f32(self.obj_model), C.c_float)
        d.obj_mat_base = ptr(_i32(self.obj_mat_base), C.c_int32)
        d.n_materials = len(self.materials)
        d.materials = mats
        d.n_lights = len(self.lights)
        d.lights = lights
        d.n_ambient = len(self.ambient)
        d.ambient_intensity = ptr(amb, C.c_float)
        d.center[:] = [float(c) for c in self.center()]
        d.n_textures = len(self.textures)
        d.textures = texs
        self._keep = keep
        return d

    def camera_struct(self) -> abi.prh_camera:
        c = abi.prh_camera()
        cam = self.camera
        c.kind = cam.get("kind", abi.PRH_CAMERA_PERSPECTIVE)
        c.position[:] = cam["position"]
        c.target[:] = cam.get("target", (0, 0, 0))
        c.up[:] = cam.get("up", (0, 1, 0))
        c.fov = cam.get("fov", 60.0)
        c.aspect = cam.get("aspect", float(F32(self.width) / F32(self.height)))
        c.near_ = cam.get("near", 0.01)
        c.far_ = cam.get("far", 1000.0)
        c.left, c.right, c.bottom, c.top = cam.get("frustum", (-1.0, 1.0, -1.0, 1.0))
        return c

    def options_struct(self, keep_linear: bool = False) -> abi.prh_options:
        o = abi.prh_options()
        o.width, o.height, o.msaa = self.width, self.height, self.msaa
        o.background[:] = self.background
        o.shadow_map = int(self.shadow_map)
        o.gamma_correct = int(self.gamma)
        o.shadow_size = int(self.shadow_size)
        o.keep_linear = int(keep_linear)
        return o


# ------------------------------------------------------------------ meshes
def _orient(pos: np.ndarray, tri: np.ndarray, outward) -> np.ndarray:
    """Order each triangle so that (b-a)x(c-a) points along `outward`, which
    renders front-facing under the reference's cull rule (render/cull.go:26-28)."""
    p = pos[:, :3].astype(np.float64)
    a, b, c = p[tri[:, 0]], p[tri[:, 1]], p[tri[:, 2]]
    n = np.cross(b - a, c - a)
    o = outward(a, b, c)
    flip = np.einsum("ij,ij->i", n, o) < 0
    tri = tri.copy()
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    return tri


def uv_sphere(n_lon: int = 32, n_lat: int = 16, color=(255, 255, 255, 255)) -> Mesh:
    """Unit UV sphere with single-vertex poles and smooth (position) normals."""
    verts, uvs = [(0.0, 1.0, 0.0)], [(0.5, 1.0)]
    for i in range(1, n_lat):
        th = math.pi * i / n_lat
        for j in range(n_lon):
            ph = 2 * math.pi * j / n_lon
            verts.append((math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)))
            uvs.append((j / n_lon, 1.0 - i / n_lat))
    verts.append((0.0, -1.0, 0.0))
    uvs.append((0.5, 0.0))
    v = np.array(verts, dtype=np.float64)
    south = len(verts) - 1

    def ring(i, j):
        return 1 + (i - 1) * n_lon + (j % n_lon)

    tris = []
    for j in range(n_lon):
        tris.append((0, ring(1, j), ring(1, j + 1)))
    for i in range(1, n_lat - 1):
        for j in range(n_lon):
            a, b, c, d = ring(i, j), ring(i + 1, j), ring(i + 1, j + 1), ring(i, j + 1)
            tris.append((a, b, c))
            tris.append((a, c, d))
    for j in range(n_lon):
        tris.append((south, ring(n_lat - 1, j + 1), ring(n_lat - 1, j)))
    tri = np.array(tris, dtype=np.int32)
    pos = np.concatenate([v, np.ones((len(v), 1))], axis=1).astype(F32)
    nrm = v / np.linalg.norm(v, axis=1, keepdims=True)
    nor = np.concatenate([nrm, np.zeros((len(v), 1))], axis=1).astype(F32)
    tri = _orient(pos, tri, lambda a, b, c: (a + b + c) / 3.0)
    col = np.tile(np.array(color, np.uint8), (len(v), 1))
    uv = np.array(uvs, F32)  # longitude, latitude (the seam wraps through Query)
    return Mesh(pos, nor, col, uv, tri, np.zeros(len(tri), np.int32))


def grid_plane(n: int, half: float, y: float, color=(255, 255, 255, 255), material: int = 0,
               zrange=None) -> Mesh:
    """n x n quad grid over x in [-half, half], z in zrange (default the same
    as x) at height y, facing +y."""
    xs = np.linspace(-half, half, n + 1)
    zs = xs if zrange is None else np.linspace(zrange[0], zrange[1], n + 1)
    X, Z = np.meshgrid(xs, zs, indexing="xy")
    V = (n + 1) * (n + 1)
    pos = np.stack([X.ravel(), np.full(V, y), Z.ravel(), np.ones(V)], axis=1).astype(F32)
    nor = np.tile(np.array([0, 1, 0, 0], F32), (V, 1))
    tris = []
    for r in range(n):
        for cc in range(n):
            a = r * (n + 1) + cc
            b, c, d = a + 1, a + n + 2, a + n + 1
            tris.append((a, b, c))
            tris.append((a, c, d))
    tri = np.array(tris, np.int32)
    tri = _orient(pos, tri, lambda a, b, c: np.tile([0.0, 1.0, 0.0], (len(a), 1)))
    col = np.tile(np.array(color, np.uint8), (V, 1))
    uv = np.stack([(X.ravel() - xs[0]) / (xs[-1] - xs[0]), (Z.ravel() - zs[0]) / (zs[-1] - zs[0])],
                  axis=1).astype(F32)
    return Mesh(pos, nor, col, uv, tri, np.full(len(tri), material, np.int32))


def model_matrix(scale: float, rot: np.ndarray, trans) -> np.ndarray:
    """Translate . Scale . Rotate as TransformContext composes it
    (math/context.go:36-105): the rotation block is float32(s*R), the last
    column the translation."""
    m = np.zeros((4, 4), F32)
    m[:3, :3] = (F32(scale) * rot.astype(F32)).astype(F32)
    m[:3, 3] = np.asarray(trans, F32)
    m[3, 3] = 1
    return m.reshape(16)


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _mat_from(kd, ks, shin, ao=False, shadow=False) -> Material:
    return Material(diffuse=(from_value(kd[0]), from_value(kd[1]), from_value(kd[2]), 255),
                    specular=(from_value(ks), from_value(ks), from_value(ks), 255),
                    shininess=float(F32(shin)), ambient_occlusion=ao, receive_shadow=shadow)


# ----------------------------------------------------------------- configs
def c1_plain(seed: int = 1, width: int = 256, height: int = 256) -> Scene:
    """Plain forward, CPU-runnable: one UV sphere on a 2-triangle ground
    plane, 256x256, fov 45, near 0.1, far 100, eye (0,1.5,4) -> origin, one
    point light at (2,3,2) intensity 1 plus ambient 0.1, Blinn-Phong with kd
    uniform in [0.2,0.9]^3, ks 0.5, shininess 32, no shadows, no AO, sRGB."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kd = rng.uniform(0.2, 0.9, 3)
    sphere = uv_sphere()
    # the ground stays in front of the eye (z < 4): the reference clips in
    # screen space after the perspective divide (render/clipping.go), so
    # geometry behind the camera is not representable
    ground = grid_plane(1, 5.0, -1.0, zrange=(-6.0, 2.5))
    ident = np.eye(4, dtype=F32).reshape(16)
    mats = [_mat_from(kd, 0.5, 32), _mat_from((0.6, 0.6, 0.6), 0.2, 8)]
    return Scene(meshes=[sphere, ground], obj_mesh=np.array([0, 1]), obj_model=np.stack([ident, ident]),
                 obj_mat_base=np.array([0, 1]), materials=mats,
                 lights=[Light(position=(2.0, 3.0, 2.0), intensity=1.0)], ambient=[0.1],
                 camera=dict(position=(0.0, 1.5, 4.0), fov=45.0, near=0.1, far=100.0),
                 width=width, height=height, gamma=True, name="c1_plain_256")


def c2_scale(seed: int = 2, width: int = 1920, height: int = 1080, n_spheres: int = 1000, grid: int = 200,
             shadow_size: int = 2048, ao: bool = True, shadows: bool = True) -> Scene:
    """Plain forward at scale: 1000 instanced UV spheres (scale log-uniform in
    [0.05,0.5], uniform in a 20^3 box, random rotations) on a 200x200-quad
    ground grid, 8 materials (kd U[0.1,0.9]^3, ks U[0,1], shininess
    log-U[4,128]), 4 point lights in the box's upper half + 1 directional
    light, all casting 2048^2 shadow maps, AO on, sRGB.  Camera outside the
    box, framing it: eye (0,4,21) -> (0,-4,0), fov 45, near 0.1, far 100."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mats = []
    for _ in range(8):
        kd = rng.uniform(0.1, 0.9, 3)
        ks = rng.uniform(0.0, 1.0)
        shin = math.exp(rng.uniform(math.log(4), math.log(128)))
        mats.append(_mat_from(kd, ks, shin, ao=ao, shadow=shadows))
    sphere = uv_sphere()
    ground = grid_plane(grid, 10.0, -10.5)
    models, bases = [], []
    for _ in range(n_spheres):
        s = math.exp(rng.uniform(math.log(0.05), math.log(0.5)))
        t = rng.uniform(-10, 10, 3)
        models.append(model_matrix(s, random_rotation(rng), t))
        bases.append(int(rng.integers(0, 8)))
    models.append(np.eye(4, dtype=F32).reshape(16))
    bases.append(0)
    lights = []
    for _ in range(4):
        p = (rng.uniform(-10, 10), rng.uniform(0, 10), rng.uniform(-10, 10))
        c = tuple(int(x) for x in rng.integers(200, 256, 3)) + (255,)
        lights.append(Light(kind=abi.PRH_LIGHT_POINT, position=tuple(float(F32(x)) for x in p), intensity=3.0,
                            color=c, cast_shadow=shadows))
    d = unit3((-0.3, -1.0, -0.2))
    lights.append(Light(kind=abi.PRH_LIGHT_DIRECTIONAL, direction=d,
                        position=(float(F32(-d[0] * 30)), float(F32(-d[1] * 30)), float(F32(-d[2] * 30))),
                        intensity=0.5, color=(255, 250, 240, 255), cast_shadow=shadows))
    return Scene(meshes=[sphere, ground], obj_mesh=np.array([0] * n_spheres + [1]), obj_model=np.stack(models),
                 obj_mat_base=np.array(bases), materials=mats, lights=lights, ambient=[0.1],
                 camera=dict(position=(0.0, 4.0, 21.0), target=(0.0, -4.0, 0.0), fov=45.0, near=0.1, far=100.0),
                 width=width, height=height, shadow_map=shadows, gamma=True, shadow_size=shadow_size,
                 name=f"c2_scale_{width}x{height}_{n_spheres}sph")


def c4_adjoint(seed: int = 4, width: int = 512, height: int = 512, n_spheres: int = 104) -> Scene:
    """Forward + adjoint: a 512x512 view of ~100k triangles (104 instanced
    UV spheres of 960 triangles, scale log-uniform in [0.3,1.5] in a 12^3
    box), 8 materials as the 1080p workload's, 4 point lights, no shadows or
    AO.  Every sphere vertex has its own colour (the albedo parameters start
    from them)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mats = []
    for _ in range(8):
        kd = rng.uniform(0.1, 0.9, 3)
        ks = rng.uniform(0.0, 1.0)
        shin = math.exp(rng.uniform(math.log(4), math.log(128)))
        mats.append(_mat_from(kd, ks, shin, ao=False, shadow=False))
    sphere = uv_sphere()
    sphere.col[:, :3] = rng.integers(64, 256, (sphere.col.shape[0], 3)).astype(np.uint8)
    models, bases = [], []
    for _ in range(n_spheres):
        s = math.exp(rng.uniform(math.log(0.3), math.log(1.5)))
        t = rng.uniform(-6, 6, 3)
        models.append(model_matrix(s, random_rotation(rng), t))
        bases.append(int(rng.integers(0, 8)))
    lights = []
    for _ in range(4):
        p = (rng.uniform(-8, 8), rng.uniform(2, 10), rng.uniform(-8, 8))
        c = tuple(int(x) for x in rng.integers(200, 256, 3)) + (255,)
        lights.append(Light(kind=abi.PRH_LIGHT_POINT, position=tuple(float(F32(x)) for x in p), intensity=4.0,
                            color=c))
    return Scene(meshes=[sphere], obj_mesh=np.zeros(n_spheres, np.int64), obj_model=np.stack(models),
                 obj_mat_base=np.array(bases), materials=mats, lights=lights, ambient=[0.1],
                 camera=dict(position=(0.0, 3.0, 18.0), target=(0.0, 0.0, 0.0), fov=45.0, near=0.1, far=100.0),
                 width=width, height=height, shadow_map=False, gamma=True, shadow_size=0,
                 name=f"c4_adjoint_{width}x{height}_{n_spheres}sph")


def c3_views(seed: int = 3, n_views: int = 64, min_tris: int = 10_000, max_tris: int = 500_000,
             min_px: int = 256 * 256, max_px: int = 1920 * 1080, shadow_size: int = 1024) -> List[Scene]:
    """Batched forward: n_views independent views, each its own seeded
    instanced-sphere scene.  Per view: pixel count log-uniform in
    [256^2, 1920x1080] with aspect uniform in [0.75, 2]; triangles
    log-uniform in [10k, 500k] (spheres of 960 triangles, scale log-uniform
    in [0.1, 0.8] in a 16^3 box); 8 materials; 1-8 shadow-casting lights,
    each point or directional, 1024^2 shadow maps; no AO; sRGB."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sphere = uv_sphere()
    views = []
    for v in range(n_views):
        area = math.exp(rng.uniform(math.log(min_px), math.log(max_px)))
        aspect = rng.uniform(0.75, 2.0)
        w = max(16, int(round(math.sqrt(area * aspect))))
        h = max(16, int(round(area / w)))
        n_tris = math.exp(rng.uniform(math.log(min_tris), math.log(max_tris)))
        n_sph = max(1, int(round(n_tris / sphere.tri.shape[0])))
        mats = []
        for _ in range(8):
            kd = rng.uniform(0.1, 0.9, 3)
            mats.append(_mat_from(kd, rng.uniform(0.0, 1.0), math.exp(rng.uniform(math.log(4), math.log(128))),
                                  shadow=True))
        models, bases = [], []
        for _ in range(n_sph):
            s = math.exp(rng.uniform(math.log(0.1), math.log(0.8)))
            models.append(model_matrix(s, random_rotation(rng), rng.uniform(-8, 8, 3)))
            bases.append(int(rng.integers(0, 8)))
        lights = []
        for _ in range(int(rng.integers(1, 9))):
            c = tuple(int(x) for x in rng.integers(200, 256, 3)) + (255,)
            if rng.uniform() < 0.75:
                p = (rng.uniform(-8, 8), rng.uniform(0, 10), rng.uniform(-8, 8))
                lights.append(Light(kind=abi.PRH_LIGHT_POINT, position=tuple(float(F32(x)) for x in p),
                                    intensity=3.0, color=c, cast_shadow=True))
            else:
                d = unit3((rng.uniform(-0.5, 0.5), -1.0, rng.uniform(-0.5, 0.5)))
                lights.append(Light(kind=abi.PRH_LIGHT_DIRECTIONAL, direction=d,
                                    position=tuple(float(F32(-x * 30)) for x in d), intensity=0.6, color=c,
                                    cast_shadow=True))
        views.append(Scene(meshes=[sphere], obj_mesh=np.zeros(n_sph, np.int64), obj_model=np.stack(models),
                           obj_mat_base=np.array(bases), materials=mats, lights=lights, ambient=[0.1],
                           camera=dict(position=(0.0, 4.0, 22.0), target=(0.0, 0.0, 0.0), fov=45.0, near=0.1,
                                       far=100.0),
                           width=w, height=h, shadow_map=True, gamma=True, shadow_size=shadow_size,
                           name=f"c3_view{v}_{w}x{h}_{n_sph}sph_{len(lights)}l"))
    return views


def c5_inverse(seed: int = 5, n_views: int = 32, n_spheres: int = 260, min_side: int = 256, max_side: int = 1024,
               shadow_size: int = 1024) -> tuple:
    """Batched inverse rendering: one scene seen from n_views cameras.  The
    scene is n_spheres instanced UV spheres (960 triangles each; scale
    log-uniform in [0.4, 1.6] in a 14^3 box) on a 20x20-quad ground, 250k
    triangles at the defaults, 8 materials (kd U[0.1,0.9]^3, ks U[0.1,1]
    grey, shininess log-U[4,64]) that receive shadows, 4 shadow-casting
    point lights in the box's upper half with 1024^2 maps, ambient 0.1.
    Views: eye on a ring of radius U[20,26] at height U[3,12] around the
    box, looking at its centre, fov 45; pixel count log-uniform in
    [256^2, 1024^2] with aspect U[0.75, 1.5].  Returns (scene, views), views
    a list of (camera dict, width, height)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mats = []
    for _ in range(8):
        kd = rng.uniform(0.1, 0.9, 3)
        mats.append(_mat_from(kd, rng.uniform(0.1, 1.0), math.exp(rng.uniform(math.log(4), math.log(64))),
                              shadow=True))
    sphere = uv_sphere()
    ground = grid_plane(20, 10.0, -8.0)
    models, bases = [], []
    for _ in range(n_spheres):
        s = math.exp(rng.uniform(math.log(0.4), math.log(1.6)))
        models.append(model_matrix(s, random_rotation(rng), rng.uniform(-7, 7, 3)))
        bases.append(int(rng.integers(0, 8)))
    models.append(np.eye(4, dtype=F32).reshape(16))
    bases.append(0)
    lights = []
    for _ in range(4):
        p = (rng.uniform(-7, 7), rng.uniform(4, 12), rng.uniform(-7, 7))
        c = tuple(int(x) for x in rng.integers(200, 256, 3)) + (255,)
        lights.append(Light(kind=abi.PRH_LIGHT_POINT, position=tuple(float(F32(x)) for x in p), intensity=4.0,
                            color=c, cast_shadow=True))
    views = []
    for _ in range(n_views):
        area = math.exp(rng.uniform(math.log(min_side * min_side), math.log(max_side * max_side)))
        aspect = rng.uniform(0.75, 1.5)
        w = max(16, int(round(math.sqrt(area * aspect))))
        h = max(16, int(round(area / w)))
        a, r, y = rng.uniform(0, 2 * math.pi), rng.uniform(20, 26), rng.uniform(3, 12)
        eye = tuple(float(F32(v)) for v in (r * math.cos(a), y, r * math.sin(a)))
        views.append((dict(position=eye, target=(0.0, -1.0, 0.0), fov=45.0, near=0.1, far=100.0), w, h))
    scene = Scene(meshes=[sphere, ground], obj_mesh=np.array([0] * n_spheres + [1]), obj_model=np.stack(models),
                  obj_mat_base=np.array(bases), materials=mats, lights=lights, ambient=[0.1],
                  camera=views[0][0], width=views[0][1], height=views[0][2], shadow_map=True, gamma=True,
                  shadow_size=shadow_size, name=f"c5_inverse_{n_views}views_{n_spheres}sph")
    return scene, views


def tiny(seed: int, width: int = 32, height: int = 24, n_tris: int = 6, vertex_color: bool = True,
         perspective: bool = True) -> Scene:
    """Small random-triangle scenes for brute-force raster checks: triangles
    in front of the camera, some crossing the viewport border (clipping),
    some overlapping (depth ties), mixed material and vertex-colour."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = rng.uniform(-1.5, 1.5, (3 * n_tris, 3))
    pos[:, 2] = rng.uniform(-1.0, 1.0, 3 * n_tris)
    nor = rng.normal(size=(3 * n_tris, 3))
    nor /= np.linalg.norm(nor, axis=1, keepdims=True)
    P = np.concatenate([pos, np.ones((3 * n_tris, 1))], 1).astype(F32)
    N = np.concatenate([nor, np.zeros((3 * n_tris, 1))], 1).astype(F32)
    col = rng.integers(0, 256, (3 * n_tris, 4)).astype(np.uint8)
    tri = np.arange(3 * n_tris, dtype=np.int32).reshape(n_tris, 3)
    tri = _orient(P, tri, lambda a, b, c: np.tile([0.0, 0.0, 1.0], (len(a), 1)))
    mat = np.where(rng.random(n_tris) < (0.5 if vertex_color else 0.0), -1, rng.integers(0, 2, n_tris)).astype(np.int32)
    mesh = Mesh(P, N, col, np.zeros((3 * n_tris, 2), F32), tri, mat)
    mats = [_mat_from(rng.uniform(0.2, 0.9, 3), 0.5, 16), _mat_from(rng.uniform(0.2, 0.9, 3), 0.3, 4)]
    cam = dict(position=(0.0, 0.0, 3.0), fov=60.0, near=0.1, far=100.0)
    if not perspective:
        cam = dict(kind=abi.PRH_CAMERA_ORTHOGRAPHIC, position=(0.0, 0.0, 3.0), near=1.0, far=-10.0,
                   frustum=(-2.0, 2.0, -1.5, 1.5))
    return Scene(meshes=[mesh], obj_mesh=np.array([0]), obj_model=np.eye(4, dtype=F32).reshape(1, 16),
                 obj_mat_base=np.array([0]), materials=mats,
                 lights=[Light(position=(1.0, 2.0, 3.0)),
                         Light(kind=abi.PRH_LIGHT_DIRECTIONAL, direction=unit3((0.2, -1.0, -0.5)), intensity=0.7)],
                 ambient=[0.15], camera=cam, width=width, height=height, gamma=True, name=f"tiny_{seed}")


# ---------------------------------------------------------------- textures
def procedural_texture(width: int, height: int, seed: int = 0, alpha: bool = False) -> np.ndarray:
    """An (height, width, 4) uint8 image with structure at every scale (so
    that each mip level differs): an 8 x 8 checker over two gradients, plus
    per-texel noise; alpha 255 or, with `alpha`, noisy too."""
    rng = np.random.Generator(np.random.PCG64(seed))
    y, x = np.mgrid[0:height, 0:width]
    cx, cy = max(1, width // 8), max(1, height // 8)
    checker = ((x // cx + y // cy) % 2) * 110
    img = np.zeros((height, width, 4), np.int64)
    img[..., 0] = x * 255 // max(1, width - 1) // 2 + checker
    img[..., 1] = y * 255 // max(1, height - 1) // 2 + 110 - checker
    img[..., 2] = rng.integers(0, 256, (height, width))
    img[..., 3] = rng.integers(96, 256, (height, width)) if alpha else 255
    return np.clip(img, 0, 255).astype(np.uint8)


def textured(seed: int = 6, width: int = 160, height: int = 120, perspective: bool = True,
             use_mipmap: bool = True, tex_sizes=((256, 256), (96, 40)), uv_repeat: float = 3.0,
             n_tris: int = 8, msaa: int = 1) -> Scene:
    """Textured materials (buffer.NewTexture images, mip-mapped unless
    `use_mipmap` is off): a ground plane whose u repeats `uv_repeat` times
    (the wrap of Query), a UV sphere (non-power-of-two texture, alpha), and
    random triangles with UVs in [-1.5, 2.5] of which several cross the
    viewport border (the clipped path), next to a uniform-colour material
    and vertex-colour triangles; two lights and two ambient sources."""
    rng = np.random.Generator(np.random.PCG64(seed))
    texs = [Texture(procedural_texture(w, h, seed=seed + i, alpha=(i == 1)), use_mipmap=use_mipmap)
            for i, (w, h) in enumerate(tex_sizes)]
    sphere = uv_sphere()
    ground = grid_plane(2, 5.0, -1.0, zrange=(-6.0, 2.5), material=0)
    # u repeats (Query wraps it); v stays in [0, 1]: Query reads 1 - v, and
    # maps a negative fraction f to 1 - f > 1, outside the image (transparent)
    ground.uv[:, 0] = (ground.uv[:, 0] * F32(uv_repeat)).astype(F32)
    pos = rng.uniform(-2.2, 2.2, (3 * n_tris, 3))
    pos[:, 1] = rng.uniform(-1.0, 2.0, 3 * n_tris)
    pos[:, 2] = rng.uniform(-1.5, 1.0, 3 * n_tris)
    P = np.concatenate([pos, np.ones((3 * n_tris, 1))], 1).astype(F32)
    nor = rng.normal(size=(3 * n_tris, 3))
    nor[:, 2] = np.abs(nor[:, 2]) + 0.5
    nor /= np.linalg.norm(nor, axis=1, keepdims=True)
    N = np.concatenate([nor, np.zeros((3 * n_tris, 1))], 1).astype(F32)
    tri = np.arange(3 * n_tris, dtype=np.int32).reshape(n_tris, 3)
    tri = _orient(P, tri, lambda a, b, c: np.tile([0.0, 0.0, 1.0], (len(a), 1)))
    mat = rng.choice(np.array([-1, 0, 1, 2], np.int32), n_tris)
    mat[:2] = (0, 1)
    col = rng.integers(0, 256, (3 * n_tris, 4)).astype(np.uint8)
    uv = rng.uniform(-1.5, 2.5, (3 * n_tris, 2)).astype(F32)
    loose = Mesh(P, N, col, uv, tri, mat)
    mats = [_mat_from((0.8, 0.8, 0.8), 0.3, 16), _mat_from(rng.uniform(0.3, 0.9, 3), 0.5, 32),
            _mat_from(rng.uniform(0.3, 0.9, 3), 0.2, 8)]
    mats[0].texture_id, mats[1].texture_id = 0, 1
    ident = np.eye(4, dtype=F32).reshape(16)
    cam = dict(position=(0.0, 1.5, 4.0), fov=50.0, near=0.1, far=100.0)
    if not perspective:
        cam = dict(kind=abi.PRH_CAMERA_ORTHOGRAPHIC, position=(0.0, 1.5, 4.0), near=1.0, far=-20.0,
                   frustum=(-2.5, 2.5, -1.9, 1.9))
    return Scene(meshes=[sphere, ground, loose], obj_mesh=np.array([0, 1, 2]),
                 obj_model=np.stack([ident, ident, ident]), obj_mat_base=np.array([1, 0, 0]), materials=mats,
                 lights=[Light(position=(2.0, 3.0, 2.0), intensity=1.0),
                         Light(kind=abi.PRH_LIGHT_DIRECTIONAL, direction=unit3((0.3, -1.0, -0.6)), intensity=0.6)],
                 ambient=[0.1, 0.05], camera=cam, width=width, height=height, gamma=True, msaa=msaa,
                 textures=texs, name=f"textured_{seed}")
