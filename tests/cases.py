"""Hand-built edge-case scenes shared by the oracle tests and the GPU parity
tests."""
import os

import numpy as np

from polyred_amd import abi, scenes


def _scene(P, tri, mats_local, width, height, camera, materials=None, lights=None, col=None):
    P = np.asarray(P, np.float32)
    n = len(P)
    N = np.tile(np.array([0, 0, 1, 0], np.float32), (n, 1))
    if col is None:
        col = np.tile(np.array([200, 100, 50, 255], np.uint8), (n, 1))
    mesh = scenes.Mesh(P, N, col, np.zeros((n, 2), np.float32), np.asarray(tri, np.int32),
                       np.asarray(mats_local, np.int32))
    return scenes.Scene(meshes=[mesh], obj_mesh=np.array([0]), obj_model=np.eye(4, dtype=np.float32).reshape(1, 16),
                        obj_mat_base=np.array([0]),
                        materials=materials if materials is not None else [scenes._mat_from((0.5, 0.6, 0.7), 0.5, 8)],
                        lights=lights if lights is not None else [scenes.Light(position=(1.0, 2.0, 3.0))],
                        ambient=[0.1], camera=camera, width=width, height=height, name="case")


def big_triangle(width=48, height=36):
    """A large front-facing triangle crossing three viewport borders (so
    clipTriangle fans it) plus a small vertex-coloured one in front."""
    P = [[-9, -1.0, 0.2, 1], [2.5, -9, 0.2, 1], [1.0, 9, 0.2, 1],
         [-0.3, -0.2, 0.6, 1], [0.4, -0.2, 0.6, 1], [0.0, 0.5, 0.6, 1]]
    tri = scenes._orient(np.asarray(P, np.float32), np.array([[0, 1, 2], [3, 4, 5]], np.int32),
                         lambda a, b, c: np.tile([0.0, 0.0, 1.0], (len(a), 1)))
    return _scene(P, tri, [0, -1], width, height, dict(position=(0.0, 0.0, 3.0), fov=60.0, near=0.1, far=100.0))


def depth_tie(width=48, height=36):
    """The same triangle twice: equal depths, the first one must win."""
    s = big_triangle(width, height)
    m = s.meshes[0]
    m.tri = np.concatenate([m.tri[1:2], m.tri[1:2]])
    m.mat = np.array([0, -1], np.int32)
    return s


ORTHO = dict(kind=abi.PRH_CAMERA_ORTHOGRAPHIC, position=(0.0, 0.0, 3.0), near=1.0, far=-10.0,
             frustum=(-1.0, 1.0, -1.0, 1.0))


def nan_depth(nan_first: bool, width=32, height=32):
    """A triangle that is valid in model space but projects onto the pixel-
    centre row y = 10.5 under an axis-aligned orthographic camera: every
    pixel centre of its box on that row gets Barycoord 0/0 = NaN, which passes
    the inside test (render/raster.go:491) with a NaN depth.  A NaN fragment
    only survives as the first fragment of a pixel (buffer/buffer.go:230,279):
    with nan_first it precedes a covering triangle, otherwise it follows it."""
    Y = -0.34375  # 16*Y + 16 = 10.5
    flat = [[-0.5, Y, 0.5, 1], [0.5, Y, 0.5, 1], [0.0, Y, -0.5, 1]]
    cover = [[-0.9, -0.9, 0.0, 1], [0.9, -0.9, 0.0, 1], [0.0, 0.6, 0.0, 1]]
    P = flat + cover if nan_first else cover + flat
    tri = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    mats = [0, -1] if nan_first else [-1, 0]
    return _scene(P, tri, mats, width, height, ORTHO)


def no_lights(width=40, height=30):
    s = scenes.tiny(7, width=width, height=height, n_tris=10)
    s.lights = []
    return s


def no_materials(width=40, height=30):
    s = scenes.tiny(8, width=width, height=height, n_tris=10)
    s.materials = []
    return s


TESTDATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "testdata")


def reference_bunny(width=1920, height=1080, msaa=2, root=TESTDATA):
    """render/raster_test.go:32-53 (newscene) as TestRender configures it
    (:55-64): bunny.obj rotated -pi/6 about y, scaled 4, translated
    (0.1, 0, -0.2); a point light (intensity 5, black colour) at
    (-2, 2.5, 6) and an ambient light 0.5; the camera at (0, 1.5, 1) looking
    at (0, 0, -0.5), fov 45, near 0.1, far 3; MSAA 2, background
    (0, 127, 255, 255), no gamma correction (the option's default)."""
    from polyred_amd import model
    g = model.load_obj(os.path.join(root, "bunny.obj"))
    g.rotate((0, 1, 0), float(np.float32(-np.pi / 6)))
    g.scale(4, 4, 4)
    g.translate(0.1, 0, -0.2)
    cam = dict(position=(0.0, 1.5, 1.0), target=(0.0, 0.0, -0.5), up=(0.0, 1.0, 0.0), fov=45.0,
               aspect=float(np.float32(width) / np.float32(height)), near=0.1, far=3.0)
    return model.scene_from_groups([g], [scenes.Light(position=(-2.0, 2.5, 6.0), intensity=5.0, color=(0, 0, 0, 255))],
                                   [0.5], cam, width, height, msaa=msaa, gamma=False,
                                   background=(0, 127, 255, 255), name=f"reference_bunny_{width}x{height}")


def _load(name, scale, translate=None):
    from polyred_amd import model
    g = model.load_obj(os.path.join(TESTDATA, name))
    g.scale(scale, scale, scale)
    if translate is not None:
        g.translate(*translate)
    return g


def _ref_scene(groups, lights, ambient, cam, w, h, **opts):
    from polyred_amd import model
    return model.scene_from_groups(groups, lights, ambient, cam, w, h, background=(0, 127, 255, 255), gamma=False,
                                   **opts)


def reference_shadow(two_lights=False):
    """render/gpudeferred_shadow_test.go:28-60 (and gpudeferred_shadow2_test.go
    with the second light): bunny and ground scaled 2, every material
    ReceiveShadow, point light(s) of intensity 3 casting shadows, ambient 0.7,
    200x200, framebuffer-sized shadow maps."""
    lights = [scenes.Light(position=(4.0, 4.0, 2.0), intensity=3.0, cast_shadow=True)]
    if two_lights:
        lights.append(scenes.Light(position=(-6.0, 4.0, 2.0), intensity=3.0, cast_shadow=True))
    cam = dict(position=(0.0, 0.6, 0.9), fov=45.0, aspect=1.0, near=0.1, far=2.0)
    s = _ref_scene([_load("bunny.obj", 2), _load("ground.obj", 2)], lights, [0.7], cam, 200, 200, shadow_map=True,
                   name="reference_shadow2" if two_lights else "reference_shadow")
    for m in s.materials:
        m.receive_shadow = True
    return s


def reference_multimaterial():
    """render/gpudeferred_multimat_test.go:24-50: the bunny and the gopher (7
    materials; its objects mixing triangles and quads keep no material) side
    by side, 160x160."""
    lights = [scenes.Light(position=(-2.0, 2.5, 6.0), intensity=3.0, color=(0, 0, 0, 255))]
    cam = dict(position=(0.0, 1.5, 1.0), target=(0.0, 0.0, -0.5), fov=45.0, aspect=1.0, near=0.1, far=3.0)
    return _ref_scene([_load("bunny.obj", 4, (-0.3, 0, -0.2)), _load("gopher.obj", 4, (0.4, 0, -0.2))], lights, [0.5],
                      cam, 160, 160, name="reference_multimaterial")


def reference_ao():
    """render/gpudeferred_ao_test.go:25-50: the bunny scaled 2 with
    AmbientOcclusion on, 150x150."""
    cam = dict(position=(0.0, 0.6, 0.9), fov=45.0, aspect=1.0, near=0.1, far=2.0)
    s = _ref_scene([_load("bunny.obj", 2)], [scenes.Light(position=(2.0, 3.0, 4.0), intensity=3.0)], [0.5], cam, 150,
                   150, name="reference_ao")
    for m in s.materials:
        m.ambient_occlusion = True
    return s


def reference_directional():
    """render/gpudeferred_directional_test.go:23-42: TestRender's bunny under
    a directional light (intensity 0.9, colour (255,240,220), direction
    (-1,-1,-1)) and ambient 0.3, 150x150."""
    from polyred_amd import model
    g = model.load_obj(os.path.join(TESTDATA, "bunny.obj"))
    g.rotate((0, 1, 0), float(np.float32(-np.pi / 6)))
    g.scale(4, 4, 4)
    g.translate(0.1, 0, -0.2)
    light = scenes.Light(kind=abi.PRH_LIGHT_DIRECTIONAL, direction=scenes.unit3((-1.0, -1.0, -1.0)), intensity=0.9,
                         color=(255, 240, 220, 255))
    cam = dict(position=(0.0, 1.5, 1.0), target=(0.0, 0.0, -0.5), fov=45.0, aspect=1.0, near=0.1, far=3.0)
    return _ref_scene([g], [light], [0.3], cam, 150, 150, name="reference_directional")


def stacked_layers(n_layers=100, width=160, height=120, shadows=True):
    """n_layers large screen-facing quads, one behind the other, each
    covering most of the viewport: every triangle touches half the tiles, so
    the (primitive, tile) pairs outgrow the first work-buffer estimate (4 per
    triangle plus 4 per tile) and the frame is drawn again after the buffers
    grow.  With a shadow-casting light the shadow pass overflows as well."""
    P, tri = [], []
    for k in range(n_layers):
        z = -0.02 * k
        b = len(P)
        P += [[-0.9, -0.85, z, 1], [0.92, -0.8, z, 1], [0.9, 0.88, z, 1], [-0.88, 0.9, z, 1]]
        tri += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
    P = np.asarray(P, np.float32)
    tri = scenes._orient(P, np.asarray(tri, np.int32), lambda a, b, c: np.tile([0.0, 0.0, 1.0], (len(a), 1)))
    mats = [scenes._mat_from((0.5, 0.6, 0.7), 0.5, 8, shadow=shadows)]
    lights = [scenes.Light(position=(0.5, 1.0, 3.0), cast_shadow=shadows)]
    s = _scene(P, tri, [0] * len(tri), width, height, dict(position=(0.0, 0.0, 1.8), fov=60.0, near=0.1, far=100.0),
               materials=mats, lights=lights)
    s.shadow_map = shadows
    s.name = f"stacked_{n_layers}"
    return s


# ------------------------------------------------------- pinning scenes
AO_STEP_CAMERA = dict(kind=abi.PRH_CAMERA_ORTHOGRAPHIC, position=(0.0, 0.0, 3.0), near=-1.0, far=-5.0,
                      frustum=(-4.0, 4.0, -2.0, 2.0))


def _quad(x0, x1, y0, y1, z):
    return [[x0, y0, z, 1], [x1, y0, z, 1], [x1, y1, z, 1], [x0, y1, z, 1]]


def ao_step(axis="x", width=640, height=320, h=4e-4):
    """Two fronto-parallel planes under an orthographic camera looking down
    -z (80 pixels per unit both ways): a low plane at z = 0 over the whole
    view and a high plane at z = h over x >= 0 (axis "x": raster columns from
    width/2 on) or y >= 0 (axis "y": raster rows from height/2 on).  With
    near -1 and far -5 the low plane's depth is 0 and the high plane's h/2.
    Both carry one material with AO on and a white uniform texture, and the
    scene has no lights, so the deferred colour of a pixel is
    u8(pow(total, 10000) * 255) (shader/blinn_cpu.go's texture colour,
    material/ao.go:20-38)."""
    low = _quad(-4.0, 4.0, -2.0, 2.0, 0.0)
    high = _quad(0.0, 4.0, -2.0, 2.0, h) if axis == "x" else _quad(-4.0, 4.0, 0.0, 2.0, h)
    P = np.asarray(low + high, np.float32)
    tri = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]], np.int32)
    mat = scenes.Material(diffuse=(128, 128, 128, 255), specular=(0, 0, 0, 255), texture=(255, 255, 255, 255),
                          shininess=1.0, ambient_occlusion=True)
    s = _scene(P, tri, [0, 0, 0, 0], width, height, dict(AO_STEP_CAMERA), materials=[mat], lights=[])
    s.gamma = False
    s.name = f"ao_step_{axis}"
    return s


def shadow_footprint(non_casting=False, back_face=False, width=160, height=120):
    """A ground square (y = 0, |x|, |z| <= 6) and a floating square occluder
    (y = 2, |x|, |z| <= 1), both facing up and receiving shadows, under one
    shadow-casting point light at (2, 9, 1); a perspective camera at (0, 5, 8)
    looking at the origin (fov 45, near 1, far 14); framebuffer-sized shadow
    maps as the reference allocates them (render/shadow.go:87-88).  With
    non_casting a second point light that casts no shadow is added, which
    shadingVisibility reports as occluding every pixel
    (render/shadow.go:233-236).  With back_face a third square (y = 3.5,
    x in [-3.5, -1.5], z in [1, 3]) faces down, away from the light and the
    camera: the back-face cull drops it from both the shadow map and the
    frame, so it casts no shadow (render/shadow.go:176-179)."""
    ground = _quad(-6.0, 6.0, -6.0, 6.0, 0.0)
    occ = _quad(-1.0, 1.0, -1.0, 1.0, 0.0)
    quads = ground + occ + (_quad(-3.5, -1.5, -3.0, -1.0, 0.0) if back_face else [])
    P = np.asarray(quads, np.float32)[:, [0, 2, 1, 3]]  # (x, y, z) -> (x, z, y): squares in y = const planes
    P[:, 2] *= -1
    P[4:8, 1] = 2.0
    P[8:, 1] = 3.5
    nq = len(P) // 4
    tri = np.array([[4 * q, 4 * q + 1, 4 * q + 2] for q in range(nq)] + [[4 * q, 4 * q + 2, 4 * q + 3]
                                                                        for q in range(nq)], np.int32)
    tri = tri[np.argsort(tri[:, 0], kind="stable")]
    up = np.tile([0.0, 1.0, 0.0], (1, 1))
    tri = scenes._orient(P, tri, lambda a, b, c: np.where((a[:, 1] > 3)[:, None], -up, up))
    n = len(P)
    nor = np.tile(np.array([0, 1, 0, 0], np.float32), (n, 1))
    nor[8:, 1] = -1
    mesh = scenes.Mesh(P, nor, np.tile(np.array([200, 200, 200, 255], np.uint8), (n, 1)),
                       np.zeros((n, 2), np.float32), tri, np.array([0, 0, 1, 1, 1, 1][: len(tri)], np.int32))
    mats = [scenes._mat_from((0.6, 0.6, 0.6), 0.2, 8, shadow=True), scenes._mat_from((0.8, 0.3, 0.3), 0.2, 8,
                                                                                    shadow=True)]
    lights = [scenes.Light(position=(2.0, 9.0, 1.0), intensity=1.0, cast_shadow=True)]
    if non_casting:
        lights.append(scenes.Light(position=(-3.0, 6.0, 4.0), intensity=0.5, cast_shadow=False))
    s = scenes.Scene(meshes=[mesh], obj_mesh=np.array([0]), obj_model=np.eye(4, dtype=np.float32).reshape(1, 16),
                     obj_mat_base=np.array([0]), materials=mats, lights=lights, ambient=[0.2],
                     camera=dict(position=(0.0, 5.0, 8.0), target=(0.0, 0.0, 0.0), fov=45.0, near=1.0, far=14.0),
                     width=width, height=height, shadow_map=True, gamma=False,
                     name="shadow_footprint" + ("_nc" if non_casting else "") + ("_bf" if back_face else ""))
    return s
