"""The seeded workload scenes: sizes, determinism, topology and the scene
contract fields (Scene.Center, material bases).  CPU only."""
import numpy as np

from polyred_amd import abi, scenes


def test_uv_sphere_topology():
    m = scenes.uv_sphere()
    assert m.pos.shape == (482, 4) and m.tri.shape == (960, 3)
    # closed 2-manifold: every undirected edge is shared by exactly 2 triangles
    e = np.sort(np.concatenate([m.tri[:, [0, 1]], m.tri[:, [1, 2]], m.tri[:, [2, 0]]]), axis=1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    assert np.all(counts == 2)
    # every face normal points outward
    p = m.pos[:, :3].astype(float)
    a, b, c = p[m.tri[:, 0]], p[m.tri[:, 1]], p[m.tri[:, 2]]
    assert np.all(np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) > 0)
    assert np.allclose(np.linalg.norm(m.nor[:, :3], axis=1), 1, atol=1e-6)


def test_c1_matches_config():
    s = scenes.c1_plain()
    assert (s.width, s.height) == (256, 256)
    assert s.n_triangles == 960 + 2
    assert s.lights[0].position == (2.0, 3.0, 2.0) and s.ambient == [0.1]
    d = s.materials[0].diffuse
    assert all(round(0.2 * 255) <= c <= round(0.9 * 255) for c in d[:3])
    assert s.materials[0].specular[:3] == (128, 128, 128) and s.materials[0].shininess == 32
    assert s.camera["fov"] == 45.0 and s.camera["position"] == (0.0, 1.5, 4.0)


def test_c2_matches_config():
    s = scenes.c2_scale()
    assert (s.width, s.height) == (1920, 1080)
    assert s.n_triangles == 1000 * 960 + 200 * 200 * 2
    assert len(s.materials) == 8 and len(s.lights) == 5 and s.shadow_size == 2048
    kinds = [l.kind for l in s.lights]
    assert kinds.count(abi.PRH_LIGHT_POINT) == 4 and kinds.count(abi.PRH_LIGHT_DIRECTIONAL) == 1
    assert all(l.cast_shadow for l in s.lights)
    assert all(m.ambient_occlusion and m.receive_shadow for m in s.materials)
    scl = np.linalg.norm(s.obj_model[:-1].reshape(-1, 4, 4)[:, :3, 0], axis=1)
    assert scl.min() >= 0.05 * 0.999 and scl.max() <= 0.5 * 1.001
    t = s.obj_model[:-1].reshape(-1, 4, 4)[:, :3, 3]
    assert np.abs(t).max() <= 10
    for l in s.lights[:4]:
        assert 0 <= l.position[1] <= 10


def test_seeded_determinism():
    a, b = scenes.c2_scale(n_spheres=50), scenes.c2_scale(n_spheres=50)
    assert np.array_equal(a.obj_model, b.obj_model) and a.materials == b.materials
    c = scenes.c2_scale(seed=3, n_spheres=50)
    assert not np.array_equal(a.obj_model, c.obj_model)


def test_center_is_union_of_untransformed_aabbs():
    s = scenes.c1_plain()
    c = s.center()
    # sphere [-1,1]^3 (model space), ground x in [-5,5], z in [-6,2.5] at y=-1,
    # point light (2,3,2), ambient at the origin
    assert np.allclose(c, [0.0, 1.0, -1.75])


def test_desc_roundtrip_fields():
    s = scenes.c1_plain()
    d = s.desc()
    assert d.n_meshes == 2 and d.n_objects == 2 and d.n_materials == 2 and d.n_lights == 1
    assert d.mesh_tri_off[2] == 962 and d.mesh_vtx_off[2] == 482 + 4
