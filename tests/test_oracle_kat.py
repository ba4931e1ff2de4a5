"""The oracle against polyred's own known-answer tests (tests/golden/
reference_kats.json holds their inputs and expected outputs) and against
closed forms.  CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

from polyred_amd import abi, scenes


def f32(*v):
    return np.ascontiguousarray(np.array(v, dtype=np.float32).reshape(-1))


def fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def approx_eq(got, want, eps):
    """math.ApproxEq in float32 (math/math.go:41-44)."""
    got = np.asarray(got, np.float32)
    want = np.asarray(want, np.float32)
    return bool(np.all(np.abs((got - want).astype(np.float32)) <= np.float32(eps)))


def test_barycoord(orc, kats):
    k = kats["barycoord"]
    out = f32(0, 0, 0)
    orc.lib().oracle_barycoord(fp(f32(*k["p"])), fp(f32(*k["v1"])), fp(f32(*k["v2"])), fp(f32(*k["v3"])), fp(out))
    assert out.tolist() == k["want"]


def test_barycoord_sums_to_one_and_vertices(orc):
    # closed form: barycentrics of a vertex are a unit vector; inside points sum to ~1
    L = orc.lib()
    a, b, c = f32(1, 2), f32(7, 3), f32(2, 9)
    out = f32(0, 0, 0)
    for v, want in ((a, [1, 0, 0]), (b, [0, 1, 0]), (c, [0, 0, 1])):
        L.oracle_barycoord(fp(v), fp(a), fp(b), fp(c), fp(out))
        assert out.tolist() == want
    rng = np.random.default_rng(3)
    for _ in range(200):
        p = f32(*rng.uniform(0, 10, 2))
        L.oracle_barycoord(fp(p), fp(a), fp(b), fp(c), fp(out))
        assert abs(float(out.sum()) - 1.0) < 1e-5
        # against exact rational barycentrics
        from fractions import Fraction as F
        P = [F(float(x)) for x in p]
        A, B, Cc = ([F(float(x)) for x in v] for v in (a, b, c))
        area = (B[0] - A[0]) * (Cc[1] - A[1]) - (B[1] - A[1]) * (Cc[0] - A[0])
        w3 = ((B[0] - A[0]) * (P[1] - A[1]) - (B[1] - A[1]) * (P[0] - A[0])) / area
        w2 = ((P[0] - A[0]) * (Cc[1] - A[1]) - (P[1] - A[1]) * (Cc[0] - A[0])) / area
        assert abs(float(out[1]) - float(w2)) < 1e-5 and abs(float(out[2]) - float(w3)) < 1e-5


def test_viewport_matrix(orc, kats):
    k = kats["viewport"]
    out = np.zeros(16, np.float32)
    orc.lib().oracle_viewport_matrix(k["w"], k["h"], fp(out))
    assert approx_eq(out, k["want"], k["eps"])


def test_view_matrix(orc, kats):
    k = kats["view_matrix"]
    out = np.zeros(16, np.float32)
    orc.lib().oracle_view_matrix(fp(f32(*k["pos"])), fp(f32(*k["target"])), fp(f32(*k["up"])), fp(out))
    assert approx_eq(out, k["want"], k["eps"])


def test_proj_matrices(orc, kats):
    k = kats["persp_proj"]
    cam = abi.prh_camera(kind=abi.PRH_CAMERA_PERSPECTIVE, fov=k["fov"], aspect=k["aspect"], near_=k["near"],
                         far_=k["far"])
    out = np.zeros(16, np.float32)
    orc.lib().oracle_proj_matrix(C.byref(cam), fp(out))
    assert approx_eq(out, k["want"], k["eps"])
    k = kats["ortho_proj"]
    cam = abi.prh_camera(kind=abi.PRH_CAMERA_ORTHOGRAPHIC, left=k["left"], right=k["right"], bottom=k["bottom"],
                         top=k["top"], near_=k["near"], far_=k["far"])
    orc.lib().oracle_proj_matrix(C.byref(cam), fp(out))
    assert approx_eq(out, k["want"], k["eps"])


def test_mulm_det_inv(orc, kats):
    L = orc.lib()
    k = kats["mulm"]
    out = np.zeros(16, np.float32)
    L.oracle_mat4_mulm(fp(f32(*k["a"])), fp(f32(*k["b"])), fp(out))
    assert out.tolist() == k["want"]
    k = kats["inv"]
    L.oracle_mat4_inv(fp(f32(*k["m"])), fp(out))
    want = [n / d for n, d in zip(k["want_num"], k["want_den"])]
    assert approx_eq(out, want, k["eps"])
    # the determinant KAT through the inverse: inv of a matrix whose det is
    # -44 is adjugate/-44; check M.inv(M) ~ I closed form
    k = kats["det"]
    m = f32(*k["m"])
    L.oracle_mat4_inv(fp(m), fp(out))
    prod = m.reshape(4, 4).astype(np.float64) @ out.reshape(4, 4).astype(np.float64)
    assert np.allclose(prod, np.eye(4), atol=1e-5)
    adj_scaled = np.linalg.inv(m.reshape(4, 4).astype(np.float64))
    assert np.allclose(out.reshape(4, 4), adj_scaled, atol=1e-6)
    assert round(np.linalg.det(m.reshape(4, 4).astype(np.float64))) == k["want"]


def test_aabb(orc, kats):
    L = orc.lib()
    k = kats["aabb_new"]
    mn, mx = f32(0, 0, 0), f32(0, 0, 0)
    L.oracle_aabb_new(3, fp(f32(*sum(k["pts"], []))), fp(mn), fp(mx))
    assert mn.tolist() == k["min"] and mx.tolist() == k["max"]

    def box(pts):
        a, b = f32(0, 0, 0), f32(0, 0, 0)
        L.oracle_aabb_new(len(pts), fp(f32(*sum(pts, []))), fp(a), fp(b))
        return a, b

    for case in kats["aabb_intersect"]["cases"]:
        a0, a1 = box(case["a"])
        b0, b1 = box(case["b"])
        assert bool(L.oracle_aabb_intersect(fp(a0), fp(a1), fp(b0), fp(b1))) == case["want"]
    k = kats["aabb_contains"]
    b0, b1 = box(k["box"])
    assert L.oracle_aabb_contains(fp(b0), fp(b1), 3, fp(f32(*sum(k["inside"], [])))) == 1
    assert L.oracle_aabb_contains(fp(b0), fp(b1), 1, fp(f32(*sum(k["outside"], [])))) == 0


def test_aabb_intersect_z_uses_max_y(orc):
    # geometry/primitive/box.go:38 compares the z extent against Max.Y: a box
    # entirely above z = Max.Z but below z = Max.Y still "intersects"
    L = orc.lib()
    assert L.oracle_aabb_intersect(fp(f32(0, 0, -1)), fp(f32(10, 10, 1)), fp(f32(1, 1, 5)), fp(f32(2, 2, 6))) == 1
    assert L.oracle_aabb_intersect(fp(f32(0, 0, -1)), fp(f32(10, 10, 1)), fp(f32(1, 1, 11)), fp(f32(2, 2, 12))) == 0


def test_tri_is_valid(orc, kats):
    for case in kats["tri_is_valid"]["cases"]:
        p = [f32(*v) for v in case["p"]]
        assert bool(orc.lib().oracle_tri_is_valid(fp(p[0]), fp(p[1]), fp(p[2]))) == case["want"]


def test_interp_worldpos(orc, kats):
    k = kats["interp_worldpos"]
    out = f32(0, 0, 0, 0)
    m1, m2, m3 = f32(*k["m1"]), f32(*k["m2"]), f32(*k["m3"])
    third = np.float32(1.0) / np.float32(3.0)
    orc.lib().oracle_interp_worldpos(fp(f32(third, third, third)), fp(m1), fp(m2), fp(m3), fp(out))
    s = float(out[0] + out[1] + out[2])
    assert k["centroid_sum"][0] < s < k["centroid_sum"][1]
    orc.lib().oracle_interp_worldpos(fp(f32(1, 0, 0)), fp(m1), fp(m2), fp(m3), fp(out))
    assert out[:3].tolist() == m1[:3].tolist()


def test_mat_at(orc, kats):
    k = kats["mat_at"]
    L = orc.lib()
    for i in k["nil"]:
        assert L.oracle_mat_at(k["n"], i) == -1
    for i in k["valid"]:
        assert L.oracle_mat_at(k["n"], i) == i
    for i in k["empty_nil"]:
        assert L.oracle_mat_at(0, i) == -1


def test_srgb(orc, kats):
    L = orc.lib()
    k = kats["srgb"]
    v = L.oracle_srgb_from_linear(k["roundtrip"])
    assert approx_eq(L.oracle_linear_from_srgb(v), k["roundtrip"], k["roundtrip_eps"])
    for x, want in k["endpoints"]:
        assert L.oracle_srgb_from_linear(x) == want
        assert L.oracle_linear_from_srgb(x) == want


def test_gamma_table_against_analytic_srgb(orc):
    """Every 8-bit value: the LUT path is within 1 LSB of the analytic sRGB
    curve, and its rounding margin is wide enough that the last bits of the
    float64 pow that builds the LUT cannot change a result."""
    L = orc.lib()
    for c in range(256):
        v = c / 255.0
        s = 12.92 * v if v <= 0.0031308 else 1.055 * v ** (1 / 2.4) - 0.055
        g = L.oracle_gamma_u8(c)
        assert abs(g - math.floor(s * 255 + 0.5)) <= 1, c
        x = float(L.oracle_srgb_from_linear(np.float32(c) / np.float32(255))) * 255 + 0.5
        assert c in (0, 255) or abs(x - round(x)) > 1e-4, (c, x)


def test_shadow_quantisation(orc):
    # u8(float32(c) * 0.5^k), render/raster.go:348-352; closed form floor(c/2^k)
    L = orc.lib()
    for c in (0, 1, 127, 128, 200, 255):
        for k in range(6):
            assert L.oracle_shadow_quant(c, k) == c >> k


def test_ao_directions(orc):
    out = np.zeros(16, np.float32)
    orc.lib().oracle_ao_dirs(fp(out))
    d = out.reshape(8, 2).astype(np.float64)
    ang = np.arange(8) * math.pi / 4
    assert np.allclose(d[:, 0], np.cos(ang), atol=1e-6) and np.allclose(d[:, 1], np.sin(ang), atol=1e-6)


def _shade_fp64(k, n, x):
    """Blinn-Phong of shader/blinn_cpu.go in float64 (independent restatement
    in numpy, used as the reference of the equivalence fixture)."""
    tex = np.array(k["texture"][:3], float)
    acc = sum(k["ambient"]) * tex
    cam = np.array(k["cam"], float)
    Vv = cam - x
    Vv /= np.linalg.norm(Vv)
    ld = np.zeros(3)
    ls = np.zeros(3)
    for l in k["lights"]:
        if l["kind"] == 0:
            d = np.array(l["position"], float) - x
            Lv = d / np.linalg.norm(d)
            I = l["intensity"] / np.linalg.norm(d)
        else:
            dd = np.array(l["direction"], float)
            Lv = -dd / np.linalg.norm(dd)
            I = l["intensity"]
        H = Lv + Vv
        H /= np.linalg.norm(H)
        ld += min(max(n @ Lv, 0), 1) * tex * I
        ls += min(max(n @ H, 0), 1) ** k["shininess"] * np.array(l["color"][:3], float) * I
    return acc + np.array(k["diffuse"][:3]) * ld / 255 + np.array(k["specular"][:3]) * ls / 255


def test_fragment_shader_equivalence_fixture(orc, kats):
    """render/shading_equiv_test.go's six fragments: FragmentShader within
    1 LSB of an FP64 Blinn-Phong, linear sums within 1e-3 relative."""
    k = kats["shading_equiv"]
    m = abi.prh_material()
    m.texture[:] = k["texture"]
    m.diffuse[:] = k["diffuse"]
    m.specular[:] = k["specular"]
    m.shininess = k["shininess"]
    lights = (abi.prh_light * 2)()
    for i, l in enumerate(k["lights"]):
        lights[i].kind = l["kind"]
        lights[i].intensity = l["intensity"]
        lights[i].color[:] = l["color"]
        if l["kind"] == 0:
            lights[i].position[:] = l["position"]
        else:
            lights[i].direction[:] = scenes.unit3(l["direction"])
    amb = f32(*k["ambient"])
    out = (C.c_uint8 * 4)()
    lin = f32(0, 0, 0)
    for nrm, pos in zip(k["normals"], k["positions"]):
        n = f32(*nrm, 0)
        x = f32(*pos, 1)
        orc.lib().oracle_fragment_shader(C.byref(m), fp(n), fp(n), fp(x), fp(f32(*k["cam"])), 2, lights, 1, fp(amb),
                                         out, fp(lin))
        want = _shade_fp64(k, n[:3].astype(float), x[:3].astype(float))
        assert np.allclose(lin, want, rtol=1e-3, atol=1e-3), (lin, want)
        q = np.clip(np.floor(want + 0.5), 0, 255)
        assert np.all(np.abs(np.array(out[:3], float) - q) <= k["tol_lsb"]), (list(out), q)
        assert out[3] == k["texture"][3]
