"""The clipper skips a screen plane when all three vertices lie in front of it
by a margin (2^-12 of the largest coordinate plus 2^-10): the pass over such
a plane must give back the polygon unchanged.  Checked here on the host with
gomath.hpp's arithmetic (g++, no GPU): Sutherland-Hodgman against the six
planes of clipping.go:31-65, with and without the skip, on random triangles
crossing the viewport and on adversarial ones (vertices just past the margin,
huge coordinates from near-plane projections, edges almost parallel to a
plane); the two polygons must match bit for bit."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "polyred_amd", "csrc")

SRC = r'''
#include <cmath>
#include "gomath.hpp"
using namespace gm;
struct Plane { V4 pos, nor; };
static bool in_front(const Plane& p, V4 v) { return dot(sub(v, p.pos), p.nor) > 0; }
static V4 isect(const Plane& p, V4 v0, V4 v1) {
    V4 u = sub(v1, v0), w = sub(v0, p.pos);
    float d = dot(p.nor, u);
    float n = -dot(p.nor, w);
    float s = n / d;
    return add(v0, scale(u, s, s, s, s));
}
// returns the vertex count; out holds x y z w per vertex
static int clip(const float* t, float W, float H, int skip, float* out) {
    Plane planes[6] = {{v4(W, 0, 0, 1), v4(-1, 0, 0, 1)}, {v4(0, 0, 0, 1), v4(1, 0, 0, 1)},
                       {v4(0, H, 0, 1), v4(0, -1, 0, 1)}, {v4(0, 0, 0, 1), v4(0, 1, 0, 1)},
                       {v4(0, 0, 1, 1), v4(0, 0, -1, 1)}, {v4(0, 0, -1, 1), v4(0, 0, 1, 1)}};
    V4 a[16], b[16];
    for (int i = 0; i < 3; i++) a[i] = v4(t[3 * i], t[3 * i + 1], t[3 * i + 2], 1);
    const V4 p1 = a[0], p2 = a[1], p3 = a[2];
    float big = 0.0f;
    for (int i = 0; i < 3; i++)
        big = std::fmax(big, std::fmax(std::fmax(std::fabs(a[i].x), std::fabs(a[i].y)), std::fabs(a[i].z)));
    const float margin = big * 0x1p-12f + 0x1p-10f;
    int na = 3;
    for (int pi = 0; pi < 6; pi++) {
        if (na == 0) break;
        const Plane& p = planes[pi];
        if (skip && dot(sub(p1, p.pos), p.nor) > margin && dot(sub(p2, p.pos), p.nor) > margin &&
            dot(sub(p3, p.pos), p.nor) > margin)
            continue;
        int nb = 0;
        for (int i = 0; i < na; i++) {
            const V4 e = a[i], s = a[i == 0 ? na - 1 : i - 1];
            const bool fe = in_front(p, e), fs = in_front(p, s);
            if (fe) {
                if (!fs && nb < 12) b[nb++] = isect(p, s, e);
                if (nb < 12) b[nb++] = e;
            } else if (fs && nb < 12) {
                b[nb++] = isect(p, s, e);
            }
        }
        for (int i = 0; i < nb; i++) a[i] = b[i];
        na = nb;
    }
    for (int i = 0; i < na; i++) out[4 * i] = a[i].x, out[4 * i + 1] = a[i].y, out[4 * i + 2] = a[i].z, out[4 * i + 3] = a[i].w;
    return na;
}
extern "C" int clip_pair(const float* tris, long n, float W, float H) {  // index of the first mismatch, or -1
    float o0[64], o1[64];
    for (long k = 0; k < n; k++) {
        const int n0 = clip(tris + 9 * k, W, H, 0, o0), n1 = clip(tris + 9 * k, W, H, 1, o1);
        if (n0 != n1) return (int)k;
        for (int i = 0; i < 4 * n0; i++) {
            unsigned u0, u1;
            __builtin_memcpy(&u0, &o0[i], 4), __builtin_memcpy(&u1, &o1[i], 4);
            if (u0 != u1 && !(o0[i] != o0[i] && o1[i] != o1[i])) return (int)k;
        }
    }
    return -1;
}
extern "C" long count_skips(const float* tris, long n, float W, float H) {  // planes skipped in all
    long c = 0;
    for (long k = 0; k < n; k++) {
        const float* t = tris + 9 * k;
        float big = 0.0f;
        for (int i = 0; i < 9; i++) big = std::fmax(big, std::fabs(t[i]));
        const float margin = big * 0x1p-12f + 0x1p-10f;
        const float lim[6][2] = {{0, W}, {0, 0}, {1, H}, {1, 0}, {2, 1}, {2, -1}};
        for (int p = 0; p < 6; p++) {
            bool all = true;
            for (int i = 0; i < 3; i++) {
                const float v = t[3 * i + (int)lim[p][0]];
                const float d = (p % 2 == 0) ? lim[p][1] - v : v - lim[p][1];
                all = all && d > margin;
            }
            c += all;
        }
    }
    return c;
}
'''

W, H = 1920.0, 1080.0


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("clipskip")
    src, so = d / "c.cpp", d / "c.so"
    src.write_text(SRC)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", f"-I{CSRC}", str(src), "-o",
                    str(so)], check=True)
    L = C.CDLL(str(so))
    f = C.POINTER(C.c_float)
    L.clip_pair.argtypes = [f, C.c_long, C.c_float, C.c_float]
    L.count_skips.argtypes = [f, C.c_long, C.c_float, C.c_float]
    L.count_skips.restype = C.c_long
    return L


def run(L, tris):
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    f = C.POINTER(C.c_float)
    return L.clip_pair(t.ctypes.data_as(f), t.shape[0], W, H), L.count_skips(t.ctypes.data_as(f), t.shape[0], W, H)


def test_random_crossing(lib):
    rng = np.random.default_rng(21)
    n = 400_000
    t = np.empty((n, 3, 3), np.float32)
    t[..., 0] = rng.uniform(-800, W + 800, (n, 3))
    t[..., 1] = rng.uniform(-800, H + 800, (n, 3))
    t[..., 2] = rng.uniform(-1.5, 1.5, (n, 3))
    bad, skips = run(lib, t)
    assert bad == -1, t[bad]
    assert skips > n  # the skip does happen


def test_huge_and_near_margin(lib):
    rng = np.random.default_rng(22)
    n = 300_000
    scale = 10.0 ** rng.uniform(0, 8, (n, 1, 1))
    t = (rng.uniform(-1, 1, (n, 3, 3)) * scale).astype(np.float32)
    t[..., 2] = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    bad, _ = run(lib, t)
    assert bad == -1, t[bad]
    # vertices just beyond the margin of one plane, the rest crossing others
    m = 50_000
    u = np.empty((m, 3, 3), np.float32)
    u[..., 0] = rng.uniform(-50, W + 50, (m, 3))
    u[..., 1] = rng.uniform(-50, H + 50, (m, 3))
    u[..., 2] = rng.uniform(-1.2, 1.2, (m, 3))
    big = np.abs(u).reshape(m, -1).max(1)
    eps = (big * 2.0 ** -12 + 2.0 ** -10).astype(np.float32)
    k = rng.integers(0, 3, m)
    u[np.arange(m), k, 0] = (W - eps * rng.uniform(1.0, 1.01, m)).astype(np.float32)  # near the right plane
    bad, _ = run(lib, u)
    assert bad == -1, u[bad]


def test_almost_parallel_edges(lib):
    rng = np.random.default_rng(23)
    n = 200_000
    t = np.empty((n, 3, 3), np.float32)
    x0 = rng.uniform(-100, W + 100, n)
    t[:, 0, 0] = x0
    t[:, 1, 0] = x0 + rng.uniform(-1e-3, 1e-3, n) * rng.choice([1, 1e3], n)
    t[:, 2, 0] = rng.uniform(-500, W + 500, n)
    t[..., 1] = rng.uniform(-300, H + 300, (n, 3))
    t[..., 2] = rng.uniform(-1.1, 1.1, (n, 3))
    bad, _ = run(lib, t)
    assert bad == -1, t[bad]
