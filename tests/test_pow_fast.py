"""The shading's float32(pow(x, y)) without the library's double-double pow
(pow32_fast in gomath.hpp): for the Blinn-Phong specular term's x in (0, 1]
and shininess y, it must give the float32 rounding of the float64 pow bit for
bit -- the series are evaluated to a relative error below 2^-40 and values
within that of a float32 rounding midpoint go to the library pow.  Checked
here against the float64 pow of the C library on random and boundary inputs
(mantissas on and beside binade and interval edges, x near 1 with large y),
with gomath.hpp built for the host by g++ (no GPU)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "polyred_amd", "csrc")

SRC = r'''
#include "gomath.hpp"
extern "C" void pow_many(const float* x, const float* y, float* out, long n) {
    for (long i = 0; i < n; i++) out[i] = gm::pow32_fast(x[i], y[i]);
}
'''


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("powfast")
    src, so = d / "p.cpp", d / "p.so"
    src.write_text(SRC)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", f"-I{CSRC}", str(src), "-o",
                    str(so)], check=True)
    L = C.CDLL(str(so))
    f = C.POINTER(C.c_float)
    L.pow_many.argtypes = [f, f, f, C.c_long]
    return L


def run(L, x, y):
    x, y = np.ascontiguousarray(x, np.float32), np.ascontiguousarray(y, np.float32)
    out = np.empty_like(x)
    f = C.POINTER(C.c_float)
    L.pow_many(x.ctypes.data_as(f), y.ctypes.data_as(f), out.ctypes.data_as(f), x.size)
    return out


def want(x, y):
    return np.power(x.astype(np.float64), y.astype(np.float64)).astype(np.float32)


def test_random(lib):
    rng = np.random.default_rng(3)
    n = 3_000_000
    x = rng.uniform(0, 1, n).astype(np.float32)
    x[: n // 3] = (2.0 ** rng.uniform(-40, 0, n // 3)).astype(np.float32)  # small bases
    y = np.exp(rng.uniform(np.log(2.0 ** -8), np.log(2048.0), n)).astype(np.float32)
    y[: n // 4] = rng.uniform(1, 256, n // 4).astype(np.float32)
    got, ref = run(lib, x, y), want(x, y)
    bad = np.nonzero(got.view(np.uint32) != ref.view(np.uint32))[0]
    assert bad.size == 0, (x[bad[:5]], y[bad[:5]], got[bad[:5]], ref[bad[:5]])
    assert (ref == 0).mean() > 0.01 and ((ref > 0) & (ref < 2.0 ** -126)).any()  # zero and subnormal results


def test_large_exponents_and_bases_near_one(lib):
    # the error budget's worst case: y up to 2048 amplifies log2's error, and
    # x near 1 puts log2 at its smallest
    rng = np.random.default_rng(5)
    n = 4_000_000
    x = (1.0 - rng.uniform(0, 2.0 ** -6, n)).astype(np.float32)
    x[: n // 2] = rng.uniform(0.5, 1.0, n // 2).astype(np.float32)
    y = rng.uniform(256, 2048, n).astype(np.float32)
    y[: n // 4] = np.round(y[: n // 4])  # integer shininess
    got, ref = run(lib, x, y), want(x, y)
    bad = np.nonzero(got.view(np.uint32) != ref.view(np.uint32))[0]
    assert bad.size == 0, (x[bad[:5]], y[bad[:5]], got[bad[:5]], ref[bad[:5]])


def test_mantissa_edges(lib):
    # bases on and beside 64 evenly spaced mantissas of several binades (the
    # sqrt(1/2) split of the mantissa included among their neighbours)
    m = 1.0 + np.arange(65) / 64.0
    base = np.concatenate([m * 2.0 ** -k for k in (1, 2, 7, 20)]).astype(np.float32)
    base = base[base <= 1.0]
    xs = np.concatenate([base, np.nextafter(base, np.float32(0)), np.nextafter(base, np.float32(2))])
    xs = xs[(xs > 0) & (xs <= 1.0)].astype(np.float32)
    ys = np.array([0.00390625, 0.3, 1.0, 2.5, 8.0, 33.3, 100.0, 517.0, 2048.0], np.float32)
    x = np.repeat(xs, ys.size)
    y = np.tile(ys, xs.size)
    got, ref = run(lib, x, y), want(x, y)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_special_and_outside(lib):
    x = np.array([1.0, 1.0, 0.0, 0.0, 0.5, 0.5, 0.5, 1e-30, np.nextafter(np.float32(1), np.float32(0)), 0.7, 0.7,
                  0.7, np.nan, 0.5, 0.0, 0.0, 0.0, -0.0, -0.0, -0.0, 0.0, -0.0], np.float32)
    y = np.array([32.0, 0.5, 5.0, 0.0, 0.0, 1.0, 150.0, 5.0, 2048.0, 3000.0, 1e-5, -2.0, 2.0, np.nan, 1e-9, 3.0,
                  np.inf, 3.0, 2.0, 0.5, -1.0, -3.0], np.float32)
    got, ref = run(lib, x, y), want(x, y)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (got, ref)


def test_exact_powers(lib):
    # 2^-k to integer powers: exact results (no rounding to decide)
    k = np.arange(1, 40)
    x = (2.0 ** -k).astype(np.float32)
    for yy in (1.0, 2.0, 3.0, 4.0):
        y = np.full(k.size, yy, np.float32)
        assert np.array_equal(run(lib, x, y), want(x, y))
