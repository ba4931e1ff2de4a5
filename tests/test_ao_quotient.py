"""The AO kernel's table paths order the taps by q = e * fl64(1 / d) and take
float32(q) as the rounded quotient, where the reference computes the float32
quotient e / d (material/ao.go:64-70) and keeps the largest.  That is exact
when q is within 2^-52 of e / d while distinct quotients of float32 e and d
(d >= 1) differ relatively by at least 2^-47 and e / d stays 2^-47 away from
every float32 rounding boundary (equal quotients may order either way in q,
and round to the same float32).  This checks these consequences on random
and adversarial operands (quotients next to float32 midpoints, near-ties),
with no GPU."""
import numpy as np

F32 = np.float32


def _q(e, d):
    return (e.astype(np.float64) * (1.0 / d.astype(np.float64))).astype(np.float64)


def test_rounded_quotient_is_float32_division():
    rng = np.random.default_rng(11)
    n = 400_000
    d = np.concatenate([np.arange(1, 100, dtype=F32), rng.uniform(1, 150, n).astype(F32)])
    mag = 2.0 ** rng.integers(-140, 10, d.size)
    e = (rng.uniform(-1, 1, d.size) * mag).astype(F32)
    want = e / d  # float32 division, correctly rounded
    got = _q(e, d).astype(F32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # adversarial: e = fl32(m * d) for m half way between two float32s
    m32 = rng.uniform(1e-3, 10, n).astype(F32)
    mid = (m32.astype(np.float64) + np.nextafter(m32, F32(np.inf)).astype(np.float64)) / 2
    dd = rng.uniform(1, 120, n).astype(F32)
    ee = (mid * dd.astype(np.float64)).astype(F32)
    want = ee / dd
    got = _q(ee, dd).astype(F32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_q_orders_quotients_exactly():
    rng = np.random.default_rng(12)
    n = 400_000
    d1 = rng.integers(1, 100, n).astype(F32)
    d2 = np.where(rng.random(n) < 0.5, rng.integers(1, 100, n), rng.uniform(1, 140, n)).astype(F32)
    e1 = (rng.uniform(-1, 1, n) * 2.0 ** rng.integers(-30, 2, n)).astype(F32)
    # near-ties: e2 / d2 within a few float32 ulps of e1 / d1
    e2 = (e1.astype(np.float64) * d2 / d1).astype(F32)
    toward = (np.sign(rng.random(n) - 0.5) * np.inf).astype(F32)
    e2 = np.where(rng.random(n) < 0.5, e2, np.nextafter(e2, toward)).astype(F32)
    assert e2.dtype == F32
    exact = np.sign(e1.astype(np.float64) * d2 - e2.astype(np.float64) * d1)  # products of float32: exact
    q1, q2 = _q(e1, d1), _q(e2, d2)
    ne = exact != 0
    assert np.array_equal(exact[ne], np.sign(q1 - q2)[ne])
    # equal quotients may round apart in q, but to the same float32
    assert np.array_equal(q1[~ne].astype(F32), q2[~ne].astype(F32))
    assert ne.sum() > n // 2 and (~ne).sum() > 1000
