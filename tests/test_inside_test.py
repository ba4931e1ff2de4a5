"""The rasters' division-free inside test (cover_z in gomath.hpp): for a
finite non-zero Sabc, fl(S / Sabc) < float32(-1e-7) (render/raster.go:491,
Barycoord's float32 quotient, math/interpolate.go:76) holds exactly when
S < M * Sabc (S > M * Sabc for Sabc < 0), M the midpoint between
float32(-1e-7) and the float32 below it, the product taken in float64.  The
claim is about float32 pairs (S, Sabc), checked here on random pairs and on
pairs built around the boundary, with no GPU."""
import numpy as np

F32 = np.float32
T = F32(-1e-7)
M = (np.float64(T) + np.float64(np.nextafter(T, F32(-np.inf)))) / 2


def test_midpoint_constant():
    assert M.hex() == "-0x1.ad7f2b0000000p-24"  # BC_EDGE_MID


def _check(s, sabc):
    s, sabc = s.astype(F32), sabc.astype(F32)
    ok = np.isfinite(sabc) & (sabc != 0) & np.isfinite(s)
    s, sabc = s[ok], sabc[ok]
    with np.errstate(all="ignore"):
        ref = (s / sabc) < T  # float32 division, correctly rounded
    ms = M * sabc.astype(np.float64)
    fast = np.where(sabc > 0, s.astype(np.float64) < ms, s.astype(np.float64) > ms)
    assert np.array_equal(ref, fast), np.nonzero(ref != fast)[0][:5]
    return ref


def test_random_pairs():
    rng = np.random.default_rng(5)
    n = 2_000_000
    sabc = (rng.choice([-1, 1], n) * 2.0 ** rng.uniform(-30, 30, n)).astype(F32)
    s = (rng.normal(size=n) * np.abs(sabc) * 2.0 ** rng.integers(-30, 3, n)).astype(F32)
    ref = _check(s, sabc)
    assert 0.2 < ref.mean() < 0.8


def test_pairs_at_the_boundary():
    rng = np.random.default_rng(6)
    n = 2_000_000
    sabc = (rng.choice([-1, 1], n) * 2.0 ** rng.uniform(-40, 40, n)).astype(F32)
    # S within a few float32 ulps of M * Sabc, either side
    s = (M * sabc.astype(np.float64)).astype(F32)
    k = rng.integers(-3, 4, n)
    for _ in range(3):
        s = np.where(k > 0, np.nextafter(s, F32(np.inf)), np.where(k < 0, np.nextafter(s, F32(-np.inf)), s))
        k = k - np.sign(k)
    ref = _check(s, sabc)
    assert 0.2 < ref.mean() < 0.8
