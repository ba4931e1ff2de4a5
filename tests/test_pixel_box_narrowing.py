"""The rasters narrow a small, well-shaped triangle's pixel box (the
reference's int(Round(min)-1) .. int(Round(max)+1), render/raster.go:473-477)
to the pixels whose centre lies within 0.24 of the triangle's AABB, when its
larger side W is in [2^-20, 64] and |Sabc| >= W^2/4.  That is only sound if every
pixel dropped fails the inside test (render/raster.go:491) as the reference
computes it.  This checks it with the oracle's Barycoord (math/interpolate.go:
66-78, float32 as the reference) over many random triangles, including ones
at the edge of the shape condition, with no GPU."""
import ctypes as C

import numpy as np

F32 = np.float32


def _fma64(a, b, c):
    # float32(fma(a, b, c)) on float32 inputs: a*b is exact in float64
    return F32(np.float64(a) * np.float64(b) + np.float64(c))


def _sabc(ax, ay, bx, by, cx, cy):
    abx, aby, acx, acy = F32(bx - ax), F32(by - ay), F32(cx - ax), F32(cy - ay)
    return _fma64(abx, acy, -F32(aby * acx))


def _boxes(t):
    (ax, ay), (bx, by), (cx, cy) = t
    mnx, mxx = min(ax, bx, cx), max(ax, bx, cx)
    mny, mxy = min(ay, by, cy), max(ay, by, cy)
    rnd = lambda v: int(np.floor(np.float64(v) + 0.5))  # noqa: E731  Go's Round for positive v
    ref = (rnd(mnx) - 1, rnd(mxx) + 1, rnd(mny) - 1, rnd(mxy) + 1)
    w = F32(max(F32(mxx - mnx), F32(mxy - mny)))
    if not (w <= 64 and w >= F32(2.0 ** -20)):
        return ref, None
    if not (abs(_sabc(ax, ay, bx, by, cx, cy)) >= F32(F32(0.25) * w * w)):
        return ref, None
    nar = (max(ref[0], int(np.ceil(F32(mnx - F32(0.74))))), min(ref[1], int(np.floor(F32(mxx - F32(0.26))))),
           max(ref[2], int(np.ceil(F32(mny - F32(0.74))))), min(ref[3], int(np.floor(F32(mxy - F32(0.26))))))
    return ref, nar


def test_dropped_pixels_fail_the_inside_test(orc):
    L = orc.lib()
    rng = np.random.default_rng(2024)
    out = (C.c_float * 3)()
    fp = lambda *v: (C.c_float * len(v))(*v)  # noqa: E731
    narrowed = dropped = 0
    for k in range(20000):
        size = float(np.exp(rng.uniform(np.log(0.05), np.log(64))))
        c0 = rng.uniform(0, 500, 2)
        pts = c0 + rng.uniform(0, size, (3, 2))
        if k % 4 == 0:  # near-slivers at the edge of the shape condition
            pts[2] = pts[0] + (pts[1] - pts[0]) * rng.uniform(0.2, 0.8) + rng.uniform(-0.3, 0.3, 2) * size * 0.3
        t = [(F32(x), F32(y)) for x, y in pts]
        ref, nar = _boxes(t)
        if nar is None:
            continue
        narrowed += 1
        (ax, ay), (bx, by), (cx, cy) = t
        for x in range(ref[0], ref[1] + 1):
            for y in range(ref[2], ref[3] + 1):
                if nar[0] <= x <= nar[1] and nar[2] <= y <= nar[3]:
                    continue
                dropped += 1
                L.oracle_barycoord(fp(x + 0.5, y + 0.5), fp(ax, ay), fp(bx, by), fp(cx, cy), out)
                bc = np.array(out[:], F32)
                assert (bc < F32(-1e-7)).any(), (t, x, y, bc)
    assert narrowed > 3000 and dropped > 30000, (narrowed, dropped)
