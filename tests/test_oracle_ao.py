"""Pins the oracle's SSAO (material/ao.go:20-73) with answers the geometry
fixes, so that the GPU's bit-exact agreement with the oracle is agreement
with the reference.  CPU only.

maxElevationAngle marches t = 0..99 along (cos a, sin a) from the pixel,
reads the fragment at (int(x + t cos a), int(y + t sin a)) (Go's int()
truncates toward zero), stops at the first tap outside the buffer, skips
taps closer than 1, and keeps the largest atan(elevation / distance), with
an uncovered fragment's depth taken as -1; AmbientOcclusionShade sums
pi/2 - max over the eight directions a = k*pi/4, divides by 4*pi and raises
the result to the 10000th power.  The closed forms below follow from that
text and from where the geometry puts each tap:

* On a depth step of height D at distance k, the largest angle along an
  axis direction crossing it is atan(D / k) (the nearest tap across), along
  a diagonal one atan(D / t0) with t0 = ceil(k / cos(pi/4)), the first t
  whose truncated offset reaches k; every direction that does not cross, or
  looks down the step, gives 0.
* On a depth ramp of slope g along +x, the +x axis gives atan(g) (every tap
  the same ratio) and the +x diagonals atan(g * max_t floor(c t) / t).
* Interior pixels are at coordinates >= 257, where an axis direction's
  minor component (|sin(float32(pi))| < 1e-7) never moves a tap off the
  pixel's row or column; the diagonal taps' offsets c*t stay at least 0.003
  from an integer for t <= 99, far above float32 rounding at these
  coordinates.

A plausible slip -- the elevation's sign, the -1 for uncovered fragments,
truncation replaced by floor at the buffer's low edges, the dist < 1 skip
(without it t = 0 gives 0/0 and Go's math.Max propagates the NaN), the
march order with its early return at the border, the normalisation or the
exponent -- fails at least one test here.
"""
import math

import numpy as np
import pytest

from . import cases

W = H = 512
D = 1e-3  # step height in depth units


def dirs(orc):
    out = np.zeros(16, np.float32)
    orc.lib().oracle_ao_dirs(orc.fptr(out))
    return out.reshape(8, 2)


def total_of(angles):
    """AmbientOcclusionShade's total before the power: sum(pi/2 - a)/(4 pi)."""
    return 1.0 - sum(angles) / (4 * math.pi)


def check(orc, depth, ok, px, py, expect, rtol=1e-4, atol=1e-9):
    ang, tot, fac = orc.ao_march(depth, ok, px, py)
    assert np.all(np.isfinite(ang)), ang
    for k in range(8):
        e = expect[k]
        assert abs(float(ang[k]) - e) <= atol + rtol * abs(e), (k, float(ang[k]), e, px, py)
    t_ref = total_of(expect)
    assert abs(float(tot) - t_ref) <= 6e-7, (float(tot), t_ref)
    # pow(total, 10000): a float32 ulp of the total moves the factor by 1e4 * 6e-8
    assert abs(float(fac) - t_ref ** 10000) <= 1e4 * 8e-7 * t_ref ** 10000 + 1e-7, (float(fac), t_ref ** 10000)
    return ang


def step_x(x0, d=D, low=0.0):
    depth = np.full((H, W), low, np.float32)
    depth[:, x0:] = low + d
    return depth


def t_diag(k, c):
    """First t with floor(c t) >= k: a diagonal's offset along a positive
    component, truncated, reaches k."""
    return math.ceil(k / c)


def t_diag_neg(k, c):
    """First t with ceil(c t) >= k: along a negative component x - c t
    truncates to x - ceil(c t) (c t is never an integer for t <= 99)."""
    return math.floor((k - 1) / c) + 1


@pytest.mark.parametrize("k", [1, 2, 3, 7, 20, 50, 69])
def test_step_low_side(orc, k):
    """A step up at column 350 seen from k pixels before it: the +x axis sees
    atan(D/k), the two +x diagonals atan(D/t0), the rest 0."""
    c = float(dirs(orc)[1, 0])
    ok = np.ones((H, W), np.uint8)
    t0 = t_diag(k, c)
    a = math.atan(D / k)
    ad = math.atan(D / t0)
    check(orc, step_x(350), ok, 350 - k, 320, [a, ad, 0, 0, 0, 0, 0, ad])


@pytest.mark.parametrize("k", [0, 1, 5, 40])
def test_step_high_side_sees_nothing(orc, k):
    """From the top of the step every elevation is <= 0: every angle is 0
    (a flipped elevation sign gives atan(D/t) here)."""
    ok = np.ones((H, W), np.uint8)
    check(orc, step_x(350), ok, 350 + k, 320, [0.0] * 8)


@pytest.mark.parametrize("k", [1, 4, 33])
def test_step_along_y(orc, k):
    """The step across rows: the buffer's raster rows (y up) are the
    fragments' Y, so +y (a = pi/2) and its diagonals cross it."""
    c = float(dirs(orc)[1, 0])
    depth = np.zeros((H, W), np.float32)
    depth[350:, :] = D
    ok = np.ones((H, W), np.uint8)
    a, ad = math.atan(D / k), math.atan(D / t_diag(k, c))
    check(orc, depth, ok, 320, 350 - k, [0, ad, a, ad, 0, 0, 0, 0])


def test_ramp(orc):
    """depth = g * x: the +x axis gives atan(g) at every tap, the +x
    diagonals atan(g * max_t floor(c t)/t) over t <= 99, everything looking
    downhill 0."""
    c = float(dirs(orc)[1, 0])
    g = 1e-4
    xs = np.arange(W, dtype=np.float64)
    depth = np.tile((g * (xs - 300)).astype(np.float32), (H, 1))
    ok = np.ones((H, W), np.uint8)
    best = max(math.floor(c * t) / t for t in range(1, 100))
    assert abs(best - 70 / 99) < 1e-12  # t = 99: floor(70.0036) = 70
    a, ad = math.atan(g), math.atan(g * best)
    check(orc, depth, ok, 320, 320, [a, ad, 0, 0, 0, 0, 0, ad], rtol=2e-4)


def test_border_stops_the_march(orc):
    """From column 509 of 512 with the step at 510: the +x march sees
    D/1 at t = 1, D/2 at t = 2 and leaves the buffer at t = 3; a march in
    the other order would return at its first tap (outside) with 0."""
    ok = np.ones((H, W), np.uint8)
    ad = math.atan(D / 2)  # the +x diagonals: floor(c t) reaches 1 at t = 2
    check(orc, step_x(510), ok, 509, 320, [math.atan(D / 1), ad, 0, 0, 0, 0, 0, ad])


def test_truncation_toward_zero_at_row_zero(orc):
    """On raster row 0 the direction a = pi has sin(a) < 0 by a few 1e-8:
    its taps' y lies in (-1, 0) and int() truncates it to row 0, inside the
    buffer, so the step 6 pixels to the left (columns 0..39 raised) is seen
    (floor would leave the buffer at t = 1 and see nothing).  The 3pi/4
    diagonal crosses at the first t with ceil(c t) >= 6, t = 8; the 5pi/4
    diagonal reads row 0 at t = 1 (y = -0.7) and stops at t = 2."""
    dd = dirs(orc)
    assert dd[4, 0] == -1.0 and -1e-6 < dd[4, 1] < 0.0, dd[4]
    assert dd[6, 1] == -1.0, dd[6]
    c = float(dd[1, 0])
    depth = np.zeros((H, W), np.float32)
    depth[:, :40] = D
    ok = np.ones((H, W), np.uint8)
    check(orc, depth, ok, 45, 0, [0, 0, 0, math.atan(D / t_diag_neg(6, c)), math.atan(D / 6), 0, 0, 0])


def test_truncation_toward_zero_at_column_zero(orc):
    """The same on column 0 for a = 3pi/2 (cos(a) within 1e-7 of 0): its taps
    stay on column 0 and see the step 6 rows below (rows 0..39 raised)."""
    dd = dirs(orc)
    assert abs(dd[6, 0]) < 1e-6
    c = float(dd[1, 0])
    depth = np.zeros((H, W), np.float32)
    depth[:40, :] = D
    ok = np.ones((H, W), np.uint8)
    check(orc, depth, ok, 0, 45, [0, 0, 0, 0, 0, 0, math.atan(D / 6), math.atan(D / t_diag_neg(6, c))])


def test_uncovered_taps_are_depth_minus_one(orc):
    """A covered pixel at depth -0.5 amid uncovered ones: every tap's
    elevation is -1 - (-0.5) < 0, so every angle is 0 (reading uncovered
    fragments as depth 0 would give atan(0.5/t))."""
    depth = np.zeros((H, W), np.float32)
    ok = np.zeros((H, W), np.uint8)
    depth[320, 320], ok[320, 320] = -0.5, 1
    check(orc, depth, ok, 320, 320, [0.0] * 8)


def test_uncovered_centre_and_truncated_offsets(orc):
    """An uncovered pixel (its depth taken as -1) inside a 3x3 uncovered
    block, everything else covered at depth 0.2, so every first covered tap
    has elevation 1.2.  Axis taps reach the ring at t = 2.  A diagonal tap's
    offset along a positive component is floor(c t), along a negative one
    ceil(c t): the (+,+) diagonal first leaves the block at t = 3 (offsets
    2, 2), the (-,+), (-,-) and (+,-) diagonals at t = 2 (an offset of 2 on
    their negative component)."""
    depth = np.full((H, W), 0.2, np.float32)
    ok = np.ones((H, W), np.uint8)
    ok[319:322, 319:322] = 0
    depth[319:322, 319:322] = 0.0
    e2, e3 = math.atan(1.2 / 2), math.atan(1.2 / 3)
    check(orc, depth, ok, 320, 320, [e2, e3, e2, e2, e2, e2, e2, e2], rtol=1e-4)


def test_flat_buffer_is_finite_and_near_one(orc):
    """Constant depth: every elevation is 0, so every angle is exactly 0 (the
    t = 0 tap is skipped: 0/0 would make Go's math.Max return NaN); the total
    is 8*(pi/2)/(4pi) = 1 up to its float32 rounding."""
    ang, tot, fac = orc.ao_march(np.full((64, 64), 0.3, np.float32), np.ones((64, 64), np.uint8), 30, 30)
    assert np.all(ang == 0)
    assert abs(float(tot) - 1.0) <= 6e-8
    assert abs(float(fac) - float(tot) ** 10000) < 1e-6


# ------------------------------------------------------------ whole frames
def frame_expect_x(orc, s, fr):
    """Closed-form totals of cases.ao_step('x') on rows 75..245."""
    c = float(dirs(orc)[1, 0])
    d = 0.5 * float(np.float32(4e-4))  # the high plane's depth (near -1, far -5 map z = h to h/2)
    x0 = s.width // 2
    exp = np.ones((s.height, s.width))
    for px in range(x0):
        k = x0 - px
        a = math.atan(d / k) if k <= 99 else 0.0
        t0 = t_diag(k, c)
        ad = math.atan(d / t0) if t0 <= 99 else 0.0
        exp[:, px] = total_of([a, ad, ad])
    return exp, (slice(75, 246), slice(0, s.width))


def frame_expect_y(orc, s, fr):
    c = float(dirs(orc)[1, 0])
    d = 0.5 * float(np.float32(4e-4))
    y0 = s.height // 2
    exp = np.ones((s.height, s.width))
    for py in range(y0):
        k = y0 - py
        a = math.atan(d / k) if k <= 99 else 0.0
        t0 = t_diag(k, c)
        ad = math.atan(d / t0) if t0 <= 99 else 0.0
        exp[py, :] = total_of([a, ad, ad])
    return exp, (slice(0, s.height), slice(75, 566))


@pytest.mark.parametrize("axis", ["x", "y"])
def test_ao_step_frame_closed_form(orc, axis):
    """The two-plane scene rendered whole: every checked pixel's AO total
    equals the closed form to a few float32 ulps, and its deferred colour
    u8(pow(total, 10000) * 255) to +-1 LSB.  The 'y' scene has the step
    across rows, so a y-flip between the fragments' Y and the buffer rows
    would put the step on the wrong side."""
    s = cases.ao_step(axis)
    fr = orc.render(s)
    d = float(fr.depth[s.height - 1, s.width - 1])
    assert abs(d - 0.5 * float(np.float32(4e-4))) < 1e-9 and fr.depth[0, 0] == 0.0
    exp, (rs, cs) = (frame_expect_x if axis == "x" else frame_expect_y)(orc, s, fr)
    got = fr.ao_total[rs, cs].astype(np.float64)
    ref = exp[rs, cs]
    err = np.abs(got - ref)
    assert err.max() <= 6e-7, (float(err.max()), np.unravel_index(err.argmax(), err.shape))
    col = np.floor(np.float32(255) * (ref ** 10000).astype(np.float32))
    dcol = np.abs(fr.pregamma[rs, cs, 0].astype(int) - col.astype(int))
    assert dcol.max() <= 1, (int(dcol.max()), np.unravel_index(dcol.argmax(), dcol.shape))
    assert (ref < 1 - 2e-6).sum() > 3000  # a band 25 pixels wide departs from 1 by >3x the tolerance
