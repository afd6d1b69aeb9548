"""The arithmetic shortcuts of the force pair body, checked on the CPU against IEEE float
arithmetic (numpy float32 division and square root are correctly rounded):

* sph_div_by: float(double(a) * RN64(1/c)) == RN32(a / c) for every float a and finite
  nonzero c. Random operands over the whole exponent range, subnormal and overflowing
  quotients, the divisors the force kernel uses (k6 per level, pair distances), and
  quotients built to lie as close to a halfway point as floats allow.
* sph_sqrt_normal: from a hardware estimate within one ulp of sqrt(x), the two residual
  tests pick the correctly rounded root for x >= 2^-96. The residual x - s'*s of two floats
  is exact in double here (its sign is what the device's fma result carries).
* sph_rcp_for_div: two double Newton steps from the hardware float reciprocal give a 1/r
  good enough for sph_div_by (exact fma emulated with rationals)."""
import numpy as np

f32, f64 = np.float32, np.float64


def div_by(a, c):
    inv = f64(1.0) / c.astype(f64)
    return (a.astype(f64) * inv).astype(f32)


def check_div(a, c):
    with np.errstate(all="ignore"):
        want = a / c
        got = div_by(a, c)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    bad = np.nonzero(got[~nan].view(np.uint32) != want[~nan].view(np.uint32))[0]
    assert len(bad) == 0, (a[~nan][bad[:5]], c[~nan][bad[:5]])


def random_floats(rng, n, lo_exp, hi_exp):
    """Floats with uniformly random exponent in [lo_exp, hi_exp) and random mantissa/sign."""
    m = rng.integers(0, 1 << 23, n, dtype=np.uint32)
    e = rng.integers(lo_exp + 127, hi_exp + 127, n).astype(np.uint32)
    s = rng.integers(0, 2, n, dtype=np.uint32) << 31
    return (s | (e << 23) | m).view(f32)


def test_div_by_random_full_range():
    rng = np.random.default_rng(11)
    for _ in range(8):
        a = random_floats(rng, 1 << 21, -126, 128)
        c = random_floats(rng, 1 << 21, -126, 128)
        check_div(a, c)


def test_div_by_subnormal_and_overflow():
    rng = np.random.default_rng(12)
    n = 1 << 20
    sub = (rng.integers(1, 1 << 23, n, dtype=np.uint32)).view(f32)  # subnormal a
    check_div(sub, random_floats(rng, n, -10, 10))
    check_div(random_floats(rng, n, -126, -100), random_floats(rng, n, 0, 30))  # subnormal a/c
    check_div(random_floats(rng, n, 100, 128), random_floats(rng, n, -30, 0))   # a/c overflows
    check_div(np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0], f32),
              np.array([3.0, 3.0, 3.0, -3.0, 3.0, 3.0], f32))


def test_div_by_force_kernel_divisors():
    """The force pair terms: (spi*ikr*ikr)/k6 with k6 = h^6 of either level (h = 1, 0.5 and
    a range of other kernel radii), and (rx*tp)/r with 0.01 <= r < 2."""
    rng = np.random.default_rng(13)
    n = 1 << 21
    for h in (1.0, 0.5, 0.7937005, 0.63, 1.25, 0.04):
        k6 = f32(h) ** 6
        a = random_floats(rng, n, -40, 20)
        check_div(a, np.full(n, k6, f32))
    r = np.sqrt(rng.uniform(1e-4, 4.0, n).astype(f32))
    check_div(random_floats(rng, n, -40, 20), r)


def test_div_by_near_halfway_quotients():
    """a = RN32(m * c) for m a halfway point (25 significant bits): the quotient lands
    within a few float ulps of m, where a wrong rounding would show first."""
    rng = np.random.default_rng(14)
    n = 1 << 21
    c = random_floats(rng, n, -20, 20)
    c = np.abs(c)
    M = (rng.integers(1 << 24, 1 << 25, n, dtype=np.int64) | 1).astype(f64)  # odd 25-bit
    m = M * np.exp2(rng.integers(-30, 30, n).astype(f64) - 24)
    a = (m * c.astype(f64)).astype(f32)
    for da in (-2, -1, 0, 1, 2):
        check_div((a.view(np.uint32) + np.uint32(da & 0xffffffff)).view(f32), c)


def sqrt_normal(x, s):
    """The device routine from estimate s (float32 arrays)."""
    su = s.view(np.uint32)
    sm, sp = (su - np.uint32(1)).view(f32), (su + np.uint32(1)).view(f32)
    rm = x.astype(f64) - sm.astype(f64) * s.astype(f64)
    rp = x.astype(f64) - sp.astype(f64) * s.astype(f64)
    out = np.where(rm <= 0, sm, s)
    return np.where(rp > 0, sp, out)


def test_sqrt_correction_from_estimates_one_ulp_off():
    rng = np.random.default_rng(15)
    for lo, hi in ((-96, -60), (-14, 2), (2, 127)):
        x = np.abs(random_floats(rng, 1 << 21, lo, hi))
        want = np.sqrt(x)
        wu = want.view(np.uint32)
        for d in (-1, 0, 1):
            s = (wu + np.uint32(d & 0xffffffff)).view(f32)
            got = sqrt_normal(x, s)
            bad = np.nonzero(got.view(np.uint32) != wu)[0]
            assert len(bad) == 0, (d, x[bad[:5]])
    # the pair-body range: 1e-4 <= r2 < 4
    x = rng.uniform(1e-4, 4.0, 1 << 21).astype(f32)
    want = np.sqrt(x)
    for d in (-1, 0, 1):
        s = (want.view(np.uint32) + np.uint32(d & 0xffffffff)).view(f32)
        assert np.array_equal(sqrt_normal(x, s).view(np.uint32), want.view(np.uint32))


def fma64(x, y, z):
    """Correctly rounded double fma (exact rational arithmetic, one rounding)."""
    from fractions import Fraction
    v = Fraction(x) * Fraction(y) + Fraction(z)
    return float(v)  # int / int true division: correctly rounded in CPython


def rcp_for_div(r, y0):
    """sph_rcp_for_div from the float estimate y0."""
    rd = float(r)
    y0 = float(y0)
    y1 = fma64(y0, fma64(-rd, y0, 1.0), y0)
    return fma64(y1, fma64(-rd, y1, 1.0), y1)


def test_reciprocal_newton_steps_from_estimates_two_ulps_off():
    """1/r by two double Newton steps from a float estimate up to two ulps (and 2^-14
    relative) off: within 2^-52.5 relative of 1/r, and every quotient through it exact."""
    from fractions import Fraction
    rng = np.random.default_rng(16)
    rs = np.concatenate([np.sqrt(rng.uniform(1e-4, 4.0, 300)).astype(f32),
                         np.abs(random_floats(rng, 100, -126, 127))])
    a = random_floats(rng, 4096, -60, 60)
    for r in rs:
        e = int((f32(1.0) / r).view(np.uint32))
        ests = [np.uint32(e + d).view(f32) for d in (-2, -1, 0, 1, 2)]
        ests.append(f32(float(f32(1.0) / r) * (1 + 2.0 ** -14)))
        for y0 in ests:
            y = rcp_for_div(r, y0)
            err = abs(Fraction(y) * Fraction(float(r)) - 1)
            assert err < Fraction(1, 2 ** 52) * Fraction(7, 10), (r, y0, float(err))
            with np.errstate(all="ignore"):
                got = (a.astype(f64) * y).astype(f32)
                want = a / np.full_like(a, r)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), r
