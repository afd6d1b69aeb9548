"""The fp32 arithmetic of the force and energy pair bodies, checked on the CPU.

The kernels divide with Markstein's residual correction (sca2012_amd/csrc/sph_arith.h):
q0 = RN(a y), q1 = RN(q0 + RN(a - b q0) y), q2 = RN(q1 + (a - b q1) y) with y = RN(1/b), five
fp32 operations, in place of the IEEE divide sequence; a numerator outside the checked range
takes the IEEE division. These tests compile that header for the host (g++, the host's fmaf
and float division are single-rounding IEEE binary32 operations like the GPU's) and compare
with a / b:

* every float numerator of whole binades (the method is invariant under scaling a by a power
  of two while nothing under- or overflows, so one binade stands for all of the range; the
  binades at both ends of the accepted range are checked too) for the constant divisors of
  the benchmark configurations (h^6 of each level, h, sqrt h) and divisors next to powers of
  two;
* every numerator the force pair term spiky * (h - r)^2 can take, over every positive float
  h - r the kernel can form, for the reference parameters of both 3D variants and the 2D dam
  break, with the host's range check that enables the fast path;
* random numerators over the whole float range (the fallback included) and divisors over
  [2^-20, 2^20];
* numerators built next to halfway points of the quotient, where a wrong rounding shows first.

The reciprocal RN(1/x) is formed on the host in double; it is checked against 1/x for every
significand. A planted mistake -- the first product q0 = RN(a y) without the corrections -- is
shown to fail the halfway cases, so the checks can see a dropped step.

sph_sqrt_normal (the pair distance): from a hardware estimate within one ulp of sqrt(x), the
two residual tests pick the correctly rounded root for x >= 2^-96.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import sca2012_amd as sa
from sca2012_amd import scenes

f32, f64 = np.float32, np.float64
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def chk(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("arith") / "libexact_arith_check.so")
    src = os.path.join(HERE, "exact_arith_check.cpp")
    subprocess.run(["g++", "-O2", "-mfma", "-ffp-contract=off", "-std=c++17", "-shared", "-fPIC",
                    "-o", so, src], check=True)
    lib = ctypes.CDLL(so)
    F, I, L, U = ctypes.c_float, ctypes.c_int, ctypes.c_long, ctypes.c_uint64
    for name, args, res in [
        ("chk_rcp_rn", [F], F), ("chk_k6_div_fast", [F, F, F], I), ("chk_div_mk", [F, F], F),
        ("chk_div_rn", [F, F], F), ("chk_div_mk_exhaustive", [F, I, I], L),
        ("chk_div_rn_random", [U, L, I, I, I, I], L), ("chk_div_mk_halfway", [U, L, I, I], L),
        ("chk_k6_term", [F, F, F, F], L), ("chk_rcp_exhaustive", [I], L)]:
        fn = getattr(lib, name)
        fn.argtypes, fn.restype = args, res
    return lib


def _divisors():
    out = set()
    for p in (sa.param_defaults(sa.V3D2), sa.param_defaults(sa.V3D4), scenes.dam_break_2d()[0]):
        for lv in ("large", "small"):
            h = f32(p[lv + "_kernel"][0])
            k6 = p[lv + "_kernel_6"][0] if lv + "_kernel_6" in p.dtype.names else p[lv + "_kernel6"][0]
            out.update([float(h), float(np.sqrt(h)), float(f32(k6))])
    for k in range(3):  # next to powers of two, where RN(1/b) errs by nearly half an ulp
        out.add(float(np.uint32((127 << 23) | k).view(f32)))
        out.add(float(np.uint32((127 << 23) | (0x7fffff - k)).view(f32)))
    return sorted(out)


def test_rcp_rn_every_significand(chk):
    for e in (-34, -20, -1, 0, 1, 19):
        assert chk.chk_rcp_exhaustive(e) == 0, e


def test_div_mk_every_numerator_of_whole_binades(chk):
    for b in _divisors():
        assert chk.chk_div_mk_exhaustive(b, 0, 1) == 0, b
        # both ends of the accepted numerator range, when the quotient stays in range
        lb = int(np.floor(np.log2(b)))
        for e in (-100, 99):
            if -119 <= e - lb <= 118:
                assert chk.chk_div_mk_exhaustive(b, e, e) == 0, (b, e)


@pytest.mark.parametrize("variant", ["3d2", "3d4", "2d"])
def test_force_pair_term_every_reachable_numerator(chk, variant):
    p = {"3d2": lambda: sa.param_defaults(sa.V3D2), "3d4": lambda: sa.param_defaults(sa.V3D4),
         "2d": lambda: scenes.dam_break_2d()[0]}[variant]()
    for lv in ("large", "small"):
        h = f32(p[lv + "_kernel"][0])
        spiky = f32(p[lv + "_spiky"][0])
        k6 = f32(p[lv + "_kernel_6"][0] if lv + "_kernel_6" in p.dtype.names else p[lv + "_kernel6"][0])
        assert chk.chk_k6_div_fast(spiky, h, k6) == 1, (variant, lv)
        u = f32(2.0 ** (np.floor(np.log2(min(float(h), 2.0 ** -7))) - 23))
        assert chk.chk_k6_term(spiky, k6, u, h) == 0, (variant, lv)


def test_k6_range_check_rejects_out_of_range_parameters(chk):
    assert chk.chk_k6_div_fast(f32(-14.3), f32(1.0), f32(1.0)) == 1
    assert chk.chk_k6_div_fast(f32(-14.3), f32(1.0), f32(1e-37)) == 0   # quotient overflows
    assert chk.chk_k6_div_fast(f32(-1e-20), f32(1.0), f32(1.0)) == 0    # numerator underflows
    assert chk.chk_k6_div_fast(f32(-1e30), f32(1e5), f32(1e30)) == 0    # numerator overflows
    assert chk.chk_k6_div_fast(f32(np.nan), f32(1.0), f32(1.0)) == 0
    assert chk.chk_k6_div_fast(f32(-14.3), f32(0.0), f32(1.0)) == 0


def test_div_rn_random_full_range(chk):
    assert chk.chk_div_rn_random(11, 4_000_000, -149, 128, -20, 21) == 0
    assert chk.chk_div_rn_random(12, 4_000_000, -110, 110, -20, 21) == 0
    assert chk.chk_div_rn_random(13, 4_000_000, -5, 5, -3, 3) == 0


def test_div_rn_special_values(chk):
    for a in (0.0, -0.0, np.inf, -np.inf, np.nan, 1e-40, -1e-40, 3e38, 1.0):
        for b in (1.0, 0.9, 1.2, 3e-6, 5e5):
            with np.errstate(all="ignore"):
                got, want = chk.chk_div_rn(f32(a), f32(b)), f32(a) / f32(b)
            assert (np.isnan(got) and np.isnan(want)) or \
                np.array_equal(f32(got).view(np.uint32), f32(want).view(np.uint32)), (a, b)
    # signed zeros through the fast path itself
    assert np.signbit(f32(chk.chk_div_mk(f32(-0.0), f32(0.9))))
    assert not np.signbit(f32(chk.chk_div_mk(f32(0.0), f32(0.9))))


def test_div_mk_next_to_halfway_quotients(chk):
    assert chk.chk_div_mk_halfway(21, 3_000_000, -20, 21) == 0
    assert chk.chk_div_mk_halfway(22, 3_000_000, -2, 2) == 0


def test_planted_mistake_uncorrected_product_is_caught():
    """q0 = RN(a RN(1/b)) alone -- the method with both correction steps dropped -- differs
    from a / b on numerators next to halfway points, the cases the checks above generate."""
    rng = np.random.default_rng(5)
    n = 1 << 20
    b = (np.uint32(127 << 23) | rng.integers(0, 1 << 23, n, dtype=np.uint32)).view(f32)
    M = (rng.integers(1 << 24, 1 << 25, n, dtype=np.int64) | 1).astype(f64) * 2.0 ** -24
    a = (M * b.astype(f64)).astype(f32)
    y = (1.0 / b.astype(f64)).astype(f32)
    assert np.count_nonzero((a * y).view(np.uint32) != (a / b).view(np.uint32)) > 1000


def sqrt_normal(x, s):
    """sph_sqrt_normal from the estimate s (float32 arrays); the residual x - s'*s of two
    floats is exact in double, so its sign is what the device's fma result carries."""
    su = s.view(np.uint32)
    sm, sp = (su - np.uint32(1)).view(f32), (su + np.uint32(1)).view(f32)
    rm = x.astype(f64) - sm.astype(f64) * s.astype(f64)
    rp = x.astype(f64) - sp.astype(f64) * s.astype(f64)
    out = np.where(rm <= 0, sm, s)
    return np.where(rp > 0, sp, out)


def random_floats(rng, n, lo_exp, hi_exp):
    """Floats with uniformly random exponent in [lo_exp, hi_exp) and random mantissa/sign."""
    m = rng.integers(0, 1 << 23, n, dtype=np.uint32)
    e = rng.integers(lo_exp + 127, hi_exp + 127, n).astype(np.uint32)
    s = rng.integers(0, 2, n, dtype=np.uint32) << 31
    return (s | (e << 23) | m).view(f32)


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
