// Correctly rounded fp32 division from fp32 operations only, shared by the kernels and the
// CPU checks of the arithmetic (which compile this header with a host compiler and
// SPH_HD defined empty).
//
// Method: Markstein's residual correction (P. Markstein, "Computation of elementary functions
// on the IBM RISC System/6000 processor", IBM J. Res. Develop. 34(1), 1990; restated in
// Muller et al., Handbook of Floating-Point Arithmetic). With y = RN(1/b), the correctly
// rounded reciprocal:
//   q0 = RN(a y)                   relative error below 2^-23 + 2^-24: within 1.5 ulp of a/b
//   q1 = RN(q0 + RN(a - b q0) y)   the error of q0 times (2^-23 + 2^-24) plus one rounding:
//                                  within 1/2 + 2^-21 ulp of a/b, i.e. a faithful quotient
//   q2 = RN(q1 + (a - b q1) y)     a - b q1 is exact in an fma (the remainder of a faithful
//                                  quotient is a float), and by Markstein's theorem (y within
//                                  half an ulp of 1/b, q1 within one ulp of a/b) q2 = RN(a/b).
// Five operations against the eleven of the IEEE divide sequence. The chain runs on |a| and
// the quotient takes the sign of a (b > 0): a zero numerator then gives a zero of a's sign
// whatever a compiler does with the signs of the intermediate zeros.
//
// Preconditions (no underflow in the residuals, no overflow): b > 0 and y = RN(1/b); a = +-0
// or 2^-100 <= |a| <= 2^100; and the quotient within [2^-120, 2^120]. sph_div_rn checks the
// numerator's range and takes the IEEE division for any other a; call sites whose operand
// ranges follow from the parameters call sph_div_mk directly when the host has checked them.
#ifndef SPH_ARITH_H
#define SPH_ARITH_H

#ifndef SPH_HD
#define SPH_HD __host__ __device__
#endif

SPH_HD inline float sph_div_mk(float a, float b, float y) {
  const float m = __builtin_fabsf(a);
  const float q0 = m * y;
  const float r0 = __builtin_fmaf(-b, q0, m);
  const float q1 = __builtin_fmaf(r0, y, q0);
  const float r1 = __builtin_fmaf(-b, q1, m);
  return __builtin_copysignf(__builtin_fmaf(r1, y, q1), a);
}

// the numerator range of sph_div_mk: +-0 or 2^-100 <= |a| <= 2^100 (false for NaN, inf)
SPH_HD inline bool sph_div_mk_ok(float a) {
  const float m = __builtin_fabsf(a);
  return (m >= 0x1p-100f || m == 0.0f) && m <= 0x1p100f;
}

// RN(a / b) for a divisor b > 0 with y = RN(1/b) and b within [2^-20, 2^20] (then every
// numerator that sph_div_mk_ok accepts keeps the quotient within [2^-120, 2^120]); other
// numerators take the IEEE division.
SPH_HD inline float sph_div_rn(float a, float b, float y) {
  return sph_div_mk_ok(a) ? sph_div_mk(a, b, y) : a / b;
}

// ---- host side: the reciprocals and the range check of the force pair term ----

// RN(1/x) for a finite nonzero float x. The quotient of two floats is never a halfway point
// between two floats, and its relative distance from every halfway point exceeds 2^-50: with
// x = X 2^k (integer X < 2^24) and a halfway point m = (2j+1) 2^(e-25),
// 1/x - m = (2^-k - (2j+1) 2^(e-25) X) / X has a numerator that is a nonzero multiple of
// 2^min(-k, e-25). The double quotient is within 2^-53 relative of 1/x, so it rounds to the
// same float as 1/x.
inline float sph_rcp_rn(float x) { return (float)(1.0 / (double)x); }

// Whether (spiky * (h - r) * (h - r)) / k6, the force pair term of a level with kernel
// radius h and k6 = h^6 (SPH-3D-2/sph_system.cu:818, :838), meets the preconditions of
// sph_div_mk for every pair the force kernel evaluates. There r = sqrt(r2) >= sqrt(1e-4) >
// 2^-7 and 0 < h - r <= h. Both h and r are floats of at least min(h, 2^-7), so a positive
// difference is at least the ulp of that, u = 2^(floor(log2 min(h, 2^-7)) - 23); the
// numerator then lies within [|spiky| u^2, |spiky| h^2] up to two roundings (the factor-2
// margins below), and the quotient within that range over k6.
inline bool sph_k6_div_fast(float spiky, float h, float k6) {
  if (!(h > 0.0f) || !(k6 >= 0x1p-126f) || !(k6 <= 0x1p127f) || !(spiky == spiky) ||
      !(__builtin_fabsf(spiky) <= 0x1p127f))
    return false;
  const double m = __builtin_fabs((double)spiky);
  const double hl = (double)h < 0x1p-7 ? (double)h : 0x1p-7;
  const double u = __builtin_ldexp(1.0, __builtin_ilogb(hl) - 23);
  const double a_min = 0.5 * m * u * u, a_max = 2.0 * m * (double)h * (double)h;
  const double q_min = a_min / (double)k6, q_max = a_max / (double)k6;
  return a_min >= 0x1p-100 && a_max <= 0x1p100 && q_min >= 0x1p-120 && q_max <= 0x1p120;
}

#endif
