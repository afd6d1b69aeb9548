// CPU checks of the fp32 division helpers in sca2012_amd/csrc/sph_arith.h (the kernels'
// own source, compiled here for the host with SPH_HD empty). Built by
// tests/test_exact_arith.py with g++ -O2 -mfma -ffp-contract=off; the host's fmaf and float
// division are IEEE binary32 operations with one rounding, as the GPU's v_fma_f32 and IEEE
// divide are. Every function returns the number of mismatches against a / b.
#include <cstdint>
#include <cstring>

#define SPH_HD
#include "../sca2012_amd/csrc/sph_arith.h"

namespace {
inline float u2f(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
inline uint32_t f2u(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
inline bool same(float x, float y) {
  if (x != x) return y != y;
  return f2u(x) == f2u(y);
}
struct Rng {  // xorshift64
  uint64_t s;
  uint64_t next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  }
};
float rnd_float(Rng &g, int e_lo, int e_hi) {  // random sign and significand, exponent in [e_lo, e_hi)
  const uint64_t x = g.next();
  const uint32_t e = (uint32_t)(e_lo + 127 + (int)((x >> 32) % (uint64_t)(e_hi - e_lo)));
  return u2f((uint32_t)((x >> 63) << 31) | (e << 23) | (uint32_t)(x & 0x7fffff));
}
}  // namespace

extern "C" {

float chk_rcp_rn(float x) { return sph_rcp_rn(x); }
int chk_k6_div_fast(float spiky, float h, float k6) { return sph_k6_div_fast(spiky, h, k6) ? 1 : 0; }
float chk_div_mk(float a, float b) { return sph_div_mk(a, b, sph_rcp_rn(b)); }
float chk_div_rn(float a, float b) { return sph_div_rn(a, b, sph_rcp_rn(b)); }

// sph_div_mk for every float a of both signs with exponent in [e_lo, e_hi], and +-0, by b
long chk_div_mk_exhaustive(float b, int e_lo, int e_hi) {
  const float y = sph_rcp_rn(b);
  long bad = 0;
  for (int e = e_lo; e <= e_hi; e++)
    for (uint32_t m = 0; m < (1u << 23); m++)
      for (uint32_t s = 0; s < 2; s++) {
        const float a = u2f((s << 31) | ((uint32_t)(e + 127) << 23) | m);
        bad += !same(sph_div_mk(a, b, y), a / b);
      }
  bad += !same(sph_div_mk(0.0f, b, y), 0.0f / b);
  bad += !same(sph_div_mk(-0.0f, b, y), -0.0f / b);
  return bad;
}

// sph_div_rn on random a (exponent in [a_lo, a_hi)) and b > 0 (exponent in [b_lo, b_hi))
long chk_div_rn_random(uint64_t seed, long n, int a_lo, int a_hi, int b_lo, int b_hi) {
  Rng g{seed | 1};
  long bad = 0;
  for (long i = 0; i < n; i++) {
    const float a = rnd_float(g, a_lo, a_hi);
    const float b = __builtin_fabsf(rnd_float(g, b_lo, b_hi));
    bad += !same(sph_div_rn(a, b, sph_rcp_rn(b)), a / b);
  }
  return bad;
}

// numerators next to a halfway point of the quotient: a = RN(m b) for m = an odd 25-bit
// significand (a halfway point between two floats) times 2^k, then a - 3 .. a + 3 ulps;
// b > 0 with exponent in [b_lo, b_hi)
long chk_div_mk_halfway(uint64_t seed, long n, int b_lo, int b_hi) {
  Rng g{seed | 1};
  long bad = 0;
  for (long i = 0; i < n; i++) {
    const float b = __builtin_fabsf(rnd_float(g, b_lo, b_hi));
    const uint64_t x = g.next();
    const double M = (double)((1u << 24) | ((uint32_t)x & 0xffffffu) | 1u);
    const double m = __builtin_ldexp(M, (int)((x >> 32) % 60) - 30 - 24);
    const float a0 = (float)(m * (double)b);
    for (int d = -3; d <= 3; d++) {
      const float a = u2f(f2u(a0) + (uint32_t)d);
      if (!sph_div_mk_ok(a)) continue;
      bad += !same(sph_div_mk(a, b, sph_rcp_rn(b)), a / b);
    }
  }
  return bad;
}

// the force pair term as the kernel forms it, for every (h - r) float in [lo, hi]:
// a = spiky * ikr * ikr, then sph_div_mk(a, k6, RN(1/k6)) against a / k6
long chk_k6_term(float spiky, float k6, float lo, float hi) {
  const float y = sph_rcp_rn(k6);
  long bad = 0;
  for (uint32_t u = f2u(lo); u <= f2u(hi); u++) {
    const float ikr = u2f(u);
    const float a = spiky * ikr * ikr;
    bad += !same(sph_div_mk(a, k6, y), a / k6);
  }
  return bad;
}

// rcp_rn against 1/x for every significand at exponent e (both signs)
long chk_rcp_exhaustive(int e) {
  long bad = 0;
  for (uint32_t m = 0; m < (1u << 23); m++)
    for (uint32_t s = 0; s < 2; s++) {
      const float x = u2f((s << 31) | ((uint32_t)(e + 127) << 23) | m);
      bad += !same(sph_rcp_rn(x), 1.0f / x);
    }
  return bad;
}

}  // extern "C"
