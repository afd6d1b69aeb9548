// Forward-mode derivative numbers for the oracle -- TEST INFRASTRUCTURE ONLY.
//
// A DualT<V> carries a value v of the oracle's scalar type V (float or double) and its
// float64 derivative d along one direction of the inputs (a learnable field, an initial
// state). Every arithmetic operation applies the chain rule to d and computes v with the
// same V operation as the oracle's V build, in the same order, so the values -- and with
// them every comparison, cell index, wall clamp and split or merge decision, which read v
// only -- are bit-identical to that build's. Running the oracle's step on DualT<V>
// therefore gives the exact derivative (up to float64 rounding of d) of the model along the
// trajectory the V build follows, with the topology it selects:
//   DualT<double>  the float64 model's derivative: the reference against which finite
//                  differences of the float64 oracle are checked;
//   DualT<float>   the derivative along the float32 trajectory, the very states (and
//                  decisions) of the GPU forward: the reference for the GPU adjoint itself.
#ifndef ORC_DUAL_H
#define ORC_DUAL_H

#include <cmath>
#include <type_traits>

template <class V> struct DualT {
  V v = 0;
  double d = 0.0;
  DualT() = default;
  // constants: zero derivative
  template <class S, class = typename std::enable_if<std::is_arithmetic<S>::value>::type>
  DualT(S x) : v((V)x), d(0.0) {}
  DualT(V x, double dx, int) : v(x), d(dx) {}
  explicit operator double() const { return (double)v; }
  explicit operator float() const { return (float)v; }
  explicit operator int() const { return (int)v; }
  explicit operator unsigned() const { return (unsigned)v; }
  DualT &operator+=(const DualT &b) { d += b.d; v += b.v; return *this; }
  DualT &operator-=(const DualT &b) { d -= b.d; v -= b.v; return *this; }
  DualT &operator*=(const DualT &b) { *this = DualT(v * b.v, d * (double)b.v + (double)v * b.d, 0); return *this; }
  DualT &operator/=(const DualT &b) {
    *this = DualT(v / b.v, (d * (double)b.v - (double)v * b.d) / ((double)b.v * (double)b.v), 0);
    return *this;
  }
};

template <class V> inline DualT<V> operator+(const DualT<V> &a, const DualT<V> &b) {
  return DualT<V>(a.v + b.v, a.d + b.d, 0);
}
template <class V> inline DualT<V> operator-(const DualT<V> &a, const DualT<V> &b) {
  return DualT<V>(a.v - b.v, a.d - b.d, 0);
}
template <class V> inline DualT<V> operator-(const DualT<V> &a) { return DualT<V>(-a.v, -a.d, 0); }
template <class V> inline DualT<V> operator*(const DualT<V> &a, const DualT<V> &b) {
  return DualT<V>(a.v * b.v, a.d * (double)b.v + (double)a.v * b.d, 0);
}
template <class V> inline DualT<V> operator/(const DualT<V> &a, const DualT<V> &b) {
  const double bv = (double)b.v;
  return DualT<V>(a.v / b.v, (a.d * bv - (double)a.v * b.d) / (bv * bv), 0);
}
// mixed with plain numbers (constants): converted to the value type first, as the V build
#define ORC_DUAL_MIXED(OP)                                                                     \
  template <class V, class S, class = typename std::enable_if<std::is_arithmetic<S>::value>::type> \
  inline DualT<V> operator OP(const DualT<V> &a, S b) { return a OP DualT<V>(b); }          \
  template <class V, class S, class = typename std::enable_if<std::is_arithmetic<S>::value>::type> \
  inline DualT<V> operator OP(S a, const DualT<V> &b) { return DualT<V>(a) OP b; }
ORC_DUAL_MIXED(+)
ORC_DUAL_MIXED(-)
ORC_DUAL_MIXED(*)
ORC_DUAL_MIXED(/)
#undef ORC_DUAL_MIXED
// comparisons read the values only
#define ORC_DUAL_CMP(OP)                                                                       \
  template <class V> inline bool operator OP(const DualT<V> &a, const DualT<V> &b) { return a.v OP b.v; } \
  template <class V, class S, class = typename std::enable_if<std::is_arithmetic<S>::value>::type> \
  inline bool operator OP(const DualT<V> &a, S b) { return a.v OP DualT<V>(b).v; }          \
  template <class V, class S, class = typename std::enable_if<std::is_arithmetic<S>::value>::type> \
  inline bool operator OP(S a, const DualT<V> &b) { return DualT<V>(a).v OP b.v; }
ORC_DUAL_CMP(<)
ORC_DUAL_CMP(>)
ORC_DUAL_CMP(<=)
ORC_DUAL_CMP(>=)
ORC_DUAL_CMP(==)
ORC_DUAL_CMP(!=)
#undef ORC_DUAL_CMP

// math used by the oracle's step: d sqrt(x) = dx / (2 sqrt(x)); floor is piecewise constant
template <class V> inline DualT<V> sqrt(const DualT<V> &x) {
  const V s = std::sqrt(x.v);
  return DualT<V>(s, x.d == 0.0 ? 0.0 : x.d / (2.0 * (double)s), 0);  // a constant stays one
}
template <class V> inline DualT<V> floor(const DualT<V> &x) { return DualT<V>(std::floor(x.v), 0.0, 0); }

typedef DualT<double> Dual;
typedef DualT<float> DualF;

#endif
