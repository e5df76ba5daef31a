// TEST INFRASTRUCTURE ONLY (oracle/): an instrumented scalar used to count the
// algorithmic floating-point operations of one element-residual evaluation.
//
// FlopReal carries a "structural zero" flag for values that are zero by
// construction (zero-initialised tensor slots of the 3D-shaped arrays, the
// third spatial dimension of a 2D problem, accumulators before their first
// term).  Operations with a structural-zero operand are not counted and keep
// the flag where the result is structurally zero, so the count is that of the
// 2D algorithm with the reference's padding work removed.  Every other +, -,
// *, / and sqrt counts one operation; negation and abs count none.
#pragma once
#include <cmath>
#include <cstdint>

struct FlopCounter {
  static inline uint64_t add = 0, mul = 0, div = 0, sqrt = 0;
  static void reset() { add = mul = div = sqrt = 0; }
  static uint64_t total() { return add + mul + div + sqrt; }
};

struct FlopReal {
  double v;
  bool z;  // structural zero
  FlopReal() : v(0.0), z(true) {}
  FlopReal(double x) : v(x), z(false) {}
  static FlopReal zero() { return FlopReal(); }
  FlopReal(double x, bool zz) : v(x), z(zz) {}
};

inline FlopReal operator+(FlopReal a, FlopReal b) {
  if (a.z && b.z) return FlopReal(a.v + b.v, true);
  if (a.z) return FlopReal(a.v + b.v, false);
  if (b.z) return FlopReal(a.v + b.v, false);
  FlopCounter::add++;
  return FlopReal(a.v + b.v, false);
}
inline FlopReal operator-(FlopReal a, FlopReal b) {
  if (a.z && b.z) return FlopReal(a.v - b.v, true);
  if (a.z || b.z) return FlopReal(a.v - b.v, false);
  FlopCounter::add++;
  return FlopReal(a.v - b.v, false);
}
inline FlopReal operator-(FlopReal a) { return FlopReal(-a.v, a.z); }
inline FlopReal operator*(FlopReal a, FlopReal b) {
  if (a.z || b.z) return FlopReal(a.v * b.v, true);
  FlopCounter::mul++;
  return FlopReal(a.v * b.v, false);
}
inline FlopReal operator/(FlopReal a, FlopReal b) {
  if (a.z) return FlopReal(a.v / b.v, true);
  FlopCounter::div++;
  return FlopReal(a.v / b.v, false);
}
inline FlopReal& operator+=(FlopReal& a, FlopReal b) { a = a + b; return a; }
inline FlopReal& operator-=(FlopReal& a, FlopReal b) { a = a - b; return a; }
inline bool operator<(FlopReal a, FlopReal b) { return a.v < b.v; }
inline bool operator>(FlopReal a, FlopReal b) { return a.v > b.v; }
inline bool operator>=(FlopReal a, FlopReal b) { return a.v >= b.v; }
inline FlopReal vsqrt(FlopReal a) {
  if (a.z) return FlopReal(std::sqrt(a.v), true);
  FlopCounter::sqrt++;
  return FlopReal(std::sqrt(a.v), false);
}
inline FlopReal fabsr(FlopReal a) { return FlopReal(std::fabs(a.v), a.z); }
inline double val(FlopReal a) { return a.v; }
