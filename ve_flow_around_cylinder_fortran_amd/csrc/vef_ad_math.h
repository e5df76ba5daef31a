// vef_ad_math.h — the element residual's Gauss-point coefficients as a smooth
// function of the interpolated fields, for the analytic (automatic-
// differentiation) Jacobian of the inexact Newton.
//
// The reference's test-node terms (momentumMainSQRT, continuityBulkMain,
// stressesBulkMainSQRT with their LSME/LSIC/LSCE terms, src/Residuals_mod.F90:
// 90-110,161-234,295-475) are linear in the test function: at a Gauss point
// the row (n, q) of the element residual is
//     WET * (phi_n * Tb_q + dphi_n/dx * Tx_q + dphi_n/dy * Ty_q),
// where T = (Tb, Tx, Ty) depends only on the 22 fields interpolated at the
// point (U, grad U, the second derivatives of U, P, grad P, S, grad S) and on
// the arclength unknown EU (WiN = WiN0*EU).  So the element Jacobian is
//     J[(n,q),(m,p)] = sum_g WET_g sum_k phi^n_k(g) sum_{f in p} dT_kq/dF_f(g) * beta^m_f(g)
// with beta^m_f the basis function (value, first or second derivative) of node
// m that field f is interpolated with: the B^T D B form, with D = dT/dF from
// forward-mode differentiation of the function below, one field direction at
// a time (Dual with one derivative component).  The function follows the
// strong form and stabilisation of the residual (setFlowKinematics,
// strongFormConEQ_SQRT, strongFormMomentum, setStabilizationTerms,
// src/Element_Calculations_mod.F90:605-773, src/Constitutive_Models_SQRT_mod.F90:
// 120-509) for DfN > 1e-12 (no discontinuity capturing) in plain algebraic
// form; its values agree with the reference's to rounding, not bit for bit,
// which is what the FP32 Jacobian of an inexact Newton needs.  Squares of
// square roots are formed directly (WiN^2 G:G/2 rather than (WiN sqrt(G:G/2))^2)
// so the derivative exists where G = 0.
#ifndef VEF_AD_MATH_H
#define VEF_AD_MATH_H

#include <hip/hip_runtime.h>

#include <cmath>

// No bitwise target on this path (it is compared with the forward differences
// to FP32 accuracy), so multiply-adds may contract into FMAs here; the pragma
// is switched back off at the end of the header, leaving the FP64 code that
// follows the include as compiled (-ffp-contract=off).
#pragma clang fp contract(fast)

namespace vefad {

template <class T>
struct Dual {
  T v, d;
  __host__ __device__ Dual() : v(0), d(0) {}
  __host__ __device__ Dual(T a) : v(a), d(0) {}
  __host__ __device__ Dual(T a, T b) : v(a), d(b) {}
};
template <class T> __host__ __device__ inline Dual<T> operator+(Dual<T> a, Dual<T> b) { return {a.v + b.v, a.d + b.d}; }
template <class T> __host__ __device__ inline Dual<T> operator-(Dual<T> a, Dual<T> b) { return {a.v - b.v, a.d - b.d}; }
template <class T> __host__ __device__ inline Dual<T> operator-(Dual<T> a) { return {-a.v, -a.d}; }
template <class T> __host__ __device__ inline Dual<T> operator*(Dual<T> a, Dual<T> b) {
  return {a.v * b.v, a.d * b.v + a.v * b.d};
}
template <class T> __host__ __device__ inline Dual<T> operator/(Dual<T> a, Dual<T> b) {
  const T q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
template <class T> __host__ __device__ inline Dual<T> operator+(Dual<T> a, T b) { return {a.v + b, a.d}; }
template <class T> __host__ __device__ inline Dual<T> operator+(T a, Dual<T> b) { return {a + b.v, b.d}; }
template <class T> __host__ __device__ inline Dual<T> operator-(Dual<T> a, T b) { return {a.v - b, a.d}; }
template <class T> __host__ __device__ inline Dual<T> operator-(T a, Dual<T> b) { return {a - b.v, -b.d}; }
template <class T> __host__ __device__ inline Dual<T> operator*(Dual<T> a, T b) { return {a.v * b, a.d * b}; }
template <class T> __host__ __device__ inline Dual<T> operator*(T a, Dual<T> b) { return {a * b.v, a * b.d}; }
template <class T> __host__ __device__ inline Dual<T> operator/(Dual<T> a, T b) { return {a.v / b, a.d / b}; }
template <class T> __host__ __device__ inline Dual<T> operator/(T a, Dual<T> b) {
  const T q = a / b.v;
  return {q, -q * b.d / b.v};
}
template <class T> __host__ __device__ inline Dual<T> dsqrt(Dual<T> a) {
  const T s = sqrt(a.v);
  return {s, a.d / (T(2) * s)};
}
template <class T> __host__ __device__ inline Dual<T> dabs(Dual<T> a) {
  return a.v > T(0) ? a : (a.v < T(0) ? -a : Dual<T>(T(0), T(0)));
}
// N-direction dual numbers: one primal value carried with N tangents, so that
// a thread differentiating in N directions evaluates the primal once.  A
// quotient or square root takes one reciprocal for all N tangents.
template <class T, int N>
struct DualV {
  T v, d[N];
  __host__ __device__ DualV() : v(0) {
    for (int j = 0; j < N; j++) d[j] = T(0);
  }
  __host__ __device__ DualV(T a) : v(a) {
    for (int j = 0; j < N; j++) d[j] = T(0);
  }
};
template <class T, int N> __host__ __device__ inline DualV<T, N> operator+(DualV<T, N> a, DualV<T, N> b) {
  a.v = a.v + b.v;
  for (int j = 0; j < N; j++) a.d[j] = a.d[j] + b.d[j];
  return a;
}
template <class T, int N> __host__ __device__ inline DualV<T, N> operator-(DualV<T, N> a, DualV<T, N> b) {
  a.v = a.v - b.v;
  for (int j = 0; j < N; j++) a.d[j] = a.d[j] - b.d[j];
  return a;
}
template <class T, int N> __host__ __device__ inline DualV<T, N> operator-(DualV<T, N> a) {
  a.v = -a.v;
  for (int j = 0; j < N; j++) a.d[j] = -a.d[j];
  return a;
}
template <class T, int N> __host__ __device__ inline DualV<T, N> operator*(DualV<T, N> a, DualV<T, N> b) {
  DualV<T, N> r;
  r.v = a.v * b.v;
  for (int j = 0; j < N; j++) r.d[j] = a.d[j] * b.v + a.v * b.d[j];
  return r;
}
template <class T, int N> __host__ __device__ inline DualV<T, N> operator/(DualV<T, N> a, DualV<T, N> b) {
  const T ib = T(1) / b.v, q = a.v * ib;
  DualV<T, N> r;
  r.v = q;
  for (int j = 0; j < N; j++) r.d[j] = (a.d[j] - q * b.d[j]) * ib;
  return r;
}
template <class T, int N> __host__ __device__ inline DualV<T, N> operator+(DualV<T, N> a, T b) { a.v = a.v + b; return a; }
template <class T, int N> __host__ __device__ inline DualV<T, N> operator+(T a, DualV<T, N> b) { b.v = a + b.v; return b; }
template <class T, int N> __host__ __device__ inline DualV<T, N> operator-(DualV<T, N> a, T b) { a.v = a.v - b; return a; }
template <class T, int N> __host__ __device__ inline DualV<T, N> operator-(T a, DualV<T, N> b) {
  b = -b;
  b.v = a + b.v;
  return b;
}
template <class T, int N> __host__ __device__ inline DualV<T, N> operator*(DualV<T, N> a, T b) {
  a.v = a.v * b;
  for (int j = 0; j < N; j++) a.d[j] = a.d[j] * b;
  return a;
}
template <class T, int N> __host__ __device__ inline DualV<T, N> operator*(T a, DualV<T, N> b) { return b * a; }
template <class T, int N> __host__ __device__ inline DualV<T, N> operator/(DualV<T, N> a, T b) {
  const T ib = T(1) / b;
  a.v = a.v / b;
  for (int j = 0; j < N; j++) a.d[j] = a.d[j] * ib;
  return a;
}
template <class T, int N> __host__ __device__ inline DualV<T, N> operator/(T a, DualV<T, N> b) {
  const T ib = T(1) / b.v, q = a * ib;
  DualV<T, N> r;
  r.v = q;
  for (int j = 0; j < N; j++) r.d[j] = -q * b.d[j] * ib;
  return r;
}
template <class T, int N> __host__ __device__ inline DualV<T, N> dsqrt(DualV<T, N> a) {
  const T s = sqrt(a.v), h = T(0.5) / s;
  DualV<T, N> r;
  r.v = s;
  for (int j = 0; j < N; j++) r.d[j] = a.d[j] * h;
  return r;
}
template <class T, int N> __host__ __device__ inline DualV<T, N> dabs(DualV<T, N> a) {
  return a.v > T(0) ? a : (a.v < T(0) ? -a : DualV<T, N>(T(0)));
}
__host__ __device__ inline double dsqrt(double a) { return sqrt(a); }
__host__ __device__ inline float dsqrt(float a) { return sqrtf(a); }
__host__ __device__ inline double dabs(double a) { return fabs(a); }
__host__ __device__ inline float dabs(float a) { return fabsf(a); }

// Field order of F[]: the interpolated quantities, then EU.
enum Field {
  kUx, kUy, kGU00, kGU10, kGU01, kGU11, kG2x0, kG2y0, kG2x1, kG2y1,  // GUjk = dU_k/dx_j; G2xk = d2U_k/dx2
  kP, kPx, kPy,
  kS00, kS01, kS11, kS00x, kS00y, kS01x, kS01y, kS11x, kS11y,
  kEU,
  kNumFields
};
constexpr int kNumInterp = kEU;  // 22 interpolated fields

// Unknown q (0 Ux, 1 Uy, 2 P, 3 Sxx, 4 Sxy, 5 Syy) and basis kind (0 value,
// 1 d/dx, 2 d/dy, 3 d2/dx2, 4 d2/dy2) of each interpolated field.
__host__ __device__ constexpr int field_unknown(int f) {
  return f <= kG2y1 ? ((f == kUx || f == kGU00 || f == kGU10 || f == kG2x0 || f == kG2y0) ? 0 : 1)
                    : (f <= kPy ? 2 : (f == kS00 || f == kS00x || f == kS00y ? 3 : (f == kS01 || f == kS01x || f == kS01y ? 4 : 5)));
}
__host__ __device__ constexpr int field_basis(int f) {
  return (f == kUx || f == kUy || f == kP || f == kS00 || f == kS01 || f == kS11) ? 0
         : (f == kGU00 || f == kGU01 || f == kPx || f == kS00x || f == kS01x || f == kS11x) ? 1
         : (f == kGU10 || f == kGU11 || f == kPy || f == kS00y || f == kS01y || f == kS11y) ? 2
         : (f == kG2x0 || f == kG2x1) ? 3 : 4;
}

template <class S>
struct Params {
  S ReN, BvN, VeN, DfN, WiN0, omb;
  int model;  // 0 L-PTT, 1 Giesekus
};

// The six test-node terms of basis kind K (0: phi, 1: dphi/dx, 2: dphi/dy) at
// a Gauss point (momentumMainSQRT, continuityBulkMain, stressesBulkMainSQRT
// with LSME, LSIC, LSCE and DfN, src/Residuals_mod.F90:90-110,161-234,
// 295-475) for phi = (b, fx, fy) = the K-th unit vector: X(ax, ay) stands for
// ax*fx + ay*fy, and every term multiplied by a zero selector is left out.
template <int K, class R, class S>
__host__ __device__ inline void test_terms(R Ux, R Uy, R S00, R S01, R S11, const R* F, R WiN, R f_T, R P00, R P01,
                                           R P11, R CE00, R CE01, R CE10, R CE11, R RC0, R RC1, R Mom0, R Mom1,
                                           R Cont, R tlsme, R tlsic, R tlsce, R kap, R rW, const Params<S>& prm,
                                           R T[6]) {
  static_assert(K >= 0 && K <= 2, "basis kind");
  if constexpr (K == 0) {
    // b = 1: the gradient-weighted parts (LSME, LSIC, the LSCE momentum
    // terms, DfN, the velocity part of the LSCE stress weight) vanish
    T[0] = RC0;
    T[1] = RC1;
    T[2] = Cont;
    const R lsce_s = tlsce * (f_T * S(0.5));
    T[3] = CE00 + lsce_s * CE00;
    T[4] = CE01 + lsce_s * CE01;
    T[5] = CE11 + lsce_s * CE11;
  } else {
    constexpr bool X = K == 1;  // fx = 1 (else fy = 1)
    const R Ud = X ? Ux : Uy;
    const R lsme_m = tlsme * prm.ReN * Ud;
    // momentum x: GUT_lsce row 1 = grad(phi)
    R w = (X ? S01 : S11) / rW;
    R u00 = -(X ? S00 : S01) - w * S01, u01 = -(w * S11);
    R u10 = -(X ? S01 : S11) + w * S00, u11 = w * S01;
    const R l0 = kap * (u00 * CE00 + u01 * CE10 + u10 * CE01 + u11 * CE11);
    // momentum y: GUT_lsce row 2 = grad(phi)
    w = -(X ? S00 : S01) / rW;
    u00 = -(w * S01);
    u01 = -(X ? S00 : S01) - w * S11;
    u10 = w * S00;
    u11 = -(X ? S01 : S11) + w * S01;
    const R l1 = kap * (u00 * CE00 + u01 * CE10 + u10 * CE01 + u11 * CE11);
    const R lsce_s = tlsce * (WiN * Ud);
    const R tw = tlsme * WiN;
    if constexpr (X) {
      T[0] = P00 + lsme_m * Mom0 + tlsic * Cont + l0;
      T[1] = P01 + lsme_m * Mom1 + l1;
      T[2] = tlsme * Mom0;
      T[3] = prm.DfN * F[kS00x] - tw * Mom0 + lsce_s * CE00;
      T[4] = prm.DfN * F[kS01x] - tw * Mom1 + lsce_s * CE01;
      T[5] = prm.DfN * F[kS11x] + lsce_s * CE11;
    } else {
      T[0] = P01 + lsme_m * Mom0 + l0;
      T[1] = P11 + lsme_m * Mom1 + tlsic * Cont + l1;
      T[2] = tlsme * Mom1;
      T[3] = prm.DfN * F[kS00y] + lsce_s * CE00;
      T[4] = prm.DfN * F[kS01y] - tw * Mom0 + lsce_s * CE01;
      T[5] = prm.DfN * F[kS11y] - tw * Mom1 + lsce_s * CE11;
    }
  }
}

// T[k][q]: k = 0 coefficient of phi, 1 of dphi/dx, 2 of dphi/dy; q the row.
// dFx, dFy: the 9 basis gradients at the point (for tsugn), H2 = Helem^2.
template <class R, class S>
__host__ __device__ inline void gauss_terms(const R* F, const S* dFx, const S* dFy, S H2, const Params<S>& prm,
                                            R T[3][6]) {
  const R Ux = F[kUx], Uy = F[kUy];
  const R GU00 = F[kGU00], GU10 = F[kGU10], GU01 = F[kGU01], GU11 = F[kGU11];
  const R S00 = F[kS00], S01 = F[kS01], S11 = F[kS11];
  const R WiN = prm.WiN0 * F[kEU];
  // kinematics: traceless gradient (N = 2), G_dot_tr
  const R htr = (GU00 + GU11) * S(0.5);
  const R Gt00 = GU00 - htr, Gt11 = GU11 - htr;
  const R Gd00 = Gt00 + Gt00, Gd01 = GU01 + GU10, Gd11 = Gt11 + Gt11;
  // C = S.S and its gradient
  const R C00 = S00 * S00 + S01 * S01, C01 = S00 * S01 + S01 * S11, C11 = S01 * S01 + S11 * S11;
  const R dC00x = S(2) * (F[kS00x] * S00 + F[kS01x] * S01);
  const R dC10y = F[kS01y] * S00 + F[kS11y] * S01 + S01 * F[kS00y] + S11 * F[kS01y];
  const R dC01x = F[kS00x] * S01 + F[kS01x] * S11 + S00 * F[kS01x] + S01 * F[kS11x];
  const R dC11y = S(2) * (F[kS01y] * S01 + F[kS11y] * S11);
  R f0, f1, f2, f_T;
  if (prm.model == 0) {
    f_T = S(1) + prm.VeN * ((C00 + C11) - S(2));
    f0 = f_T;
    f1 = f_T;
    f2 = R(S(0));
  } else {
    f_T = R(S(1) - S(2) * prm.VeN);
    f0 = R(S(1) - prm.VeN);
    f1 = R(S(1) - S(2) * prm.VeN);
    f2 = R(prm.VeN);
  }
  const R a = prm.omb / WiN;
  const R T00 = a * (C00 - S(1)), T01 = a * C01, T11 = a * (C11 - S(1));
  const R P00 = T00 + prm.BvN * Gd00 - F[kP];
  const R P01 = T01 + prm.BvN * Gd01;
  const R P11 = T11 + prm.BvN * Gd11 - F[kP];
  // upper-convected derivative of S
  const R rW = S00 + S11;
  const R W = (-(S00 * GU01) + S01 * (Gt00 - Gt11) + S11 * GU10) / rW;  // omega(1,2) from GUT_tr
  const R U00 = (Ux * F[kS00x] + Uy * F[kS00y]) - (S00 * Gt00 + S01 * GU10) - W * S01;
  const R U01 = (Ux * F[kS01x] + Uy * F[kS01y]) - (S00 * GU01 + S01 * Gt11) - W * S11;
  const R U10 = (Ux * F[kS01x] + Uy * F[kS01y]) - (S01 * Gt00 + S11 * GU10) + W * S00;
  const R U11 = (Ux * F[kS11x] + Uy * F[kS11y]) - (S01 * GU01 + S11 * Gt11) + W * S01;
  const R det = S00 * S11 - S01 * S01;
  const R SI00 = S11 / det, SI01 = -S01 / det, SI11 = S00 / det;
  R t00 = f1 * S00 - f0 * SI00, t01 = f1 * S01 - f0 * SI01, t11 = f1 * S11 - f0 * SI11;
  if (prm.model != 0) {
    t00 = t00 + f2 * (S00 * C00 + S01 * C01);
    t01 = t01 + f2 * (S00 * C01 + S01 * C11);
    t11 = t11 + f2 * (S01 * C01 + S11 * C11);
  }
  const R CE00 = WiN * U00 + t00 * S(0.5), CE01 = WiN * U01 + t01 * S(0.5);
  const R CE10 = WiN * U10 + t01 * S(0.5), CE11 = WiN * U11 + t11 * S(0.5);
  // momentum and continuity
  const R dP0 = -F[kPx] + a * (dC00x + dC10y) + prm.BvN * (F[kG2x0] + F[kG2y0]);
  const R dP1 = -F[kPy] + a * (dC01x + dC11y) + prm.BvN * (F[kG2x1] + F[kG2y1]);
  const R RC0 = prm.ReN * (Ux * GU00 + Uy * GU10), RC1 = prm.ReN * (Ux * GU01 + Uy * GU11);
  const R Mom0 = RC0 - dP0, Mom1 = RC1 - dP1;
  const R Cont = GU00 + GU11;
  // stabilisation (tau's)
  R tsugn = dabs(Ux * dFx[0] + Uy * dFy[0]);
  for (int n = 1; n < 9; n++) tsugn = tsugn + dabs(Ux * dFx[n] + Uy * dFy[n]);
  const R ddT = T00 * T00 + S(2) * (T01 * T01) + T11 * T11 + a * a;
  const R ddG = Gd00 * Gd00 + S(2) * (Gd01 * Gd01) + Gd11 * Gd11;
  const R Ha = dsqrt(a * a + S(0.5) * ddT) / dsqrt(S(1) + S(0.5) * ddG);
  const R r2 = prm.ReN * tsugn, r4 = Ha / H2;
  const S r3 = prm.BvN / H2;
  const R tlsme = S(1) / dsqrt(r2 * r2 + r3 * r3 + r4 * r4);
  const R tlsic = H2 / tlsme;
  const R ddC = C00 * C00 + S(2) * (C01 * C01) + C11 * C11;
  const S t7 = prm.DfN / H2;
  const R tlsce = S(1) / dsqrt(WiN * WiN * tsugn * tsugn + WiN * WiN * (S(0.5) * ddG) + f1 * f1 +
                               f2 * f2 * (S(0.5) * ddC) + t7 * t7);
  const R kap = tlsce * prm.omb / WiN;  // LSCEMomentumSQRT: tlsce*UCD_lsce*(1-BvN)/WiN
  // test-node terms for phi = (1, 0, 0), (0, 1, 0), (0, 0, 1): the reference's
  // terms with the basis selectors b = phi, fx = dphi/dx, fy = dphi/dy set to
  // their 0/1 values at compile time, so the products by zero (which IEEE
  // arithmetic may not drop by itself) are not formed
  test_terms<0>(Ux, Uy, S00, S01, S11, F, WiN, f_T, P00, P01, P11, CE00, CE01, CE10, CE11, RC0, RC1, Mom0, Mom1,
                Cont, tlsme, tlsic, tlsce, kap, rW, prm, T[0]);
  test_terms<1>(Ux, Uy, S00, S01, S11, F, WiN, f_T, P00, P01, P11, CE00, CE01, CE10, CE11, RC0, RC1, Mom0, Mom1,
                Cont, tlsme, tlsic, tlsce, kap, rW, prm, T[1]);
  test_terms<2>(Ux, Uy, S00, S01, S11, F, WiN, f_T, P00, P01, P11, CE00, CE01, CE10, CE11, RC0, RC1, Mom0, Mom1,
                Cont, tlsme, tlsic, tlsce, kap, rW, prm, T[2]);
}

}  // namespace vefad

#pragma clang fp contract(off)

#endif
