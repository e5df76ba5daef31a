// vef_assembly.hip — MI355X (gfx950) residual and forward-difference Jacobian
// assembly for the 2D square-root-conformation viscoelastic cylinder problem,
// behind the C ABI of include/vef.h.
//
// What runs where (one assembly, in stream order)
//   vef_bulk_kernel<7,1,1>  residual pass: one wavefront per 7 elements, lane =
//                           (Gauss point, element); element residual rows R_0
//                           into r0bank[e][54].
//   vef_edge_kernel         outlet-edge traction (openBC), 3 edge Gauss points
//                           x 56 evaluations per outlet edge, into ebank.
//   vef_bulk_kernel<1,7,56> Jacobian pass, one launch per element colour:
//                           eight single-wavefront workgroups per element, 63
//                           lanes = 9 Gauss points x 7 of the 54 nodal
//                           perturbations and the EU perturbation.  Each lane
//                           evaluates the Gauss-point strong form and the 9
//                           test-node rows; an ordered LDS reduction over the
//                           Gauss points gives each evaluation's 54 rows, and
//                           the quotients (R_k - R_0)/h go straight to the CSR
//                           values when the node pair belongs to this element
//                           only, are stored (first colour) or added (second)
//                           when it belongs to two elements, and go to cbank
//                           when it belongs to more; dR/dEU goes to acbank.
//   vef_shared_kernel       values of node pairs of three or more elements:
//                           their cbank entries summed in ascending element
//                           order (the order of the reference's ADD_VALUES).
//   vef_edge_add_kernel     outlet-edge quotients added after the bulk ones,
//                           in ascending edge order.
//   vef_rhs_kernel          b and the EU column: per node row, elements then
//                           edges in ascending order.
// Every sum starts from +0 as ADD_VALUES into zeroed storage does, so even
// signed zeros match.
//
// The arithmetic of each lane restates src/Element_Calculations_mod.F90,
// src/Constitutive_Models_SQRT_mod.F90 and src/Residuals_mod.F90 in 2D form:
// the reference's 3x3 tensors have a zero third row and column here, and the
// terms they contribute are exact zeros whose addition leaves every value
// unchanged, so they are dropped.  The order of every remaining operation is
// the reference's (left-to-right sums, sequential dot products, complex
// divisions with zero imaginary part rounded as (a*c)/(c*c)), and the file is
// compiled with -ffp-contract=off, so the FP64 results equal the sequential
// reference evaluation bit for bit.
#include <hip/hip_runtime.h>
#include <type_traits>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "vef.h"
#include "vef_stability_math.h"
#include "vef_ad_math.h"

namespace {

constexpr double kEpsJac = 1.0e-8;  // perturbVariable, src/Tools_mod.F90:831-840
constexpr int kNeval = 56;          // 1 base + 54 nodal + 1 EU evaluations

// ----------------------------------------------------------------------------
// Gauss tables (src/GaussIntegration/Gauss_Quadrangle_mod.F90), host-built.
// ----------------------------------------------------------------------------
struct GaussTables {
  double W2[9];
  double W1[3];
  double bf[9][9], dc[9][9], de[9][9], dcc[9][9], dee[9][9], dce[9][9];  // [gp][node]
  double bfE[4][3][9], dcE[4][3][9], deE[4][3][9];                       // [edge][gp][node]
};
__constant__ GaussTables c_tab;

// setQuadraticQuadrangle, Gauss_Quadrangle_mod.F90:405-478: Lagrange products
// of the 1D quadratics, evaluated in the source's operation order.
void q2_basis(double c, double e, double* bfn, double* dfdc, double* dfde, double* d2fdc2, double* d2fde2,
              double* d2fdce) {
  const double a = 1.0 - c, b = 1.0 + c, p = 1.0 - e, q = 1.0 + e;
  bfn[0] = a * c * p * e / 4.0;
  bfn[1] = -(b * c * p * e / 4.0);
  bfn[2] = b * c * q * e / 4.0;
  bfn[3] = -(a * c * q * e / 4.0);
  bfn[4] = -(b * a * p * e / 2.0);
  bfn[5] = b * c * q * p / 2.0;
  bfn[6] = b * a * q * e / 2.0;
  bfn[7] = -(a * c * q * p / 2.0);
  bfn[8] = b * a * q * p;
  const double c2m = 1.0 - 2.0 * c, c2p = 1.0 + 2.0 * c, cm2 = -2.0 * c;
  const double e2m = 1.0 - 2.0 * e, e2p = 1.0 + 2.0 * e, em2 = -2.0 * e;
  dfdc[0] = c2m * p * e / 4.0;
  dfdc[1] = -(c2p * p * e / 4.0);
  dfdc[2] = c2p * q * e / 4.0;
  dfdc[3] = -(c2m * q * e / 4.0);
  dfdc[4] = -(cm2 * p * e / 2.0);
  dfdc[5] = c2p * q * p / 2.0;
  dfdc[6] = cm2 * q * e / 2.0;
  dfdc[7] = -(c2m * q * p / 2.0);
  dfdc[8] = cm2 * q * p;
  d2fdc2[0] = -2.0 * p * e / 4.0;
  d2fdc2[1] = -(2.0 * p * e / 4.0);
  d2fdc2[2] = 2.0 * q * e / 4.0;
  d2fdc2[3] = -(-2.0 * q * e / 4.0);
  d2fdc2[4] = -(-2.0 * p * e / 2.0);
  d2fdc2[5] = 2.0 * q * p / 2.0;
  d2fdc2[6] = -2.0 * q * e / 2.0;
  d2fdc2[7] = -(-2.0 * q * p / 2.0);
  d2fdc2[8] = -2.0 * q * p;
  d2fdce[0] = c2m * e2m / 4.0;
  d2fdce[1] = -(c2p * e2m / 4.0);
  d2fdce[2] = c2p * e2p / 4.0;
  d2fdce[3] = -(c2m * e2p / 4.0);
  d2fdce[4] = -(cm2 * e2m / 2.0);
  d2fdce[5] = c2p * em2 / 2.0;
  d2fdce[6] = cm2 * e2p / 2.0;
  d2fdce[7] = -(c2m * em2 / 2.0);
  d2fdce[8] = cm2 * em2;
  dfde[0] = a * c * e2m / 4.0;
  dfde[1] = -(b * c * e2m / 4.0);
  dfde[2] = b * c * e2p / 4.0;
  dfde[3] = -(a * c * e2p / 4.0);
  dfde[4] = -(b * a * e2m / 2.0);
  dfde[5] = b * c * em2 / 2.0;
  dfde[6] = b * a * e2p / 2.0;
  dfde[7] = -(a * c * em2 / 2.0);
  dfde[8] = b * a * em2;
  d2fde2[0] = a * c * -2.0 / 4.0;
  d2fde2[1] = -(b * c * -2.0 / 4.0);
  d2fde2[2] = b * c * 2.0 / 4.0;
  d2fde2[3] = -(a * c * 2.0 / 4.0);
  d2fde2[4] = -(b * a * -2.0 / 2.0);
  d2fde2[5] = b * c * -2.0 / 2.0;
  d2fde2[6] = b * a * 2.0 / 2.0;
  d2fde2[7] = -(a * c * -2.0 / 2.0);
  d2fde2[8] = b * a * -2.0;
}

void build_gauss_tables(GaussTables& T) {
  // 3-point rule with the centre abscissa shifted by 1e-8 (Gauss_Quadrangle_mod.F90:69-72,166-169)
  const double P1[3] = {0.0 + 1.0e-8, +std::sqrt(3.0 / 5.0), -std::sqrt(3.0 / 5.0)};
  const double W1[3] = {8.0 / 9.0, 5.0 / 9.0, 5.0 / 9.0};
  for (int i = 0; i < 3; i++) T.W1[i] = W1[i];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      int g = i * 3 + j;  // xi outer, eta inner (:93-99)
      T.W2[g] = W1[i] * W1[j];
      q2_basis(P1[i], P1[j], T.bf[g], T.dc[g], T.de[g], T.dcc[g], T.dee[g], T.dce[g]);
    }
  for (int ed = 0; ed < 4; ed++)
    for (int g = 0; g < 3; g++) {
      double c = (ed == 1) ? 1.0 : (ed == 3) ? -1.0 : P1[g];
      double e = (ed == 0) ? -1.0 : (ed == 2) ? 1.0 : P1[g];
      double t1[9], t2[9], t3[9];  // second derivatives: not used on edges
      q2_basis(c, e, T.bfE[ed][g], T.dcE[ed][g], T.deE[ed][g], t1, t2, t3);
    }
}

struct DevParams {
  double ReN, BvN, VeN, DfN, WiN0;
  double dt;    // +huge(1d0) off the transient path (src/Physical_mod.F90:345)
  double omb;   // 1 - BvN
  int model;    // 0 L-PTT, 1 Giesekus
  int dcs_on;   // DfN <= 1e-12
};

// Geometry of one element at one Gauss point (basis2D, Element_Calculations_mod.F90:308-361;
// Helem from setStabilizationTerms :718-727).
struct Geo {
  double dFx[9], dFy[9], d2x[9], d2y[9];
  double WET, H2;
  double H4, yH4;     // H2*H2 and RN(1/(H2*H2)) (denominator of the complex division by H2)
  double r3, t7;      // BvN/H2 and DfN/H2
  int in_range;       // every dFx, dFy is 0 or has magnitude in [2^-64, 2^64)
};

// Complex division with zero imaginary parts: re((a+0i)/(c+0i)) = (a*c)/(c*c).
__device__ __forceinline__ double cdiv(double a, double c) { return (a * c) / (c * c); }

// Division by a denominator shared by many numerators, in two modes.
//
// Fast (EXACT = false): with y = RN(1/d), q0 = RN(a*y), e = RN(d*q0 - a) and
// q1 = RN(q0 - e*y) (two FMAs), q1 is the correctly rounded a/d: the
// correction term q0 - (d*q0 - a)*y differs from a/d by |1 - d*y|*|a/d - q0|,
// far below the distance of any quotient of two doubles from a rounding
// midpoint (Markstein's correction step; checked exhaustively over all pairs
// of significands at precisions 4..17).  The argument is scale-free, so it
// holds while every quantity stays in the normal range with room for the
// residual: d in [2^-60, 2^40] and |q0| in [2^-901, 2^900), or a = +-0 (then
// q0 = +-0, e = +0 and q1 = q0 keeps the sign of a/d because d > 0).  Instead
// of branching, each division widens the exponent interval [lo, hi] of its q0
// in a DivCheck, and a denominator outside its range sets hi to kBadExp.  The
// frexp exponent of an infinity or a NaN is 0, which the interval accepts, so
// the caller also tests the quotients for finiteness (gauss_point: one test of
// their sum).  The caller repeats the whole evaluation in the exact mode when
// the interval or that test says so.
//
// Exact (EXACT = true): the IEEE division a/d.
struct Recip {
  double d, y;
};
struct DivCheck {
  int lo, hi;    // binades of the tracked quotients q0: must stay in [-900, 900]
  int blo, bhi;  // binades of the base quantities: must stay in [-63, 64]
  __device__ __forceinline__ bool failed() const { return lo < -900 || hi > 900 || blo < -63 || bhi > 64; }
};
constexpr int kBadExp = 4096;

// The binade of x (0 for +-0), or kBadExp for an infinity or a NaN, whose
// frexp exponent (0) would pass every range check.
__device__ __forceinline__ int binade(double x) {
  return __builtin_isfinite(x) ? __builtin_amdgcn_frexp_exp(x) : kBadExp;
}

// Base quantities of the untracked divisions: x = 0 or 2^-64 <= |x| < 2^64
// (finiteness tested by the caller).
template <bool EXACT>
__device__ __forceinline__ void base_range(double x, DivCheck& dc) {
  if constexpr (!EXACT) {
    const int ex = __builtin_amdgcn_frexp_exp(x);
    dc.blo = min(dc.blo, ex);
    dc.bhi = max(dc.bhi, ex);
  }
}

template <bool EXACT>
__device__ __forceinline__ void check_denominator(double d, DivCheck& dc) {
  if constexpr (!EXACT) {
    if (!(d >= 0x1p-60 && d <= 0x1p40)) dc.hi = kBadExp;
  }
}
template <bool EXACT>
__device__ __forceinline__ Recip make_recip(double d, DivCheck& dc) {
  if constexpr (EXACT) return Recip{d, 0.0};
  check_denominator<EXACT>(d, dc);
  return Recip{d, 1.0 / d};
}
template <bool EXACT>
__device__ __forceinline__ double qdiv(double a, const Recip& r, DivCheck& dc) {
  if constexpr (EXACT) {
    return a / r.d;
  } else {
    const double q0 = a * r.y;
    const double e = fma(r.d, q0, -a);
    const int ex = __builtin_amdgcn_frexp_exp(q0);  // 0 for +-0 (and for inf/NaN: tested by the caller)
    dc.lo = min(dc.lo, ex);
    dc.hi = max(dc.hi, ex);
    return fma(-e, r.y, q0);
  }
}

// Fast division whose window is guaranteed by the ranges of its inputs rather
// than tracked (the test-node divisions of node_terms; the bound is derived
// there).
template <bool EXACT>
__device__ __forceinline__ double qdiv_ranged(double a, const Recip& r) {
  if constexpr (EXACT) {
    return a / r.d;
  } else {
    const double q0 = a * r.y;
    const double e = fma(r.d, q0, -a);
    return fma(-e, r.y, q0);
  }
}

template <bool EXACT>
__device__ __forceinline__ double fd_quotient(double rk, double r0, DivCheck& dc) {
  // (dRes - Res)/eps_jac with a complex numerator: ((dR-R)*h)/(h*h)
  constexpr double hh = kEpsJac * kEpsJac;
  return qdiv<EXACT>((rk - r0) * kEpsJac, Recip{hh, 1.0 / hh}, dc);
}
// The same quotient decided per entry: the two-FMA division where its q0 lies
// in the window, the IEEE division otherwise.  The IEEE division sits behind
// a branch on the ballot of the active lanes, uniform among them, with an
// empty volatile statement in it: without one the compiler computes the
// division for every entry and selects (about 11 more FP64 instructions per
// quotient).
template <bool EXACT>
__device__ __forceinline__ double fd_quotient_entry(double rk, double r0) {
  constexpr double hh = kEpsJac * kEpsJac;
  const double a = (rk - r0) * kEpsJac;
  if constexpr (EXACT) {
    return a / hh;
  } else {
    constexpr double y = 1.0 / hh;
    const double q0 = a * y;
    const double e = fma(hh, q0, -a);
    double q = fma(-e, y, q0);
    const int ex = binade(q0);
    const bool outside = ex < -900 || ex > 900;
    if (__builtin_amdgcn_ballot_w64(outside) != 0) {
      asm volatile("" ::: "memory");  // keeps the block a branch: not speculated
      if (outside) q = a / hh;
    }
    return q;
  }
}
__device__ __forceinline__ double fd_quotient_ieee(double rk, double r0) {
  DivCheck dc{0, 0, 0, 0};
  return fd_quotient<true>(rk, r0, dc);
}


// basis2D for one Gauss point from the nine node coordinates; fills Geo.
__device__ void geometry(const double* X /*[9][2]*/, const double* dc, const double* de, const double* dcc,
                         const double* dee, const double* dce, double wgt, Geo& G) {
  double dxdc = dc[0] * X[0], dydc = dc[0] * X[1], dxde = de[0] * X[0], dyde = de[0] * X[1];
#pragma unroll
  for (int n = 1; n < 9; n++) {
    dxdc = dxdc + dc[n] * X[2 * n];
    dydc = dydc + dc[n] * X[2 * n + 1];
    dxde = dxde + de[n] * X[2 * n];
    dyde = dyde + de[n] * X[2 * n + 1];
  }
  const double detJ = dxdc * dyde - dxde * dydc;
  const double dcdx = dyde / detJ;
  const double dedx = -(dydc / detJ);
  const double dcdy = -(dxde / detJ);
  const double dedy = dxdc / detJ;
  double H = 0.0;
#pragma unroll
  for (int n = 0; n < 9; n++) {
    const double fx = dcdx * dc[n] + dedx * de[n];
    const double fy = dcdy * dc[n] + dedy * de[n];
    G.dFx[n] = fx;
    G.dFy[n] = fy;
    G.d2x[n] = (dcdx * dcdx) * dcc[n] + (dedx * dedx) * dee[n] + 2.0 * dcdx * dedx * dce[n];
    G.d2y[n] = (dcdy * dcdy) * dcc[n] + (dedy * dedy) * dee[n] + 2.0 * dcdy * dedy * dce[n];
    H = (n == 0) ? fx * fx + fy * fy : H + fx * fx + fy * fy;
  }
  const double Helem = 1.0 / sqrt(H);
  G.H2 = Helem * Helem;
  G.WET = wgt * fabs(detJ);
  G.H4 = G.H2 * G.H2;
  G.yH4 = 1.0 / G.H4;
  int lo = 0, hi = 0;
#pragma unroll
  for (int n = 0; n < 9; n++) {
    lo = min(lo, min(binade(G.dFx[n]), binade(G.dFy[n])));
    hi = max(hi, max(binade(G.dFx[n]), binade(G.dFy[n])));
  }
  G.in_range = lo >= -63 && hi <= 64;
}

// Gauss-point quantities that the test-node terms read (GaussPointQuantities,
// src/Global_Variables_mod.F90:56-84, 2D part).
struct GP {
  double Ux, Uy;
  double Mom0, Mom1, Cont;
  double RC0, RC1;  // ReN*Convection(i)
  double P00, P01, P10, P11;  // P_tot
  double CE00, CE01, CE10, CE11;  // Con_EQ
  double S00, S01, S11;
  Recip rW;    // S11+S22, the omega denominator
  Recip rW2;   // WiN*WiN
  double dS00x, dS00y, dS01x, dS01y, dS11x, dS11y;
  double WiN, f_T;
  double tlsme, tlsic, tlsce, tdcs, mag_re;
};

// setFlowUnknowns .. setStabilizationTerms for one state.  v(n, q) reads the
// lane's nodal value of unknown q at node n; bf are the basis values at the
// Gauss point; H2 = Helem**2.
// MODEL: 0 L-PTT, 1 Giesekus.  DCS: discontinuity capturing on (DfN <= 1e-12).
// INERTIA: ReN /= 0.  With ReN = 0 the convective and (ReN*tsugn)**2 terms are
// signed zeros whose additions leave every value unchanged, so they are not
// computed; likewise S^3 and C:C (multiplied by f(3) = 0 for L-PTT) and the DCS
// terms (tdcs = 0 when DfN > 1e-12).
template <int MODEL, bool DCS, bool INERTIA, bool EXACT, class V>
__device__ __forceinline__ void gauss_point(const V& v, const double* bf, const Geo& G, double EU,
                                            const DevParams& prm, GP& o, DivCheck& dc) {
  // Interpolation, each sum sequential over the nodes (dot_product).
  double Ux, Uy, GU00, GU10, GU01, GU11, G2x0, G2y0, G2x1, G2y1, P, dPx, dPy;
  double S00, S01, S11, dS00x, dS00y, dS01x, dS01y, dS11x, dS11y;
  {
    const double b = bf[0], fx = G.dFx[0], fy = G.dFy[0], gx = G.d2x[0], gy = G.d2y[0];
    const double u0 = v(0, 0), u1 = v(0, 1), u2 = v(0, 2), u3 = v(0, 3), u4 = v(0, 4), u5 = v(0, 5);
    Ux = u0 * b; GU00 = u0 * fx; GU10 = u0 * fy; G2x0 = u0 * gx; G2y0 = u0 * gy;
    Uy = u1 * b; GU01 = u1 * fx; GU11 = u1 * fy; G2x1 = u1 * gx; G2y1 = u1 * gy;
    P = u2 * b; dPx = u2 * fx; dPy = u2 * fy;
    S00 = u3 * b; dS00x = u3 * fx; dS00y = u3 * fy;
    S01 = u4 * b; dS01x = u4 * fx; dS01y = u4 * fy;
    S11 = u5 * b; dS11x = u5 * fx; dS11y = u5 * fy;
  }
#pragma unroll 2
  for (int n = 1; n < 9; n++) {
    const double b = bf[n], fx = G.dFx[n], fy = G.dFy[n], gx = G.d2x[n], gy = G.d2y[n];
    const double u0 = v(n, 0), u1 = v(n, 1), u2 = v(n, 2), u3 = v(n, 3), u4 = v(n, 4), u5 = v(n, 5);
    Ux = Ux + u0 * b; GU00 = GU00 + u0 * fx; GU10 = GU10 + u0 * fy; G2x0 = G2x0 + u0 * gx; G2y0 = G2y0 + u0 * gy;
    Uy = Uy + u1 * b; GU01 = GU01 + u1 * fx; GU11 = GU11 + u1 * fy; G2x1 = G2x1 + u1 * gx; G2y1 = G2y1 + u1 * gy;
    P = P + u2 * b; dPx = dPx + u2 * fx; dPy = dPy + u2 * fy;
    S00 = S00 + u3 * b; dS00x = dS00x + u3 * fx; dS00y = dS00y + u3 * fy;
    S01 = S01 + u4 * b; dS01x = dS01x + u4 * fx; dS01y = dS01y + u4 * fy;
    S11 = S11 + u5 * b; dS11x = dS11x + u5 * fx; dS11y = dS11y + u5 * fy;
  }
  const double S10 = S01;

  // setContinuationParameter: WiN = WiN0*EU(Nex)
  const double WiN = prm.WiN0 * EU;

  // setFlowKinematics (N = 2): traceless gradient, G_dot_tr
  const double tr = GU00 + GU11;
  const double htr = (tr * 2.0) / (2.0 * 2.0);  // (traceTensor*I)/N as a complex division
  const double Gt00 = GU00 - htr, Gt01 = GU01, Gt10 = GU10, Gt11 = GU11 - htr;
  const double Gd00 = Gt00 + Gt00, Gd01 = Gt01 + Gt10, Gd10 = Gt10 + Gt01, Gd11 = Gt11 + Gt11;

  // setCTensorSQRT: C = S.S, dC/dx_k = dS_k.S + S.dS_k
  const double C00 = S00 * S00 + S01 * S10, C01 = S00 * S01 + S01 * S11;
  const double C10 = S10 * S00 + S11 * S10, C11 = S10 * S01 + S11 * S11;
  const double dC00x = (dS00x * S00 + dS01x * S10) + (S00 * dS00x + S01 * dS01x);
  const double dC01x = (dS00x * S01 + dS01x * S11) + (S00 * dS01x + S01 * dS11x);
  const double dC10x = (dS01x * S00 + dS11x * S10) + (S10 * dS00x + S11 * dS01x);
  const double dC11x = (dS01x * S01 + dS11x * S11) + (S10 * dS01x + S11 * dS11x);
  const double dC00y = (dS00y * S00 + dS01y * S10) + (S00 * dS00y + S01 * dS01y);
  const double dC01y = (dS00y * S01 + dS01y * S11) + (S00 * dS01y + S01 * dS11y);
  const double dC10y = (dS01y * S00 + dS11y * S10) + (S10 * dS00y + S11 * dS01y);
  const double dC11y = (dS01y * S01 + dS11y * S11) + (S10 * dS01y + S11 * dS11y);

  // conEQ_L_PTT_SQRT / conEQ_Giesekus_SQRT: f(1..3), f_T (g = 1, dg/dxi = 0)
  double f0, f1, f2, f_T;
  if constexpr (MODEL == 0) {
    const double trC = C00 + C11;
    f_T = 1.0 + prm.VeN * (trC - 2.0);
    f0 = f_T;
    f1 = f_T;
    f2 = 0.0;
  } else {
    f_T = 1.0 - 2.0 * prm.VeN;
    f0 = 1.0 - prm.VeN;
    f1 = 1.0 - 2.0 * prm.VeN;
    f2 = prm.VeN;
  }

  // setStressTensorSQRT: T = ((1-BvN)/WiN)(C - I), P_tot = T + BvN G_dot_tr - P I
  const double a = prm.omb / WiN;
  const double T00 = a * (C00 - 1.0), T01 = a * (C01 - 0.0), T10 = a * (C10 - 0.0), T11 = a * (C11 - 1.0);
  const double T22 = a * (0.0 - 1.0);
  o.P00 = T00 + prm.BvN * Gd00 - P;
  o.P01 = T01 + prm.BvN * Gd01;
  o.P10 = T10 + prm.BvN * Gd10;
  o.P11 = T11 + prm.BvN * Gd11 - P;

  // setUpperConvectedDerivativeSQRT: UCD = U.grad S - S.GU_tr - W.S
  const Recip rW = make_recip<EXACT>(S00 + S11, dc);
  const double W12 = qdiv<EXACT>(-(S00 * Gt01) + S01 * (Gt00 - Gt11) + S11 * Gt10, rW, dc);
  const double W21 = -W12;
  double U00 = Ux * dS00x + Uy * dS00y, U01 = Ux * dS01x + Uy * dS01y;
  double U10 = U01, U11 = Ux * dS11x + Uy * dS11y;
  U00 = U00 - (S00 * Gt00 + S01 * Gt10) - W12 * S10;
  U01 = U01 - (S00 * Gt01 + S01 * Gt11) - W12 * S11;
  U10 = U10 - (S10 * Gt00 + S11 * Gt10) - W21 * S00;
  U11 = U11 - (S10 * Gt01 + S11 * Gt11) - W21 * S01;

  // setSInverse (inverseTensor2D)
  const Recip rS = make_recip<EXACT>(S00 * S11 - S01 * S10, dc);
  const double SI00 = qdiv<EXACT>(S11, rS, dc), SI01 = qdiv<EXACT>(-S01, rS, dc);
  const double SI10 = qdiv<EXACT>(-S10, rS, dc), SI11 = qdiv<EXACT>(S00, rS, dc);
  // the tracked quotients so far, summed for the finiteness test below
  const double fin_part = EXACT ? 0.0 : (W12 + (SI00 + SI01)) + (SI10 + SI11);

  // term_S = f3 S^3 + f2 S - f1 S^-1 ; Con_EQ = WiN*UCD + term_S/2
  double tS00, tS01, tS10, tS11;
  if constexpr (MODEL == 0) {
    tS00 = f1 * S00 - f0 * SI00;
    tS01 = f1 * S01 - f0 * SI01;
    tS10 = f1 * S10 - f0 * SI10;
    tS11 = f1 * S11 - f0 * SI11;
  } else {
    const double S3_00 = S00 * C00 + S01 * C10, S3_01 = S00 * C01 + S01 * C11;
    const double S3_10 = S10 * C00 + S11 * C10, S3_11 = S10 * C01 + S11 * C11;
    tS00 = f2 * S3_00 + f1 * S00 - f0 * SI00;
    tS01 = f2 * S3_01 + f1 * S01 - f0 * SI01;
    tS10 = f2 * S3_10 + f1 * S10 - f0 * SI10;
    tS11 = f2 * S3_11 + f1 * S11 - f0 * SI11;
  }
  const double CE00 = WiN * U00 + tS00 / 2.0, CE01 = WiN * U01 + tS01 / 2.0;
  const double CE10 = WiN * U10 + tS10 / 2.0, CE11 = WiN * U11 + tS11 / 2.0;

  // strongFormMomentum: Momentum = ReN*(U.grad U) - (-grad P + div T + BvN lap U)
  const double divT0 = a * dC00x + a * dC10y;
  const double divT1 = a * dC01x + a * dC11y;
  const double dP0 = -dPx + divT0 + prm.BvN * (G2x0 + G2y0);
  const double dP1 = -dPy + divT1 + prm.BvN * (G2x1 + G2y1);
  if constexpr (INERTIA) {
    const double Cv0 = Ux * GU00 + Uy * GU10;
    const double Cv1 = Ux * GU01 + Uy * GU11;
    o.RC0 = prm.ReN * Cv0;
    o.RC1 = prm.ReN * Cv1;
    o.Mom0 = o.RC0 - dP0;
    o.Mom1 = o.RC1 - dP1;
  } else {
    o.RC0 = 0.0;
    o.RC1 = 0.0;
    o.Mom0 = -dP0;
    o.Mom1 = -dP1;
  }
  o.Cont = GU00 + GU11;  // strongFormContinuity

  // setStabilizationTerms
  double tsugn = fabs(Ux * G.dFx[0] + Uy * G.dFy[0]);
#pragma unroll
  for (int n = 1; n < 9; n++) tsugn = tsugn + fabs(Ux * G.dFx[n] + Uy * G.dFy[n]);
  const double H2 = G.H2;
  const double ddT = T00 * T00 + T01 * T10 + T10 * T01 + T11 * T11 + T22 * T22;
  const double ddG = Gd00 * Gd00 + Gd01 * Gd10 + Gd10 * Gd01 + Gd11 * Gd11;
  double Ha = sqrt(a * a + 0.5 * ddT);
  Ha = cdiv(Ha, sqrt(1.0 + 0.5 * ddG));
  // (ReN/dt)**2 with dt = +huge is exactly zero and adds nothing.
  const double r3 = G.r3;
  check_denominator<EXACT>(G.H4, dc);
  const double r4 = qdiv<EXACT>(Ha * H2, Recip{G.H4, G.yH4}, dc);
  double tlsme;
  if constexpr (INERTIA) {
    const double r2 = prm.ReN * tsugn;
    tlsme = cdiv(1.0, sqrt(r2 * r2 + r3 * r3 + r4 * r4));
  } else {
    tlsme = cdiv(1.0, sqrt(r3 * r3 + r4 * r4));
  }
  o.tlsme = tlsme;
  o.tlsic = cdiv(H2, tlsme);
  double tdcs = 0.0, mag_re = 0.0, mag_im = 0.0;
  if constexpr (DCS) {
    const double ddS = S00 * S00 + S01 * S10 + S10 * S01 + S11 * S11;
    tdcs = cdiv(H2, sqrt(0.5 * ddS));
    const double mc = 0.5 * (CE00 * CE00 + CE01 * CE10 + CE10 * CE01 + CE11 * CE11);
    if (mc >= 0.0) mag_re = sqrt(mc);
    else mag_im = sqrt(-mc);
  }
  o.tdcs = tdcs;
  o.mag_re = mag_re;
  // (WiN/dt)**2 with dt = +huge is exactly zero and adds nothing.
  const double t2 = WiN * tsugn, t3 = WiN * sqrt(0.5 * ddG), t4 = f1;
  double sum = t2 * t2 + t3 * t3 + t4 * t4;
  if constexpr (MODEL != 0) {
    const double ddC = C00 * C00 + C01 * C10 + C10 * C01 + C11 * C11;
    const double t5 = f2 * sqrt(0.5 * ddC);
    sum = sum + t5 * t5;
  }
  if constexpr (DCS) {
    const double zr = cdiv(tdcs * mag_re, H2), zi = cdiv(tdcs * mag_im, H2);
    sum = sum + (zr * zr - zi * zi);
  }
  const double t7 = G.t7;
  sum = sum + t7 * t7;
  o.tlsce = cdiv(1.0, sqrt(sum));
  // Inputs of the test-node divisions (node_terms), range-checked once here.
  base_range<EXACT>(S00, dc);
  base_range<EXACT>(S01, dc);
  base_range<EXACT>(S11, dc);
  base_range<EXACT>(o.tlsce, dc);
  base_range<EXACT>(WiN, dc);
  base_range<EXACT>(prm.omb, dc);
  if (!EXACT && !G.in_range) dc.hi = kBadExp;
  if constexpr (!EXACT) {
    // The range checks above take an infinity or a NaN for 0: one finiteness
    // test of the sum of every checked quantity and tracked quotient catches
    // them (inf - inf is NaN; the checked ranges keep finite sums far below
    // overflow).
    const double all = fin_part + (((S00 + S01) + (S11 + o.tlsce)) + ((WiN + prm.omb) + r4));
    if (!__builtin_isfinite(all)) dc.hi = kBadExp;
  }

  o.Ux = Ux; o.Uy = Uy;
  o.CE00 = CE00; o.CE01 = CE01; o.CE10 = CE10; o.CE11 = CE11;
  o.S00 = S00; o.S01 = S01; o.S11 = S11;
  o.rW = rW;
  o.rW2 = make_recip<EXACT>(WiN * WiN, dc);
  o.dS00x = dS00x; o.dS00y = dS00y; o.dS01x = dS01x; o.dS01y = dS01y; o.dS11x = dS11x; o.dS11y = dS11y;
  o.WiN = WiN; o.f_T = f_T;
}

// Test-node rows of one Gauss point (momentumMainSQRT, continuityBulkMain,
// stressesBulkMainSQRT and their LSME/LSIC/LSCE/DCS terms, Residuals_mod.F90:
// 90-110,161-234,295-490), already multiplied by WET.
//
// Why the divisions here need no tracking.  Write Z(a, b) for {0} together
// with the magnitudes in [2^a, 2^b].  In the fast pass gauss_point has checked
// S00, S01, S11, tlsce, WiN, omb and (through Geo::in_range) fx, fy to lie in
// Z(-64, 64), and make_recip has checked rW and rW2 to lie in [2^-60, 2^40],
// so y = RN(1/d) lies in [2^-41, 2^61].  A rounded product of values in
// Z(a, b) and Z(c, d) lies in Z(a+c, b+d); a rounded sum or difference of
// values in Z(a, b) lies in Z(a-52, b+1) (both operands are multiples of
// 2^(a-52)).  Hence: the omega numerators and m00, m10 lie in Z(-180, 129),
// their q0 in Z(-221, 190) and w in Z(-220, 189); the u's (-m - w*S or
// -(w*S)) lie in Z(-336, 254); the l numerators ((tlsce*u)*omb)*WiN in
// Z(-528, 446) and their q0 in Z(-569, 507).  Every q0 is inside the window
// [2^-901, 2^900) of the two-FMA division, and nothing leaves the normal
// range, so each quotient is the IEEE one.
template <int MODEL, bool DCS, bool INERTIA, bool EXACT>
__device__ __forceinline__ void node_terms(const GP& g, double bifn, double fx, double fy, const DevParams& prm,
                                           double WET, double* out6) {
  const double S10 = g.S01;
  const double Udf = g.Ux * fx + g.Uy * fy;
  const double lsme_m = INERTIA ? g.tlsme * prm.ReN * Udf : 0.0;
  double T0, T1;
  {
    // i = 1: GUT_lsce row 1 = grad(phi)
    const double lsic = (g.tlsic * fx) * g.Cont;
    const double w = qdiv_ranged<EXACT>(g.S01 * fx + g.S11 * fy, g.rW);  // omegaNominator2D / omegaDenominator2D
    const double m00 = g.S00 * fx + g.S01 * fy, m10 = S10 * fx + g.S11 * fy;
    const double u00 = -m00 - w * S10, u01 = -(w * g.S11);
    const double u10 = -m10 - (-w) * g.S00, u11 = -((-w) * g.S01);
    const double l00 = qdiv_ranged<EXACT>((g.tlsce * u00) * prm.omb * g.WiN, g.rW2);
    const double l01 = qdiv_ranged<EXACT>((g.tlsce * u01) * prm.omb * g.WiN, g.rW2);
    const double l10 = qdiv_ranged<EXACT>((g.tlsce * u10) * prm.omb * g.WiN, g.rW2);
    const double l11 = qdiv_ranged<EXACT>((g.tlsce * u11) * prm.omb * g.WiN, g.rW2);
    const double lsce = l00 * g.CE00 + l01 * g.CE10 + l10 * g.CE01 + l11 * g.CE11;
    if constexpr (INERTIA) T0 = g.RC0 * bifn + (g.P00 * fx + g.P01 * fy) + lsme_m * g.Mom0 + lsic + lsce;
    else T0 = (g.P00 * fx + g.P01 * fy) + lsic + lsce;
  }
  {
    // i = 2: GUT_lsce row 2 = grad(phi)
    const double lsic = (g.tlsic * fy) * g.Cont;
    const double w = qdiv_ranged<EXACT>(-(g.S00 * fx) + g.S01 * (-fy), g.rW);
    const double m01 = g.S00 * fx + g.S01 * fy, m11 = S10 * fx + g.S11 * fy;
    const double u00 = -(w * S10), u01 = -m01 - w * g.S11;
    const double u10 = -((-w) * g.S00), u11 = -m11 - (-w) * g.S01;
    const double l00 = qdiv_ranged<EXACT>((g.tlsce * u00) * prm.omb * g.WiN, g.rW2);
    const double l01 = qdiv_ranged<EXACT>((g.tlsce * u01) * prm.omb * g.WiN, g.rW2);
    const double l10 = qdiv_ranged<EXACT>((g.tlsce * u10) * prm.omb * g.WiN, g.rW2);
    const double l11 = qdiv_ranged<EXACT>((g.tlsce * u11) * prm.omb * g.WiN, g.rW2);
    const double lsce = l00 * g.CE00 + l01 * g.CE10 + l10 * g.CE01 + l11 * g.CE11;
    if constexpr (INERTIA) T1 = g.RC1 * bifn + (g.P10 * fx + g.P11 * fy) + lsme_m * g.Mom1 + lsic + lsce;
    else T1 = (g.P10 * fx + g.P11 * fy) + lsic + lsce;
  }
  // continuity row
  const double T2 = g.Cont * bifn + g.tlsme * (fx * g.Mom0 + fy * g.Mom1);
  // stress rows (xx, xy, yy)
  double lsce_s = g.WiN * Udf;
  lsce_s = lsce_s + ((g.f_T * bifn) * 2.0) / (2.0 * 2.0);
  lsce_s = g.tlsce * lsce_s;
  const double d00 = g.dS00x * fx + g.dS00y * fy;
  const double d01 = g.dS01x * fx + g.dS01y * fy;
  const double d11 = g.dS11x * fx + g.dS11y * fy;
  double T3 = g.CE00 * bifn + prm.DfN * d00 + (g.tlsme * (-(fx * g.Mom0)) * g.WiN) + lsce_s * g.CE00;
  double T4 = g.CE01 * bifn + prm.DfN * d01 + (g.tlsme * (-(fy * g.Mom0) - fx * g.Mom1) * g.WiN) + lsce_s * g.CE01;
  double T5 = g.CE11 * bifn + prm.DfN * d11 + (g.tlsme * (-(fy * g.Mom1)) * g.WiN) + lsce_s * g.CE11;
  if constexpr (DCS) {
    T3 = T3 + (g.tdcs * d00) * g.mag_re;
    T4 = T4 + (g.tdcs * d01) * g.mag_re;
    T5 = T5 + (g.tdcs * d11) * g.mag_re;
  }
  out6[0] = T0 * WET;
  out6[1] = T1 * WET;
  out6[2] = T2 * WET;
  out6[3] = T3 * WET;
  out6[4] = T4 * WET;
  out6[5] = T5 * WET;
}

// Nodal unknowns of one evaluation: u[6*inod + q] in LDS.
struct PlainState {
  const double* u;  // [54]
  __device__ __forceinline__ double operator()(int n, int q) const { return u[n * 6 + q]; }
};

// ----------------------------------------------------------------------------
// Bulk kernel: one wavefront per workgroup.  A workgroup holds EPB elements x
// KPB evaluations = 7 slots; lane t < 63 is Gauss point g = t / 7 of slot
// s = t % 7 (element el = s / KPB, evaluation k = kg*KPB + s % KPB, where kg
// is the workgroup's evaluation group of its element).  NEVAL = 56 assembles
// residual and FD Jacobian (8 workgroups per element), NEVAL = 1 the residual
// only (7 elements per workgroup).
// ----------------------------------------------------------------------------
// Where the Jacobian entries go.  For element e and its node pair (jn, in),
// code[e*81 + jn*9 + in] is
//   pos<<2 | kCodeStore  when the element is the only one containing both
//       nodes, or the first (by colour) of the two that do: the 36 entries of
//       the pair are written directly, at rowptr[6r] + jq*len(r) + pos*6 + iq
//       (r = node jn; every row of a node has the same length);
//   pos<<2 | kCodeAdd    for the second of two: added to what the first wrote
//       (the elements of one colour share no such pair, and the colours are
//       launched in order);
//   cidx<<2 | kCodeBank  when three or more elements contain both: to
//       cbank[cidx*36 + jq*6 + iq] for the ordered sum of vef_shared_kernel.
// The dR/dEU quotients go to acbank[e*54 + row].
constexpr int kCodeStore = 0, kCodeBank = 1, kCodeAdd = 2;
struct FdOut {
  const double* r0;      // [n_el][54] base residual rows (residual pass)
  const int32_t* code;   // [n_el][81]
  const int64_t* rowptr;
  double* values;
  double* cbank;
  double* acbank;
  double* geo;           // [n_el][9] Geo records: written by the residual pass when a
                         // Jacobian pass follows (else NULL), read by the Jacobian pass
  const int32_t* elist;  // Jacobian passes: the elements of this launch (one colour)
  int32_t n_list;
};
constexpr int kGeoWords = (int)(sizeof(Geo) / sizeof(double));
static_assert(sizeof(Geo) % sizeof(double) == 0, "Geo is a whole number of 8-byte words");

template <int EPB, int KPB, int NEVAL, int MODEL, bool DCS, bool INERTIA>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8)))
vef_bulk_kernel(const double* __restrict__ coords, const int32_t* __restrict__ conn, const double* __restrict__ x,
                int64_t eu_index, int32_t n_el, DevParams prm, double* __restrict__ bank, FdOut fo) {
  constexpr int NSLOT = EPB * KPB;
  static_assert(NSLOT == 7, "seven slots per wavefront");
  constexpr int BPE = NEVAL / KPB;  // workgroups per element group
  constexpr bool FD = NEVAL > 1;
  __shared__ double sS[NSLOT][54];  // the nodal unknowns each slot evaluates at
  __shared__ double sX[EPB][18];
  __shared__ Geo sG[EPB][9];
  __shared__ double sRed[9][NSLOT][6];  // one buffer: 10.1 KB of LDS per workgroup -> 16 per CU
  __shared__ double sR0[FD ? 54 : 1];
  __shared__ int32_t sCode[FD ? 81 : 1];
  __shared__ int64_t sRb[FD ? 9 : 1];
  __shared__ int32_t sRl[FD ? 9 : 1];

  const int tid = threadIdx.x;
  // Workgroups are dispatched round-robin over the 8 XCDs (each with its own
  // L2).  With BPE > 1 the BPE workgroups of one element are given ids that
  // land on the same XCD (ids 64m + 8j + x, j = 0..7, for element 8m + x), so
  // the element's inputs are fetched once per L2 and its bank rows, written in
  // 7-evaluation pieces, merge in one L2 before reaching HBM.
  int64_t egroup;
  int kg;
  if (BPE == 8) {
    const int64_t b = blockIdx.x;
    egroup = (b >> 6) * 8 + (b & 7);
    kg = (int)((b >> 3) & 7);
  } else {
    egroup = blockIdx.x / BPE;
    kg = blockIdx.x % BPE;
  }
  if (FD && egroup >= fo.n_list) return;  // padding workgroups of the XCD-aware numbering (uniform)
  const int64_t e0 = FD ? (int64_t)fo.elist[egroup] : egroup * EPB;
  const double EU = x[eu_index];
  const int64_t elast = (int64_t)n_el - 1;

  // Gather element unknowns (copyToLocal) and build each slot's state: for
  // evaluation k of an element, unknown j is u_j for j > p, fl(u_p+h) at p and
  // fl(fl(u_j+h)-h) for j < p, p = k-1 — the in-place perturb/restore sequence
  // of storeBulk (Equations_mod.F90:267-306) as perturbation p sees it.
  // Every gather here issues all of a lane's loads before the first is used
  // (indices clamped into range, only the LDS writes predicated), so their
  // latencies overlap instead of adding up.
  {
    constexpr int NG = NSLOT * 54, NI = (NG + 63) / 64;
    int32_t gn[NI];
#pragma unroll
    for (int i = 0; i < NI; i++) {
      const int t = min(tid + 64 * i, NG - 1);
      const int64_t e = e0 + (t / 54) / KPB;
      gn[i] = conn[(e < elast ? e : elast) * 9 + (t % 54) / 6];
    }
    double xv[NI];
#pragma unroll
    for (int i = 0; i < NI; i++) {
      const int t = min(tid + 64 * i, NG - 1);
      xv[i] = x[(int64_t)gn[i] * 6 + (t % 54) % 6];
    }
#pragma unroll
    for (int i = 0; i < NI; i++) {
      const int t = tid + 64 * i;
      if (t < NG) {
        const int ss = t / 54, j = t % 54;
        const int el = ss / KPB;
        const int k = kg * KPB + ss % KPB;
        double u = (e0 + el < n_el) ? xv[i] : 0.0;
        if (NEVAL > 1) {
          const int p = k - 1;
          const double up = u + kEpsJac;
          u = (j < p) ? up - kEpsJac : ((j == p) ? up : u);
        }
        sS[ss][j] = u;
      }
    }
  }
  {
    constexpr int NG = EPB * 18, NI = (NG + 63) / 64;
    int32_t gn[NI];
#pragma unroll
    for (int i = 0; i < NI; i++) {
      const int t = min(tid + 64 * i, NG - 1);
      const int64_t e = e0 + t / 18;
      gn[i] = conn[(e < elast ? e : elast) * 9 + (t % 18) / 2];
    }
    double cv[NI];
#pragma unroll
    for (int i = 0; i < NI; i++) {
      const int t = min(tid + 64 * i, NG - 1);
      cv[i] = coords[(int64_t)gn[i] * 2 + (t & 1)];
    }
#pragma unroll
    for (int i = 0; i < NI; i++) {
      const int t = tid + 64 * i;
      if (t < NG) {
        const int el = t / 18, c = t % 18;
        sX[el][c] = (e0 + el < n_el) ? cv[i] : ((c & 1) ? 1.0 : 0.0);
      }
    }
  }
  if constexpr (FD) {
    // The element's codes, base residual rows, row offsets and Gauss-point
    // geometry (basis2D, Helem**2 and the two quotients by it, which do not
    // depend on the state: the residual pass computed them once per element
    // and wrote them to fo.geo — the same bits: same code, same coordinates).
    // e0 < n_el here.
    constexpr int NW = 9 * kGeoWords, NI = (NW + 63) / 64;
    const int32_t c0 = fo.code[e0 * 81 + tid], c1 = fo.code[e0 * 81 + min(tid + 64, 80)];
    const double r0v = fo.r0[e0 * 54 + min(tid, 53)];
    const int64_t r6 = (int64_t)conn[e0 * 9 + min(tid, 8)] * 6;
    const double* gsrc = fo.geo + e0 * 9 * kGeoWords;
    double gv[NI];
#pragma unroll
    for (int i = 0; i < NI; i++) gv[i] = gsrc[min(tid + 64 * i, NW - 1)];
    const int64_t rb = fo.rowptr[r6], re = fo.rowptr[r6 + 1];
    sCode[tid] = c0;
    if (tid + 64 < 81) sCode[tid + 64] = c1;
    if (tid < 54) sR0[tid] = r0v;
    if (tid < 9) {
      sRb[tid] = rb;
      sRl[tid] = (int32_t)(re - rb);
    }
    double* gdst = reinterpret_cast<double*>(&sG[0][0]);
#pragma unroll
    for (int i = 0; i < NI; i++)
      if (tid + 64 * i < NW) gdst[tid + 64 * i] = gv[i];
  }
  __syncthreads();
  // Gauss-point geometry (basis2D, Helem**2 and the two quotients by it): the
  // residual pass computes it, 63 lanes at a time, and hands it to the
  // Jacobian pass through fo.geo when one follows.
  if constexpr (!FD) {
    for (int t = tid; t < EPB * 9; t += 64) {
      const int el = t / 9, g = t % 9;
      Geo& G = sG[el][g];
      geometry(sX[el], c_tab.dc[g], c_tab.de[g], c_tab.dcc[g], c_tab.dee[g], c_tab.dce[g], c_tab.W2[g], G);
      G.r3 = prm.BvN / G.H2;
      G.t7 = prm.DfN / G.H2;
    }
    if (fo.geo) {
      __syncthreads();
      const int64_t nv = (e0 + EPB <= n_el ? EPB : n_el - e0) * 9 * kGeoWords;
      const double* src = reinterpret_cast<const double*>(&sG[0][0]);
      double* dst = fo.geo + e0 * 9 * kGeoWords;
      for (int t = tid; t < nv; t += 64) dst[t] = src[t];
    }
    __syncthreads();
  }

  const bool active = tid < 9 * NSLOT;
  const int g = active ? tid / NSLOT : 0;
  const int s = tid % NSLOT;
  const int el = s / KPB;
  const int k = kg * KPB + s % KPB;
  const bool valid = active && (e0 + el < n_el);

  // One pass over the wavefront's 7 slots x 9 Gauss points: Gauss-point
  // quantities, then per test node the 6 terms, their ordered sum over the
  // Gauss points and its destination.  The fast pass divides with shared
  // reciprocals.  Whether its divisions equal IEEE ones is settled before the
  // first store: the Gauss-point divisions are tracked and the test-node
  // divisions bounded by the checked ranges of their inputs (node_terms), so
  // if any lane's DivCheck fails the wavefront goes to the exact pass before
  // writing anything (the branch is uniform: one wavefront per workgroup),
  // and otherwise its sums equal the exact pass's bit for bit and only the
  // forward-difference quotients remain, each decided on its own
  // (fd_quotient_entry).  Every output is therefore written exactly once,
  // which the in-place sums of the two-contributor entries need.
  auto evaluate = [&](auto exact_tag) -> bool {
    constexpr bool EXACT = decltype(exact_tag)::value;
    DivCheck dc{0, 0, 0, 0};
    GP q;
    if (valid) {
      const double EUk = (NEVAL > 1 && k == NEVAL - 1) ? EU + kEpsJac : EU;
      PlainState st{sS[s]};
      gauss_point<MODEL, DCS, INERTIA, EXACT>(st, c_tab.bf[g], sG[el][g], EUk, prm, q, dc);
    }
    if constexpr (!EXACT) {
      if (__ballot(dc.failed()) != 0) return true;
    }
    const double WET = sG[el][g].WET;

    for (int n = 0; n < 9; n++) {
      __syncthreads();  // the previous sums have been read
      // The destination of this lane's quotient for test node n, and, for the
      // second of two contributors, the first one's value: loaded here so that
      // the test-node terms below hide its latency.
      const int c = tid / NSLOT, ss = tid % NSLOT;
      const int eel = ss / KPB, kk = kg * KPB + ss % KPB;
      const bool sum_lane = tid < 6 * NSLOT && e0 + eel < n_el;
      double* dst = nullptr;
      double first = 0.0;
      bool add = false;
      if constexpr (FD) {
        const int in = (kk - 1) / 6, iq = (kk - 1) % 6;
        const int32_t code = sCode[n * 9 + min(max(in, 0), 8)];
        const int kind = code & 3;
        double* pv = fo.values + (sRb[n] + (int64_t)c * sRl[n] + (int64_t)(code >> 2) * 6 + iq);
        double* pc = fo.cbank + ((int64_t)(code >> 2) * 36 + c * 6 + iq);
        double* pa = fo.acbank + (e0 * 54 + n * 6 + c);
        dst = kk == NEVAL - 1 ? pa : (kind == kCodeBank ? pc : pv);
        add = sum_lane && kk > 0 && kk != NEVAL - 1 && kind == kCodeAdd;
        if (add) first = *pv;
      }
      if (valid) {
        double r[6];
        node_terms<MODEL, DCS, INERTIA, EXACT>(q, c_tab.bf[g][n], sG[el][g].dFx[n], sG[el][g].dFy[n], prm, WET,
                                               r);
#pragma unroll
        for (int cc = 0; cc < 6; cc++) sRed[g][s][cc] = r[cc];
      }
      __syncthreads();
      // Ordered sum over the Gauss points: Res = Res + TERM_RES*WET, ig = 1..9.
      if (sum_lane) {
        double acc = sRed[0][ss][c];
#pragma unroll
        for (int gg = 1; gg < 9; gg++) acc = acc + sRed[gg][ss][c];
        if (!FD) {
          bank[(e0 + eel) * 54 + n * 6 + c] = acc;
        } else if (kk > 0) {
          // (dRes - Res)/eps_jac for column kk-1 (or EU for kk = 55), as
          // ADD_VALUES into zeroed storage (+0 for -0).  The ordered sums of
          // cbank and acbank start from +0 as well, so every destination but
          // the second of two contributors takes 0 + q (the destination is
          // selected, not branched on); the second adds its 0 + q to the
          // first's: (0 + a) + (0 + b) = (0 + a) + b = (0 + b) + a, the
          // reference's sum in either element order.
          const double qv = 0.0 + fd_quotient_entry<EXACT>(acc, sR0[n * 6 + c]);
          *dst = add ? first + qv : qv;
        }
      }
    }
    return false;
  };
  const bool exact = evaluate(std::false_type{});
  // A fence between the passes: without it the compiler interleaves the exact
  // pass with the fast one's node loop, and at the 128-VGPR limit of this
  // occupancy that spills.
  __syncthreads();
  if (exact) evaluate(std::true_type{});
}

// ----------------------------------------------------------------------------
// Analytic FP32 Jacobian (VEF_RES_JAC_FP32): one 192-thread workgroup (three
// wavefronts) per element.
//  (1) The 22 fields at the 9 Gauss points in FP32.
//  (2) 171 lanes = 9 Gauss points x 19 directions evaluate the Gauss-point
//      coefficients T(F) of the test functions (vef_ad_math.h) with
//      one-component dual numbers, giving D = WET * dT/dF in LDS.  The four
//      second derivatives of U are not differentiated: they enter T only
//      through BvN * (d2U_k/dx2 + d2U_k/dy2) beside -dP/dx_k in the strong
//      momentum residual, so dT/d(d2U_k/dx2) = dT/d(d2U_k/dy2) =
//      -BvN * dT/d(dP/dx_k) exactly; the 19 directions are the other 18
//      fields and EU.
//  (3) Every matrix entry is sum_g sum_k phi^n_k(g) E_g[k][q,m,p] with
//      E_g[k][q,m,p] = sum_{f in p} D[g][f][k,q] beta^m_f(g).  Per Gauss point
//      162 threads, one per (k, q, m), form the six E_g[k][q,m,0..5] from 18 D
//      values and the node's 5 basis values (branch-free, FMAs), into one of
//      two LDS buffers; then the outer sum runs on the matrix cores as
//      J += Phi_g^T E_g, a 16 x 4 by 4 x 336 product (test nodes padded to 16,
//      the three basis kinds and a zero row, the 324 entries of a row block and
//      the 6 of dR/dEU padded to 336) in 21 v_mfma_f32_16x16x4f32 tiles, 7 per
//      wavefront, whose FP32 accumulators stay in registers over the 9 Gauss
//      points.  With E double-buffered one barrier per Gauss point suffices:
//      a wavefront's products from buffer g % 2 precede its arrival at the
//      barrier after which buffer g % 2 is written again.
//  Stored through the same scatter map as the forward-difference pass, dR/dEU
//  into acbank.
// ----------------------------------------------------------------------------
namespace jac32 {
constexpr int kThreads = 192;
constexpr int kDirs = 19;                 // differentiated directions per Gauss point
constexpr int kECols = 336;               // E columns: 324 entries (q*54 + m*6 + p), 6 of dR/dEU, padding
constexpr int kTiles = kECols / 16;       // 21
constexpr int kTilesPerWave = kTiles / (kThreads / 64);  // 7
static_assert(kTiles % (kThreads / 64) == 0, "the E tiles split evenly over the wavefronts");
static_assert(9 * kDirs <= kThreads, "one lane per (Gauss point, direction)");
static_assert(18 * 9 + 18 <= kThreads, "one thread per (k, q, m) and per dR/dEU entry");
// Field of direction d: the fields in their order without the four second
// derivatives of U (kG2x0 .. kG2y1), then EU.
__host__ __device__ constexpr int dir_field(int d) { return d < vefad::kG2x0 ? d : d + 4; }
// Directions of the fields E needs (dir_field inverse).
constexpr int dUx = 0, dUy = 1, dGU00 = 2, dGU10 = 3, dGU01 = 4, dGU11 = 5;
constexpr int dP = vefad::kP - 4, dPx = vefad::kPx - 4, dPy = vefad::kPy - 4;
constexpr int dS00 = vefad::kS00 - 4, dS01 = vefad::kS01 - 4, dS11 = vefad::kS11 - 4;
constexpr int dS00x = vefad::kS00x - 4, dS00y = vefad::kS00y - 4, dS01x = vefad::kS01x - 4;
constexpr int dS01y = vefad::kS01y - 4, dS11x = vefad::kS11x - 4, dS11y = vefad::kS11y - 4;
constexpr int dEU = vefad::kEU - 4;
static_assert(dir_field(dP) == vefad::kP && dir_field(dS11y) == vefad::kS11y && dir_field(dEU) == vefad::kEU &&
              dEU == kDirs - 1, "direction numbering");
}  // namespace jac32

__global__ void __launch_bounds__(jac32::kThreads) __attribute__((amdgpu_waves_per_eu(4, 8)))
vef_jac32_kernel(const int32_t* __restrict__ conn, const double* __restrict__ x, int64_t eu_index, int32_t n_el,
                 vefad::Params<float> prm, FdOut fo) {
#pragma clang fp contract(fast)  // FP32 path: compared to FP32 accuracy, FMAs allowed
  using namespace jac32;
  constexpr int NF = vefad::kNumFields;
  __shared__ float sU[54];
  __shared__ float sBas[9][5][9];  // [g][basis kind][node]
  __shared__ float sWET[9], sH2[9];
  __shared__ float sF[9][NF];
  __shared__ float sD[9][kDirs][18];  // WET * dT[k][q]/dF for the direction's field
  __shared__ __align__(16) float sM[2][3 * kECols];  // E of Gauss points g % 2: [k][column]
  __shared__ int32_t sCode[81];
  __shared__ int64_t sRb[9];
  __shared__ int32_t sRl[9];
  const int tid = threadIdx.x;
  if ((int64_t)blockIdx.x >= fo.n_list) return;  // uniform over the workgroup
  const int64_t e = fo.elist[blockIdx.x];
  const float EU = (float)x[eu_index];
  if (tid < 54) sU[tid] = (float)x[(int64_t)conn[e * 9 + tid / 6] * 6 + tid % 6];
  if (tid < 81) sCode[tid] = fo.code[e * 81 + tid];
  if (tid < 9) {
    const int64_t r6 = (int64_t)conn[e * 9 + tid] * 6;
    const int64_t rb = fo.rowptr[r6];
    sRb[tid] = rb;
    sRl[tid] = (int32_t)(fo.rowptr[r6 + 1] - rb);
  }
  // the Gauss points' basis derivatives, WET and Helem^2 from the geometry
  // records the residual pass wrote (words 0..37 of each record), as FP32
  {
    constexpr int NW = 9 * 38, NI = (NW + kThreads - 1) / kThreads;
    const double* src = fo.geo + e * 9 * kGeoWords;
#pragma unroll
    for (int i = 0; i < NI; i++) {
      const int t = tid + kThreads * i;
      if (t < NW) {
        const int g = t / 38, w = t % 38;
        const float v = (float)src[g * kGeoWords + w];
        if (w < 36) sBas[g][1 + w / 9][w % 9] = v;
        else if (w == 36) sWET[g] = v;
        else sH2[g] = v;
      }
    }
    if (tid < 81) sBas[tid / 9][0][tid % 9] = (float)c_tab.bf[tid / 9][tid % 9];
  }
  // the padding columns of E stay zero
  for (int t = tid; t < 2 * 3 * (kECols - 330); t += kThreads) {
    const int b = t / (3 * (kECols - 330)), r = t % (3 * (kECols - 330));
    sM[b][(r / (kECols - 330)) * kECols + 330 + r % (kECols - 330)] = 0.0f;
  }
  __syncthreads();
  // (1) the fields at the Gauss points (unknown and basis kind of each field
  //     from a table: selecting them per lane compiled into branch chains)
  constexpr unsigned char kFieldQ[vefad::kNumInterp] = {
#define Q(f) (unsigned char)vefad::field_unknown(f)
      Q(0), Q(1), Q(2), Q(3), Q(4), Q(5), Q(6), Q(7), Q(8), Q(9), Q(10),
      Q(11), Q(12), Q(13), Q(14), Q(15), Q(16), Q(17), Q(18), Q(19), Q(20), Q(21)};
#undef Q
  constexpr unsigned char kFieldK[vefad::kNumInterp] = {
#define K(f) (unsigned char)vefad::field_basis(f)
      K(0), K(1), K(2), K(3), K(4), K(5), K(6), K(7), K(8), K(9), K(10),
      K(11), K(12), K(13), K(14), K(15), K(16), K(17), K(18), K(19), K(20), K(21)};
#undef K
  static_assert(vefad::kNumInterp == 22, "the tables list 22 fields");
  for (int t = tid; t < 9 * vefad::kNumInterp + 9; t += kThreads) {
    if (t < 9 * vefad::kNumInterp) {
      const int g = t / vefad::kNumInterp, f = t % vefad::kNumInterp;
      const int q = kFieldQ[f], k = kFieldK[f];
      float acc = 0.0f;
#pragma unroll
      for (int n = 0; n < 9; n++) acc += sU[n * 6 + q] * sBas[g][k][n];
      sF[g][f] = acc;
    } else {
      sF[t - 9 * vefad::kNumInterp][vefad::kEU] = EU;
    }
  }
  __syncthreads();
  // (2) D = WET * dT/dF, one direction per lane
  if (tid < 9 * kDirs) {
    const int g = tid / kDirs, d = tid % kDirs, fd = dir_field(d);
    vefad::DualV<float, 1> F[NF];
#pragma unroll
    for (int i = 0; i < NF; i++) {
      F[i].v = sF[g][i];
      F[i].d[0] = i == fd ? 1.0f : 0.0f;
    }
    vefad::DualV<float, 1> T[3][6];
    vefad::gauss_terms(F, sBas[g][1], sBas[g][2], sH2[g], prm, T);
    const float w = sWET[g];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int q = 0; q < 6; q++) sD[g][d][k * 6 + q] = w * T[k][q].d[0];
  }
  // (3) per Gauss point: E_g (VALU), then J += Phi_g^T E_g (matrix cores).
  //     v_mfma_f32_16x16x4f32 lane layout: A[i][k] in lane i + 16k, B[k][j] in
  //     lane j + 16k, D[i][j] in register i % 4 of lane j + 16(i / 4).
  //     Wavefront w owns the tiles w, w + 3, ..., w + 18.
  const int wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;  // A row / B column, and k
  // E-forming role: thread (kq, m) for tid < 162, dR/dEU entry kq for 162 <= tid < 180
  const bool e_entry = tid < 162, e_eu = tid >= 162 && tid < 180;
  const int kq = e_entry ? tid / 9 : tid - 162, em = tid % 9;
  const int e_row = (kq / 6) * kECols;
  const int e_col = e_entry ? (kq % 6) * 54 + em * 6 : 324 + kq % 6;
  const float mBvN = -prm.BvN;
  typedef float floatx4 __attribute__((ext_vector_type(4)));
  floatx4 accj[kTilesPerWave];
#pragma unroll
  for (int t = 0; t < kTilesPerWave; t++) accj[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  __syncthreads();  // D complete
  for (int g = 0; g < 9; g++) {
    float* E = sM[g & 1];
    if (e_entry) {
      const float* Dg = &sD[g][0][kq];
#define DV(dd) Dg[(dd) * 18]
      const float b0 = sBas[g][0][em], b1 = sBas[g][1][em], b2 = sBas[g][2][em];
      const float b34 = sBas[g][3][em] + sBas[g][4][em];
      const float2 v01{DV(dUx) * b0 + DV(dGU00) * b1 + DV(dGU10) * b2 + mBvN * DV(dPx) * b34,
                       DV(dUy) * b0 + DV(dGU01) * b1 + DV(dGU11) * b2 + mBvN * DV(dPy) * b34};
      const float2 v23{DV(dP) * b0 + DV(dPx) * b1 + DV(dPy) * b2,
                       DV(dS00) * b0 + DV(dS00x) * b1 + DV(dS00y) * b2};
      const float2 v45{DV(dS01) * b0 + DV(dS01x) * b1 + DV(dS01y) * b2,
                       DV(dS11) * b0 + DV(dS11x) * b1 + DV(dS11y) * b2};
#undef DV
      float2* dst = reinterpret_cast<float2*>(&E[e_row + e_col]);  // e_col is even: 8-byte aligned
      dst[0] = v01;
      dst[1] = v23;
      dst[2] = v45;
    } else if (e_eu) {
      E[e_row + e_col] = sD[g][dEU][kq];
    }
    __syncthreads();
    // Phi_g^T: node li, basis kind lk (zero for the padding rows and the
    // fourth k); loads are unconditional from clamped addresses, so no lane
    // mask is switched around them.  B's k = 3 row meets A's zero column: it
    // is read from row 2 (finite), which adds exact zeros.
    const float av = sBas[g][min(lk, 2)][min(li, 8)];
    const float a = (li < 9 && lk < 3) ? av : 0.0f;
    float bv[kTilesPerWave];
#pragma unroll
    for (int t = 0; t < kTilesPerWave; t++) bv[t] = E[min(lk, 2) * kECols + (wave + 3 * t) * 16 + li];
#pragma unroll
    for (int t = 0; t < kTilesPerWave; t++) accj[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv[t], accj[t], 0, 0, 0);
  }
  // the element matrix to LDS (over D, whose last reads preceded the last
  // barrier): lane (li, lk) holds rows n = 4 lk + r of column (wave + 3 t) * 16 + li
  float* sJ = &sD[0][0][0];  // [n][330]
  static_assert(sizeof(sD) >= 9 * 330 * sizeof(float), "the element matrix fits in D's space");
#pragma unroll
  for (int t = 0; t < kTilesPerWave; t++) {
    const int col = (wave + 3 * t) * 16 + li;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int n = 4 * lk + r;
      if (n < 9 && col < 330) sJ[n * 330 + col] = accj[t][r];
    }
  }
  __syncthreads();
  // stores: one thread per (test node n, row q, trial node m): the 6-entry run
  // of row (n, q) in node m's block, stored into values (single or first of
  // two contributors), added there (second of two) or stored to cbank; then
  // the 54 entries of dR/dEU to acbank
  for (int it = tid; it < 9 * 6 * 9 + 54; it += kThreads) {
    if (it < 9 * 6 * 9) {
      const int n = it / 54, q = (it / 9) % 6, m = it % 9;
      const float2* src = reinterpret_cast<const float2*>(&sJ[n * 330 + q * 54 + m * 6]);  // even index
      const float2 v01 = src[0], v23 = src[1], v45 = src[2];
      const int32_t cd = sCode[n * 9 + m];
      double* dst = (cd & 3) == kCodeBank ? fo.cbank + ((int64_t)(cd >> 2) * 36 + q * 6)
                                          : fo.values + (sRb[n] + (int64_t)q * sRl[n] + (int64_t)(cd >> 2) * 6);
      // +0 for -0, as ADD_VALUES into zeroed storage
      double v[6] = {0.0 + (double)v01.x, 0.0 + (double)v01.y, 0.0 + (double)v23.x,
                     0.0 + (double)v23.y, 0.0 + (double)v45.x, 0.0 + (double)v45.y};
      if ((cd & 3) == kCodeAdd) {
#pragma unroll
        for (int i = 0; i < 6; i++) v[i] = dst[i] + v[i];
      }
#pragma unroll
      for (int i = 0; i < 6; i++) dst[i] = v[i];
    } else {
      const int i = it - 9 * 6 * 9, n = i / 6;
      fo.acbank[e * 54 + i] = 0.0 + (double)sJ[n * 330 + 324 + i % 6];
    }
  }
}

// ----------------------------------------------------------------------------
// Outlet edges: openBC traction (Residuals_mod.F90:599-624) at the 3 edge
// Gauss points, with setEdgeQuantities normals (Element_Calculations_mod.F90:887-916).
// Bank layout: ebank[(edge*18 + n*2 + i)*56 + k], rows (node n, momentum i).
// ----------------------------------------------------------------------------
// Unnormalised outward normal of edge ied (setEdgeQuantities, Quadrangle,
// Element_Calculations_mod.F90:887-905): edges 1 (eta = -1) and 3 (eta = +1)
// run along xi, 2 and 4 along eta; 1 and 2 take (t_y, -t_x), 3 and 4
// (-t_y, t_x) of the tangent t.  Written with selects: with a four-way switch
// on ied, hipcc (ROCm 7.2) produced wrong normals for edge 4 in the drag
// kernel although the optimised IR was right.
__device__ __forceinline__ void edge_normal(int ied, double dxdc, double dydc, double dxde, double dyde,
                                            double& n0, double& n1) {
  const bool along_xi = (ied & 1) != 0;
  const double tx = along_xi ? dxdc : dxde;
  const double ty = along_xi ? dydc : dyde;
  const bool first = ied <= 2;
  n0 = first ? ty : -ty;
  n1 = first ? -tx : tx;
}

struct EdgeGeo {
  double dFx[9], dFy[9];
  double n0, n1, WET;
};

__global__ void __launch_bounds__(192)
vef_edge_kernel(const double* __restrict__ coords, const int32_t* __restrict__ conn, const int32_t* __restrict__ edges,
                const double* __restrict__ x, int64_t eu_index, int32_t n_edges, DevParams prm,
                double* __restrict__ ebank) {
  __shared__ double sU[3 * 54];
  __shared__ double sX[18];
  __shared__ EdgeGeo sG[3];
  __shared__ double sRed[2][3][kNeval][2];
  const int tid = threadIdx.x;
  const int ed_idx = blockIdx.x;
  const int32_t e = edges[2 * ed_idx];
  const int ied = edges[2 * ed_idx + 1];
  const int ed = ied - 1;
  const double EU = x[eu_index];
  for (int j = tid; j < 54; j += blockDim.x) {
    const double u = x[(int64_t)conn[(int64_t)e * 9 + j / 6] * 6 + (j % 6)];
    const double up = u + kEpsJac;
    sU[j] = u;
    sU[108 + j] = up;
    sU[54 + j] = up - kEpsJac;
  }
  for (int c = tid; c < 18; c += blockDim.x) sX[c] = coords[(int64_t)conn[(int64_t)e * 9 + c / 2] * 2 + (c & 1)];
  __syncthreads();
  if (tid < 3) {
    const int g = tid;
    const double* dc = c_tab.dcE[ed][g];
    const double* de = c_tab.deE[ed][g];
    double dxdc = dc[0] * sX[0], dydc = dc[0] * sX[1], dxde = de[0] * sX[0], dyde = de[0] * sX[1];
    for (int n = 1; n < 9; n++) {
      dxdc = dxdc + dc[n] * sX[2 * n];
      dydc = dydc + dc[n] * sX[2 * n + 1];
      dxde = dxde + de[n] * sX[2 * n];
      dyde = dyde + de[n] * sX[2 * n + 1];
    }
    const double detJ = dxdc * dyde - dxde * dydc;
    const double dcdx = dyde / detJ, dedx = -(dydc / detJ), dcdy = -(dxde / detJ), dedy = dxdc / detJ;
    for (int n = 0; n < 9; n++) {
      sG[g].dFx[n] = dcdx * dc[n] + dedx * de[n];
      sG[g].dFy[n] = dcdy * dc[n] + dedy * de[n];
    }
    double n0, n1;
    edge_normal(ied, dxdc, dydc, dxde, dyde, n0, n1);
    const double dL = sqrt(n0 * n0 + n1 * n1);
    sG[g].n0 = n0 / dL;
    sG[g].n1 = n1 / dL;
    sG[g].WET = c_tab.W1[g] * dL;
  }
  __syncthreads();
  const bool active = tid < 3 * kNeval;
  const int g = active ? tid / kNeval : 0;
  const int k = tid % kNeval;
  double P00 = 0, P01 = 0, P10 = 0, P11 = 0, P = 0;
  if (active) {
    const int p = k - 1;
    const double EUk = (k == kNeval - 1) ? EU + kEpsJac : EU;
    const double* bf = c_tab.bfE[ed][g];
    auto v = [&](int n, int qq) {
      const int j = n * 6 + qq;
      const int sel = (j < p) ? 54 : ((j == p) ? 108 : 0);
      return sU[sel + j];
    };
    double GU00, GU10, GU01, GU11, S00, S01, S11;
    {
      const double fx = sG[g].dFx[0], fy = sG[g].dFy[0], b = bf[0];
      GU00 = v(0, 0) * fx; GU10 = v(0, 0) * fy; GU01 = v(0, 1) * fx; GU11 = v(0, 1) * fy;
      P = v(0, 2) * b; S00 = v(0, 3) * b; S01 = v(0, 4) * b; S11 = v(0, 5) * b;
    }
    for (int n = 1; n < 9; n++) {
      const double fx = sG[g].dFx[n], fy = sG[g].dFy[n], b = bf[n];
      GU00 = GU00 + v(n, 0) * fx; GU10 = GU10 + v(n, 0) * fy;
      GU01 = GU01 + v(n, 1) * fx; GU11 = GU11 + v(n, 1) * fy;
      P = P + v(n, 2) * b; S00 = S00 + v(n, 3) * b; S01 = S01 + v(n, 4) * b; S11 = S11 + v(n, 5) * b;
    }
    const double S10 = S01;
    const double WiN = prm.WiN0 * EUk;
    const double tr = GU00 + GU11;
    const double htr = (tr * 2.0) / (2.0 * 2.0);
    const double Gt00 = GU00 - htr, Gt01 = GU01, Gt10 = GU10, Gt11 = GU11 - htr;
    const double Gd00 = Gt00 + Gt00, Gd01 = Gt01 + Gt10, Gd10 = Gt10 + Gt01, Gd11 = Gt11 + Gt11;
    const double C00 = S00 * S00 + S01 * S10, C01 = S00 * S01 + S01 * S11;
    const double C10 = S10 * S00 + S11 * S10, C11 = S10 * S01 + S11 * S11;
    const double a = prm.omb / WiN;
    P00 = a * (C00 - 1.0) + prm.BvN * Gd00 - P;
    P01 = a * (C01 - 0.0) + prm.BvN * Gd01;
    P10 = a * (C10 - 0.0) + prm.BvN * Gd10;
    P11 = a * (C11 - 1.0) + prm.BvN * Gd11 - P;
  }
  for (int n = 0; n < 9; n++) {
    const int buf = n & 1;
    if (active) {
      const double bifn = c_tab.bfE[ed][g][n];
      const double n0 = sG[g].n0, n1 = sG[g].n1, WET = sG[g].WET;
      // TERM_RES(i) = -dot_product(normal, P_tot(:,i) + P*I(:,i))*bifn
      const double T0 = -((n0 * (P00 + P) + n1 * P10) * bifn);
      const double T1 = -((n0 * P01 + n1 * (P11 + P)) * bifn);
      sRed[buf][g][k][0] = T0 * WET;
      sRed[buf][g][k][1] = T1 * WET;
    }
    __syncthreads();
    if (tid < 2 * kNeval) {
      const int i = tid / kNeval, kk = tid % kNeval;
      const double acc = (sRed[buf][0][kk][i] + sRed[buf][1][kk][i]) + sRed[buf][2][kk][i];
      ebank[((int64_t)ed_idx * 18 + n * 2 + i) * kNeval + kk] = acc;
    }
  }
}

// ----------------------------------------------------------------------------
// Scatter (storeResidual / storeJacobian / storeJacobian_Ac with ADD_VALUES,
// Storage_mod.F90:10-115) as deterministic ordered sums.  Node pair (r, c) of
// the pattern has the list of bulk contributions (element e, local index jn
// of r, local index in of c), ascending in e, packed as e*81 + jn*9 + in, then
// the outlet-edge contributions packed as ed*81 + jn*9 + in, ascending in ed.
// ----------------------------------------------------------------------------

// Entries of node pairs shared by several elements: thread (shared pair su,
// jq, iq) sums the contributions of its elements in ascending element order.
// The contributions of pair su are cbank rows ustart[su] .. ustart[su+1]-1,
// already in that order; the first four loads are issued together.
__global__ void __launch_bounds__(256)
vef_shared_kernel(int64_t n_shared, const int64_t* __restrict__ sbase, const int32_t* __restrict__ slen,
                  const int32_t* __restrict__ ustart, const double* __restrict__ cbank, double* __restrict__ values) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_shared * 36) return;
  const int64_t su = t / 36;
  const int jq = (int)((t % 36) / 6), iq = (int)(t % 6);
  const int32_t u0 = ustart[su], u1 = ustart[su + 1];
  const double* c = cbank + (int64_t)u0 * 36 + jq * 6 + iq;
  const int32_t cnt = u1 - u0;  // >= 2
  const double c0 = c[0], c1 = c[36];
  const double c2 = cnt > 2 ? c[72] : 0.0, c3 = cnt > 3 ? c[108] : 0.0;
  double acc = (0.0 + c0) + c1;
  if (cnt > 2) acc = acc + c2;
  if (cnt > 3) acc = acc + c3;
  for (int32_t u = 4; u < cnt; u++) acc = acc + c[(int64_t)u * 36];
  values[sbase[su] + (int64_t)jq * slen[su] + iq] = acc;
}

// Outlet-edge Jacobian entries, added after every bulk contribution in
// ascending edge order: thread (edge pair ep, momentum row jq, iq).
__global__ void __launch_bounds__(256)
vef_edge_add_kernel(int64_t n_epairs, const int64_t* __restrict__ ep_pair, const int64_t* __restrict__ ebase,
                    const int32_t* __restrict__ elen, const int32_t* __restrict__ estart,
                    const int32_t* __restrict__ econtrib, const double* __restrict__ ebank,
                    double* __restrict__ values) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_epairs * 12) return;
  const int64_t ep = t / 12;
  const int jq = (int)((t % 12) / 6), iq = (int)(t % 6);
  const int64_t pp = ep_pair[ep];
  const int64_t slot = ebase[ep] + (int64_t)jq * elen[ep] + iq;
  double v = values[slot];
  for (int32_t c = estart[pp]; c < estart[pp + 1]; c++) {
    const int32_t w = econtrib[c];
    const int64_t ed = w / 81;
    const int jn = (w / 9) % 9, in = w % 9;
    const double* R = ebank + (ed * 18 + jn * 2 + jq) * kNeval;
    v = v + fd_quotient_ieee(R[1 + in * 6 + iq], R[0]);
  }
  values[slot] = v;
}

// Residual b and the dR/dEU column: one thread per node row, summing over the
// elements (then outlet edges) that contain the node, in ascending order.
__global__ void __launch_bounds__(256)
vef_rhs_kernel(int64_t n_node_rows, int jac, const int64_t* __restrict__ rowptr, const int64_t* __restrict__ diag_pair,
               const int32_t* __restrict__ cstart, const int32_t* __restrict__ contrib,
               const int32_t* __restrict__ estart, const int32_t* __restrict__ econtrib,
               const double* __restrict__ r0bank, const double* __restrict__ acbank, const double* __restrict__ ebank,
               double* __restrict__ values, double* __restrict__ b) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n_node_rows) return;
  const int64_t r = row / 6;
  const int jq = (int)(row % 6);
  const int64_t pp = diag_pair[r];
  double rb = 0.0, ra = 0.0;
  const int32_t c0 = cstart[pp];
  for (int32_t c = c0; c < cstart[pp + 1]; c++) {
    const int32_t w = contrib[c];
    const int64_t idx = (w / 81) * 54 + ((w / 9) % 9) * 6 + jq;
    rb = rb + r0bank[idx];
    if (jac) ra = ra + acbank[idx];
  }
  if (jq < 2) {
    for (int32_t c = estart[pp]; c < estart[pp + 1]; c++) {
      const int32_t w = econtrib[c];
      const int64_t ed = w / 81;
      const int jn = (w / 9) % 9;
      const double* R = ebank + (ed * 18 + jn * 2 + jq) * kNeval;
      rb = rb + R[0];
      if (jac) ra = ra + fd_quotient_ieee(R[kNeval - 1], R[0]);
    }
  }
  b[row] = rb;
  if (jac) values[rowptr[row + 1] - 1] = ra;
}

// ----------------------------------------------------------------------------
// Post-processing: drag on a boundary (calculateEdgeQuantities with ieq_ed = 1,
// src/Post_Process_mod.F90:1796-1890) and the nodal stress tensor
// (nodalStressTensorSQRT, :2176-2236).
// ----------------------------------------------------------------------------
constexpr int kDragEdgesPerBlock = 21;  // 63 lanes = 21 edges x 3 edge Gauss points

// Lane (edge, g): dot_product(-normal, P_tot(:,1))*WET at edge Gauss point g,
// then lane g = 0 of each edge forms term_el = ((0 + t0) + t1) + t2.
__global__ void __launch_bounds__(64)
vef_drag_edge_kernel(const double* __restrict__ coords, const int32_t* __restrict__ conn,
                     const int32_t* __restrict__ edges, const double* __restrict__ x, int64_t eu_index,
                     int32_t n_edges, DevParams prm, double* __restrict__ term_el) {
  __shared__ double sT[kDragEdgesPerBlock][3];
  const int tid = threadIdx.x;
  const int le = tid / 3, g = tid % 3;
  const int64_t k = (int64_t)blockIdx.x * kDragEdgesPerBlock + le;
  const bool valid = tid < 3 * kDragEdgesPerBlock && k < n_edges;
  if (valid) {
    const int32_t e = edges[2 * k];
    const int ied = edges[2 * k + 1];
    const int ed = ied - 1;
    double X[18], u[54];
    for (int n = 0; n < 9; n++) {
      const int64_t gn = conn[(int64_t)e * 9 + n];
      X[2 * n] = coords[gn * 2];
      X[2 * n + 1] = coords[gn * 2 + 1];
      for (int q = 0; q < 6; q++) u[n * 6 + q] = x[gn * 6 + q];
    }
    const double* dc = c_tab.dcE[ed][g];
    const double* de = c_tab.deE[ed][g];
    const double* bf = c_tab.bfE[ed][g];
    double dxdc = dc[0] * X[0], dydc = dc[0] * X[1], dxde = de[0] * X[0], dyde = de[0] * X[1];
    for (int n = 1; n < 9; n++) {
      dxdc = dxdc + dc[n] * X[2 * n];
      dydc = dydc + dc[n] * X[2 * n + 1];
      dxde = dxde + de[n] * X[2 * n];
      dyde = dyde + de[n] * X[2 * n + 1];
    }
    const double detJ = dxdc * dyde - dxde * dydc;
    const double dcdx = dyde / detJ, dedx = -(dydc / detJ), dcdy = -(dxde / detJ), dedy = dxdc / detJ;
    double GU00 = 0, GU10 = 0, GU01 = 0, GU11 = 0, P = 0, S00 = 0, S01 = 0, S11 = 0;
    for (int n = 0; n < 9; n++) {
      const double fx = dcdx * dc[n] + dedx * de[n];
      const double fy = dcdy * dc[n] + dedy * de[n];
      const double b = bf[n];
      if (n == 0) {
        GU00 = u[0] * fx; GU10 = u[0] * fy; GU01 = u[1] * fx; GU11 = u[1] * fy;
        P = u[2] * b; S00 = u[3] * b; S01 = u[4] * b; S11 = u[5] * b;
      } else {
        GU00 = GU00 + u[n * 6] * fx; GU10 = GU10 + u[n * 6] * fy;
        GU01 = GU01 + u[n * 6 + 1] * fx; GU11 = GU11 + u[n * 6 + 1] * fy;
        P = P + u[n * 6 + 2] * b; S00 = S00 + u[n * 6 + 3] * b;
        S01 = S01 + u[n * 6 + 4] * b; S11 = S11 + u[n * 6 + 5] * b;
      }
    }
    const double S10 = S01;
    const double WiN = prm.WiN0 * x[eu_index];
    const double tr = GU00 + GU11;
    const double htr = (tr * 2.0) / (2.0 * 2.0);
    const double Gt00 = GU00 - htr, Gt10 = GU10, Gt01 = GU01;
    const double Gd00 = Gt00 + Gt00, Gd10 = Gt10 + Gt01;
    const double C00 = S00 * S00 + S01 * S10;
    const double C10 = S10 * S00 + S11 * S10;
    const double a = prm.omb / WiN;
    const double P00 = a * (C00 - 1.0) + prm.BvN * Gd00 - P;
    const double P10 = a * (C10 - 0.0) + prm.BvN * Gd10;
    double n0, n1;
    edge_normal(ied, dxdc, dydc, dxde, dyde, n0, n1);
    const double dL = sqrt(n0 * n0 + n1 * n1);
    n0 = n0 / dL;
    n1 = n1 / dL;
    const double WET = c_tab.W1[g] * dL;
    const double d = (0.0 + (-n0) * P00) + (-n1) * P10;  // the third term, (-0)*0, adds nothing
    sT[le][g] = d * WET;
  }
  __syncthreads();
  if (valid && g == 0) term_el[k] = ((0.0 + sT[le][0]) + sT[le][1]) + sT[le][2];
}

// term = term + term_el over the edges in order (one thread; a few hundred edges).
__global__ void vef_drag_sum_kernel(int32_t n_edges, const double* __restrict__ term_el, double* __restrict__ Fd) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double t = 0.0;
  for (int32_t k = 0; k < n_edges; k++) t = t + term_el[k];
  Fd[0] = t;
}

// T = ((1-BvN)/WiN)*(C - I), C = S.S summed from zero over k = 1..3 (g1 = g2 = 1).
__global__ void __launch_bounds__(256)
vef_nodal_stress_kernel(int64_t n_nodes, const double* __restrict__ x, double omb, double WiN,
                        double* __restrict__ T) {
  const int64_t gnod = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gnod >= n_nodes) return;
  const double S00 = x[gnod * 6 + 3], S01 = x[gnod * 6 + 4], S11 = x[gnod * 6 + 5];
  const double S10 = S01;
  const double C00 = (0.0 + S00 * S00) + S01 * S10;
  const double C01 = (0.0 + S00 * S01) + S01 * S11;
  const double C11 = (0.0 + S10 * S01) + S11 * S11;
  const double a = omb / WiN;
  T[gnod * 4 + 0] = a * (1.0 * C00 - 1.0 * 1.0);
  T[gnod * 4 + 1] = a * (1.0 * C01 - 1.0 * 0.0);
  T[gnod * 4 + 2] = a * (1.0 * C11 - 1.0 * 1.0);
  T[gnod * 4 + 3] = a * (1.0 * 0.0 - 1.0 * 1.0);
}

// ----------------------------------------------------------------------------
// Dirichlet rows (applyDirichletBC / applyDirichletBCFullyDeveloped,
// src/Dirichlet_mod.F90:104-190): b[r] = x[r] - value and, with the Jacobian,
// MatZeroRows with diagonal 1.  One wavefront per row; the rows are distinct.
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
vef_dirichlet_kernel(int64_t n_rows, const int64_t* __restrict__ rows, const double* __restrict__ target,
                     const double* __restrict__ x, int jac, const int64_t* __restrict__ rowptr,
                     const int32_t* __restrict__ col, double* __restrict__ b, double* __restrict__ values) {
  const int64_t i = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (i >= n_rows) return;
  const int64_t r = rows[i];
  if (lane == 0) b[r] = x[r] - target[i];
  if (!jac) return;
  for (int64_t p = rowptr[r] + lane; p < rowptr[r + 1]; p += 64) values[p] = (col[p] == r) ? 1.0 : 0.0;
}

// ----------------------------------------------------------------------------
// Linear stability assembly (Cylinder_Stability_2D): A and B of the
// generalised eigenproblem by forward differences in the disturbance
// (storeBulkStability, src/Stability/Equations_Stability_mod.F90:135-226).
//
// Everything that depends on the base state alone is computed once per
// element by vef_stab_prep_kernel: the Gauss-point geometry, the base
// quantities (base_point) and, per Gauss point and test node, the test-node
// coefficients (node_coef).  The reference recomputes them for each of the 54
// disturbances; they do not depend on the disturbance, so the values are the
// same bits.  Per element the prep block holds, in doubles:
//   [0, 288)      Base of Gauss points 0..8 (32 each)
//   [288, 648)    per Gauss point: dFx[9], dFy[9], d2x[9], d2y[9], WET, 3 pad
//   [648, 1944)   NodeCoef of (test node n, Gauss point g) at 648 + 16*(9n + g)
// ----------------------------------------------------------------------------
constexpr int kPrepBase = 0, kPrepGeo = 9 * 32, kPrepCoef = kPrepGeo + 9 * 40;
constexpr int kPrepPerElement = kPrepCoef + 81 * 16;
static_assert(sizeof(vefstab::Base) == 32 * sizeof(double), "Base is 32 doubles");
static_assert(sizeof(vefstab::NodeCoef) == 16 * sizeof(double), "NodeCoef is 16 doubles");

// Seven elements per wavefront, lane (element el, Gauss point g).
__global__ void __launch_bounds__(64)
vef_stab_prep_kernel(const double* __restrict__ coords, const int32_t* __restrict__ conn,
                     const double* __restrict__ x, int64_t eu_index, int32_t n_el, vefstab::Params prm,
                     double* __restrict__ prep) {
  __shared__ double sU[7][54];
  __shared__ double sX[7][18];
  __shared__ Geo sG[63];
  const int tid = threadIdx.x;
  const int64_t e0 = (int64_t)blockIdx.x * 7;
  const int64_t elast = (int64_t)n_el - 1;
  // All of a lane's gather loads are issued before the first is used (indices
  // clamped into range, only the LDS writes predicated).
  {
    constexpr int NU = 7 * 54, NX = 7 * 18, IU = (NU + 63) / 64, IX = (NX + 63) / 64;
    int32_t gu[IU], gx[IX];
#pragma unroll
    for (int i = 0; i < IU; i++) {
      const int t = min(tid + 64 * i, NU - 1);
      const int64_t e = e0 + t / 54;
      gu[i] = conn[(e < elast ? e : elast) * 9 + (t % 54) / 6];
    }
#pragma unroll
    for (int i = 0; i < IX; i++) {
      const int t = min(tid + 64 * i, NX - 1);
      const int64_t e = e0 + t / 18;
      gx[i] = conn[(e < elast ? e : elast) * 9 + (t % 18) / 2];
    }
    double vu[IU], vx[IX];
#pragma unroll
    for (int i = 0; i < IU; i++) vu[i] = x[(int64_t)gu[i] * 6 + (min(tid + 64 * i, NU - 1) % 54) % 6];
#pragma unroll
    for (int i = 0; i < IX; i++) vx[i] = coords[(int64_t)gx[i] * 2 + (min(tid + 64 * i, NX - 1) & 1)];
#pragma unroll
    for (int i = 0; i < IU; i++) {
      const int t = tid + 64 * i;
      if (t < NU) sU[t / 54][t % 54] = vu[i];
    }
#pragma unroll
    for (int i = 0; i < IX; i++) {
      const int t = tid + 64 * i;
      if (t < NX) sX[t / 18][t % 18] = vx[i];
    }
  }
  __syncthreads();
  const int el = tid / 9, g = tid % 9;
  if (tid >= 63 || e0 + el >= n_el) return;  // no barrier follows
  const double EU = x[eu_index];
  Geo& G = sG[tid];
  geometry(sX[el], c_tab.dc[g], c_tab.de[g], c_tab.dcc[g], c_tab.dee[g], c_tab.dce[g], c_tab.W2[g], G);
  vefstab::Base b;
  vefstab::base_point(sU[el], c_tab.bf[g], G.dFx, G.dFy, G.H2, EU, prm, b);
  double* dst = prep + (e0 + el) * kPrepPerElement;
  const double* bs = reinterpret_cast<const double*>(&b);
#pragma unroll
  for (int i = 0; i < 32; i += 2)
    *reinterpret_cast<double2*>(dst + kPrepBase + g * 32 + i) = make_double2(bs[i], bs[i + 1]);
  double* geo = dst + kPrepGeo + g * 40;
#pragma unroll
  for (int n = 0; n < 9; n++) {
    geo[n] = G.dFx[n];
    geo[9 + n] = G.dFy[n];
    geo[18 + n] = G.d2x[n];
    geo[27 + n] = G.d2y[n];
  }
  geo[36] = G.WET;
  geo[37] = 0.0;
  geo[38] = 0.0;
  geo[39] = 0.0;
  for (int n = 0; n < 9; n++) {
    vefstab::NodeCoef c;
    vefstab::node_coef(b, c_tab.bf[g][n], G.dFx[n], G.dFy[n], prm, c);
    const double* cs = reinterpret_cast<const double*>(&c);
    double* cd = dst + kPrepCoef + (n * 9 + g) * 16;
#pragma unroll
    for (int i = 0; i < 16; i += 2) *reinterpret_cast<double2*>(cd + i) = make_double2(cs[i], cs[i + 1]);
  }
}

// (dRes - Res)/eps_jac = ((r*h)/(h*h)) with Res = 0: the two-FMA division by
// the shared reciprocal of h*h (correctly rounded while |q0| lies in
// [2^-901, 2^900) or r*h = +-0; h*h = 1e-16 is inside [2^-60, 2^40]), the IEEE
// division for any other lane.
__device__ __forceinline__ double stab_quotient(double r) {
  constexpr double hh = vefstab::kEpsJac * vefstab::kEpsJac;
  constexpr double y = 1.0 / hh;
  const double a = r * vefstab::kEpsJac;
  const double q0 = a * y;
  const double m = fabs(q0);
  if (!((m >= 0x1p-901 && m < 0x1p900) || a == 0.0)) return a / hh;
  const double e = fma(hh, q0, -a);
  return fma(-e, y, q0);
}

// The forward differences: eight single-wavefront workgroups per element, ids
// 64m + 8j + x for element 8m + x and group j (one XCD per element); slot s of
// 7 is perturbation p = 7j + s < 54 of unknown (p/6, p%6).  The base and
// geometry part of the element's prep block is copied to LDS; lane (g, s)
// forms the disturbance quantities of Gauss point g (Dist) into LDS; then
// lane (n, s) forms the 6 A and 6 B terms of test node n at each Gauss point
// from the coefficients, read from the prep block in global memory (seven
// lanes read each record, and the element's eight workgroups share it through
// the caches; staged in LDS instead, the nine records a load instruction
// touches sit 128 bytes apart, in the same banks), and sums them over the
// Gauss points in order (Res = Res + TERM*WET, ig = 1..9) in registers,
// giving Res(eps e_p) and the quotient Res/eps (the zero disturbance has zero
// residuals), stored like the base Jacobian: straight into the values of node
// pairs with one contributing element, else into the element's bank rows for
// vef_shared_kernel.  The stability pattern is the base one without the
// arclength column (the last of each row) and row, so the codes of the base
// scatter map apply with the stability row pointers.
__global__ void __launch_bounds__(64)
vef_stab_kernel(const double* __restrict__ prep, const int32_t* __restrict__ conn, int32_t n_el, vefstab::Params prm,
                const int32_t* __restrict__ code, const int64_t* __restrict__ rowptr_s, double* __restrict__ A,
                double* __restrict__ B, double* __restrict__ cbank_a, double* __restrict__ cbank_b) {
  static_assert(sizeof(vefstab::Dist) == 24 * sizeof(double), "Dist is 24 doubles");
  // Base and geometry while the disturbances are formed, then the disturbances
  // of (g, s) at stride kDist (26 doubles: the seven records one load
  // instruction reads start in distinct banks).
  constexpr int kDist = 26;
  __shared__ __align__(16) union {
    double pre[kPrepCoef];  // Base[9] then geometry[9][40]
    double dist[63 * kDist];
  } sW;
  __shared__ double sT[9][3];  // tlsme, WiN, WET per Gauss point
  __shared__ int32_t sCode[81];
  __shared__ int64_t sRb[9];
  __shared__ int32_t sRl[9];
  const int tid = threadIdx.x;
  const int64_t blk = blockIdx.x;
  const int64_t e = (blk >> 6) * 8 + (blk & 7);
  const int kg = (int)((blk >> 3) & 7);
  if (e >= n_el) return;  // uniform over the workgroup
  const double2* src = reinterpret_cast<const double2*>(prep + e * kPrepPerElement);
  {
    // all loads first, then the LDS writes
    constexpr int NP = kPrepCoef / 2, IP = (NP + 63) / 64;
    double2 pv[IP];
#pragma unroll
    for (int i = 0; i < IP; i++) pv[i] = src[min(tid + 64 * i, NP - 1)];
    const int32_t c0 = code[e * 81 + tid], c1 = code[e * 81 + min(tid + 64, 80)];
#pragma unroll
    for (int i = 0; i < IP; i++)
      if (tid + 64 * i < NP) reinterpret_cast<double2*>(sW.pre)[tid + 64 * i] = pv[i];
    sCode[tid] = c0;
    if (tid + 64 < 81) sCode[tid + 64] = c1;
  }
  const vefstab::NodeCoef* gK = reinterpret_cast<const vefstab::NodeCoef*>(prep + e * kPrepPerElement + kPrepCoef);
  if (tid < 9) {
    const int64_t r6 = (int64_t)conn[e * 9 + tid] * 6;
    const int64_t rb = rowptr_s[r6];
    sRb[tid] = rb;
    sRl[tid] = (int32_t)(rowptr_s[r6 + 1] - rb);
  }
  __syncthreads();
  const int hi = tid / 7, s = tid % 7;  // (g, s) in the first phase, (n, s) in the second
  const int p = kg * 7 + s;
  const bool valid = tid < 63 && p < 54;
  // The first two coefficient records of test node hi, loaded before the
  // disturbance is formed; each later one is loaded a Gauss point ahead.
  vefstab::NodeCoef kc, kn;
  if (valid) {
    kc = gK[hi * 9];
    kn = gK[hi * 9 + 1];
  }
  vefstab::Dist D;
  if (valid) {
    const vefstab::Base& b = reinterpret_cast<const vefstab::Base*>(sW.pre + kPrepBase)[hi];
    const double* geo = sW.pre + kPrepGeo + hi * 40;
    vefstab::disturbance(b, p / 6, p % 6, c_tab.bf[hi], geo, geo + 9, geo + 18, geo + 27, prm, D);
  }
  if (tid < 9) {
    sT[tid][0] = reinterpret_cast<const vefstab::Base*>(sW.pre + kPrepBase)[tid].tlsme;
    sT[tid][1] = reinterpret_cast<const vefstab::Base*>(sW.pre + kPrepBase)[tid].WiN;
    sT[tid][2] = sW.pre[kPrepGeo + tid * 40 + 36];
  }
  __syncthreads();  // the base and geometry have been read
  if (valid) *reinterpret_cast<vefstab::Dist*>(sW.dist + (hi * 7 + s) * kDist) = D;
  __syncthreads();
  if (!valid) return;  // no barrier follows
  const int n = hi;
  double ra[6], rb[6];
  {
    double a[6], bb[6];
    auto dist = [&](int g) -> const vefstab::Dist& {
      return *reinterpret_cast<const vefstab::Dist*>(sW.dist + (g * 7 + s) * kDist);
    };
    vefstab::node_terms_c(kc, dist(0), sT[0][0], sT[0][1], sT[0][2], prm, ra, rb);
#pragma unroll
    for (int g = 1; g < 9; g++) {
      kc = kn;
      if (g < 8) kn = gK[n * 9 + g + 1];
      vefstab::node_terms_c(kc, dist(g), sT[g][0], sT[g][1], sT[g][2], prm, a, bb);
#pragma unroll
      for (int c = 0; c < 6; c++) {
        ra[c] = ra[c] + a[c];
        rb[c] = rb[c] + bb[c];
      }
    }
  }
  const int in = p / 6, iq = p % 6;
  const int32_t cd = sCode[n * 9 + in];
  if ((cd & 1) == 0) {
    const int64_t o = sRb[n] + (int64_t)(cd >> 1) * 6 + iq;
    const int64_t rl = sRl[n];
#pragma unroll
    for (int c = 0; c < 6; c++) {
      A[o + c * rl] = 0.0 + stab_quotient(ra[c]);
      B[o + c * rl] = 0.0 + stab_quotient(rb[c]);
    }
  } else {
    const int64_t o = (int64_t)(cd >> 1) * 36 + iq;
#pragma unroll
    for (int c = 0; c < 6; c++) {
      cbank_a[o + c * 6] = stab_quotient(ra[c]);
      cbank_b[o + c * 6] = stab_quotient(rb[c]);
    }
  }
}

// applyDirichletBCStability (src/Dirichlet_mod.F90:287-316): MatZeroRows of A
// with diagonal 1 and of B with diagonal 0, a wavefront per row.  Columns are
// read from the base pattern, whose row r holds the stability row's columns
// followed by the arclength column.
__global__ void __launch_bounds__(256)
vef_stab_dirichlet_kernel(int64_t n_rows, const int64_t* __restrict__ rows, const int64_t* __restrict__ rowptr_s,
                          const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col, double* __restrict__ A,
                          double* __restrict__ B) {
  const int64_t i = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (i >= n_rows) return;
  const int64_t r = rows[i];
  const int64_t ps = rowptr_s[r], len = rowptr_s[r + 1] - ps, pb = rowptr[r];
  for (int64_t k = lane; k < len; k += 64) {
    A[ps + k] = (col[pb + k] == r) ? 1.0 : 0.0;
    B[ps + k] = 0.0;
  }
}

// ----------------------------------------------------------------------------
// Arclength row (arclength / pseudoArclength, src/Extra_Equations_mod.F90:698-828)
// ----------------------------------------------------------------------------
// Increment <= 2: MatZeroRowsColumns(N-1) with diagonal 1.  Column N-1 is the
// last stored entry of every row; row N-1 is dense.
__global__ void vef_arc_pin_kernel(int64_t N, const int64_t* __restrict__ rowptr, double* __restrict__ values) {
  const int64_t last = rowptr[N - 1];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    if (i < N - 1) values[rowptr[i + 1] - 1] = 0.0;
    values[last + i] = (i == N - 1) ? 1.0 : 0.0;
  }
}

__global__ void vef_arc_pin_rhs_kernel(int64_t N, const double* __restrict__ x, int32_t increment, double LN0,
                                       double dArN, double* __restrict__ b) {
  if (threadIdx.x == 0 && blockIdx.x == 0) b[N - 1] = x[N - 1] - (LN0 + (double)(increment - 1) * dArN);
}

constexpr int kArcChunk = 4096;  // nodes per partial sum
// Largest number of node chunks vef_arclength_noalloc handles (16,777,216 nodes).
constexpr int kArcScratchBlocks = 4096;

// Per-field partial sums of t*(x - xo) over fixed node chunks; fixed
// reduction tree, so the result is deterministic.
__global__ void __launch_bounds__(256)
vef_arc_dot_kernel(int64_t n_nodes, const double* __restrict__ x, const double* __restrict__ xo,
                   const double* __restrict__ t, double* __restrict__ partials) {
  __shared__ double red[6][256];
  const int64_t g0 = (int64_t)blockIdx.x * kArcChunk;
  const int64_t g1 = (g0 + kArcChunk < n_nodes) ? g0 + kArcChunk : n_nodes;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t g = g0 + threadIdx.x; g < g1; g += 256)
#pragma unroll
    for (int q = 0; q < 6; q++) acc[q] = acc[q] + t[g * 6 + q] * (x[g * 6 + q] - xo[g * 6 + q]);
#pragma unroll
  for (int q = 0; q < 6; q++) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int q = 0; q < 6; q++) red[q][threadIdx.x] = red[q][threadIdx.x] + red[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 6) partials[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = red[threadIdx.x][0];
}

__global__ void vef_arc_rhs_kernel(int64_t N, int nblk, const double* __restrict__ partials,
                                   const double* __restrict__ x, const double* __restrict__ xo,
                                   const double* __restrict__ t, double dSo, double* __restrict__ b) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double value = 0.0;
  for (int q = 0; q < 6; q++) {
    double d = 0.0;
    for (int i = 0; i < nblk; i++) d = d + partials[(int64_t)q * nblk + i];
    value = value + d;
  }
  b[N - 1] = value + t[N - 1] * (x[N - 1] - xo[N - 1]) - dSo;
}

__global__ void vef_copy_kernel(double* __restrict__ dst, const double* __restrict__ src, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void vef_fill_kernel(double* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

}  // namespace

// Scratch of vef_arclength_noalloc: the per-field partial sums of the
// pseudo-arclength dot product (6 * kArcScratchBlocks doubles = 192 KiB), one
// copy per GPU the library runs on, held until the process ends.  Host code
// never passes `scratch` itself to a kernel: it uses the device address that
// vef_create resolves once per device (g_scratch below).
static __device__ __align__(256) double scratch[6 * kArcScratchBlocks];

// =============================================================================
// Host side
// =============================================================================
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch[kMaxDevices] = {};   // device address of `scratch`, per device
bool g_scratch_done[kMaxDevices] = {};

// Resolves the current device's address of `scratch` once (the caller has
// that device current).  nullptr if it cannot.
double* scratch_address(int* device) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  *device = dev;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_done[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch[dev] = (double*)ptr;
    else (void)hipGetLastError();
    g_scratch_done[dev] = true;
  }
  return g_scratch[dev];
}
}  // namespace

struct vef_handle {
  int32_t n_nodes = 0, n_el = 0, n_edges = 0;
  int64_t N = 0, nnz = 0;
  vef_params params{};
  DevParams dprm{};
  std::vector<int64_t> h_rowptr;
  std::vector<int32_t> h_col;
  // device
  double* d_coords = nullptr;
  int32_t* d_conn = nullptr;
  int32_t* d_edges = nullptr;
  int64_t* d_rowptr = nullptr;
  int32_t* d_col = nullptr;
  int64_t* d_diag_pair = nullptr;
  int64_t n_pairs = 0;
  // Scatter of the Jacobian passes (FdOut codes): direct stores, in-place
  // sums of two-contributor entries over the element colours, and the ordered
  // sums of the entries of three or more elements
  int32_t* d_code_fd = nullptr;  // [n_el][81]
  int32_t* d_elist = nullptr;    // elements by colour, ascending within a colour
  std::vector<int32_t> colour_off;  // colour c: d_elist[colour_off[c] .. colour_off[c+1])
  int64_t* d_sbase = nullptr;    // pairs of three or more elements: first slot
  int32_t* d_slen = nullptr;     //   row length
  int32_t* d_ustart_m = nullptr; //   first cbank row
  int64_t n_shared_m = 0, n_cidx_m = 0;
  // Scatter of the stability assembly: every shared pair through its banks
  int32_t* d_code = nullptr;     // [n_el][81]: pos<<1, or (cidx<<1)|1 for a shared pair
  int32_t* d_ustart = nullptr;   //   first bank row of each shared pair
  int64_t n_shared = 0, n_cidx = 0;
  int64_t* d_ep_pair = nullptr;  // pairs with outlet-edge contributions
  int64_t* d_ebase = nullptr;
  int32_t* d_elen = nullptr;
  int64_t n_epairs = 0;
  double* d_r0bank = nullptr;    // [n_el][54] base residual rows
  double* d_geob = nullptr;      // [n_el][9] Geo records of the residual pass, for the Jacobian pass
  double* d_acbank = nullptr;    // [n_el][54] dR/dEU quotients
  double* d_cbank = nullptr;     // [n_cidx_m][36] quotients of entries of three or more elements
  int32_t* d_cstart = nullptr;
  int32_t* d_contrib = nullptr;
  int32_t* d_estart = nullptr;
  int32_t* d_econtrib = nullptr;
  double* d_ebank = nullptr;  // n_edges*18*56
  double* d_arc = nullptr;    // per-field partial sums of the arclength dot product
  int64_t* d_bc_rows = nullptr;     // Dirichlet rows, distinct
  double* d_bc_values = nullptr;    // their target values
  int64_t n_bc = 0;
  int64_t bc_cap = 0;               // capacity of the two buffers
  // linear stability assembly (allocated on first use)
  std::vector<int64_t> h_sbase_s;   // shared pairs: first slot in the stability pattern
  std::vector<int32_t> h_slen_s;    //   its row length there
  bool stab_ready = false;
  int64_t nnz_s = 0;
  int64_t* d_rowptr_s = nullptr;
  int64_t* d_sbase_s = nullptr;
  int32_t* d_slen_s = nullptr;
  double* d_cbank_a = nullptr;
  double* d_cbank_b = nullptr;
  double* d_stab_prep = nullptr;  // per element: base, geometry and test-node coefficients
  int64_t* d_sbc_rows = nullptr;    // stability Dirichlet rows, distinct
  int64_t n_sbc = 0;
  int32_t* d_drag_edges = nullptr;  // boundary edges of the drag integral, [n][2]
  double* d_drag_terms = nullptr;   // term_el per edge
  int32_t n_drag_edges = 0;
  int arc_blocks = 0;
  int device = -1;                   // the device vef_create ran on
  double* arc_scratch = nullptr;     // that device's address of `scratch` (vef_arclength_noalloc)
  double* d_tmp_x = nullptr;  // host-assemble staging
  double* d_tmp_b = nullptr;
  double* d_tmp_v = nullptr;
  int last_flag = -1;
  std::string err;
  // Optional kernel timing: event pairs around the element kernel and around
  // the scatter kernels of each assembly.
  bool profiling = false;
  std::vector<hipEvent_t> ev_bulk, ev_scatter, ev_other;  // start/stop pairs
};

namespace {
std::string g_create_err;

int set_err(vef_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  else g_create_err = msg;
  return code;
}

#define VEF_HIP(h, call)                                                                  \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return set_err(h, VEF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int check_params(vef_handle* h, const vef_params* p) {
  if (!p) return set_err(h, VEF_ERR_ARG, "params is NULL");
  if (p->model != VEF_MODEL_LPTT && p->model != VEF_MODEL_GIESEKUS)
    return set_err(h, VEF_ERR_UNSUPPORTED, "model must be VEF_MODEL_LPTT or VEF_MODEL_GIESEKUS");
  if (p->reserved != 0) return set_err(h, VEF_ERR_ARG, "params.reserved must be 0");
  if (!(p->WiN0 != 0.0) || !std::isfinite(p->WiN0)) return set_err(h, VEF_ERR_ARG, "WiN0 must be finite and nonzero");
  return VEF_OK;
}

DevParams make_dev_params(const vef_params& p) {
  DevParams d;
  d.ReN = p.ReN;
  d.BvN = p.BvN;
  d.VeN = p.VeN;
  d.DfN = p.DfN;
  d.WiN0 = p.WiN0;
  d.dt = DBL_MAX;
  d.omb = 1.0 - p.BvN;
  d.model = p.model;
  d.dcs_on = (p.DfN > 1.0e-12) ? 0 : 1;
  return d;
}

template <class T>
int dev_upload(vef_handle* h, T** dst, const T* src, size_t n) {
  if (n == 0) n = 1;
  VEF_HIP(h, hipMalloc((void**)dst, n * sizeof(T)));
  if (src) VEF_HIP(h, hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return VEF_OK;
}

int upload_tables(vef_handle* h) {
  // Gauss tables live in each device's constant memory; upload on the current device.
  GaussTables T;
  std::memset(&T, 0, sizeof(T));
  build_gauss_tables(T);
  VEF_HIP(h, hipMemcpyToSymbol(HIP_SYMBOL(c_tab), &T, sizeof(T)));
  return VEF_OK;
}
}  // namespace

namespace {
template <int MODEL, bool DCS, bool INERTIA>
void launch_bulk_t(vef_handle* h, int jac, double* values, const double* x, int64_t eu, hipStream_t st, double* geo) {
  FdOut fo{h->d_r0bank, h->d_code_fd, h->d_rowptr, values, h->d_cbank, h->d_acbank, geo, nullptr, 0};
  if (jac) {
    // one launch per colour, in colour order; 8 workgroups per element, the
    // colour's elements padded to a multiple of 8 (XCD-aware ids)
    for (size_t c = 0; c + 1 < h->colour_off.size(); c++) {
      fo.elist = h->d_elist + h->colour_off[c];
      fo.n_list = h->colour_off[c + 1] - h->colour_off[c];
      hipLaunchKernelGGL((vef_bulk_kernel<1, 7, kNeval, MODEL, DCS, INERTIA>),
                         dim3((unsigned)((fo.n_list + 7) / 8) * 64), dim3(64), 0, st, h->d_coords, h->d_conn, x, eu,
                         h->n_el, h->dprm, (double*)nullptr, fo);
    }
  } else
    hipLaunchKernelGGL((vef_bulk_kernel<7, 1, 1, MODEL, DCS, INERTIA>), dim3((unsigned)(h->n_el + 6) / 7), dim3(64), 0,
                       st, h->d_coords, h->d_conn, x, eu, h->n_el, h->dprm, h->d_r0bank, fo);
}

template <int MODEL, bool DCS>
void launch_bulk_i(vef_handle* h, int jac, double* values, const double* x, int64_t eu, hipStream_t st, double* geo) {
  if (h->dprm.ReN != 0.0) launch_bulk_t<MODEL, DCS, true>(h, jac, values, x, eu, st, geo);
  else launch_bulk_t<MODEL, DCS, false>(h, jac, values, x, eu, st, geo);
}

template <int MODEL>
void launch_bulk_d(vef_handle* h, int jac, double* values, const double* x, int64_t eu, hipStream_t st, double* geo) {
  if (h->dprm.dcs_on) launch_bulk_i<MODEL, true>(h, jac, values, x, eu, st, geo);
  else launch_bulk_i<MODEL, false>(h, jac, values, x, eu, st, geo);
}

// jac = 0: the residual pass, which writes the geometry bank when geo is not
// NULL; jac = 1: the Jacobian pass, which reads it (geo must be the bank the
// residual pass of the same assembly wrote).
void launch_bulk(vef_handle* h, int jac, double* values, const double* x, int64_t eu, hipStream_t st, double* geo) {
  if (h->dprm.model == 0) launch_bulk_d<0>(h, jac, values, x, eu, st, geo);
  else launch_bulk_d<1>(h, jac, values, x, eu, st, geo);
}
}  // namespace

extern "C" {

const char* vef_last_error(const vef_handle* h) { return h ? h->err.c_str() : g_create_err.c_str(); }

int64_t vef_num_rows(const vef_handle* h) { return h ? h->N : -1; }
int64_t vef_nnz(const vef_handle* h) { return h ? h->nnz : -1; }
const int64_t* vef_device_rowptr(const vef_handle* h) { return h ? h->d_rowptr : nullptr; }
const int32_t* vef_device_colind(const vef_handle* h) { return h ? h->d_col : nullptr; }

int vef_get_pattern(const vef_handle* h, int64_t* rowptr, int32_t* colind) {
  if (!h || !rowptr || !colind) return VEF_ERR_ARG;
  std::memcpy(rowptr, h->h_rowptr.data(), sizeof(int64_t) * (h->N + 1));
  std::memcpy(colind, h->h_col.data(), sizeof(int32_t) * h->nnz);
  return VEF_OK;
}

void vef_destroy(vef_handle* h) {
  if (!h) return;
  for (hipEvent_t e : h->ev_other) (void)hipEventDestroy(e);
  for (hipEvent_t e : h->ev_scatter) (void)hipEventDestroy(e);
  void* ptrs[] = {h->d_coords, h->d_conn,   h->d_edges,  h->d_rowptr,  h->d_col,     h->d_diag_pair,
                  h->d_cstart, h->d_contrib, h->d_estart, h->d_econtrib, h->d_code,    h->d_sbase,
                  h->d_slen,   h->d_ustart,  h->d_ep_pair,  h->d_ebase,   h->d_elen,
                  h->d_code_fd, h->d_elist, h->d_ustart_m,
                  h->d_r0bank, h->d_acbank,  h->d_cbank,  h->d_ebank,    h->d_arc,     h->d_tmp_x,
                  h->d_drag_edges, h->d_drag_terms, h->d_bc_rows, h->d_bc_values,
                  h->d_tmp_b,  h->d_tmp_v, h->d_rowptr_s, h->d_sbase_s, h->d_slen_s,
                  h->d_cbank_a, h->d_cbank_b, h->d_sbc_rows, h->d_stab_prep, h->d_geob};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete h;
}

int vef_set_params(vef_handle* h, const vef_params* p) {
  if (!h) return VEF_ERR_ARG;
  int rc = check_params(h, p);
  if (rc) return rc;
  h->params = *p;
  h->dprm = make_dev_params(*p);
  return VEF_OK;
}

int vef_create(int32_t n_nodes, int32_t n_el, const double* coords, const int32_t* conn, int32_t n_edges,
               const int32_t* edges, const vef_params* params, vef_handle** out) {
  if (!out) return set_err(nullptr, VEF_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (n_nodes <= 0 || n_el <= 0 || !coords || !conn) return set_err(nullptr, VEF_ERR_ARG, "empty mesh");
  if (n_edges < 0 || (n_edges > 0 && !edges)) return set_err(nullptr, VEF_ERR_ARG, "bad outlet edge list");
  if ((int64_t)n_el * 81 >= INT32_MAX || (int64_t)n_edges * 81 >= INT32_MAX)
    return set_err(nullptr, VEF_ERR_UNSUPPORTED, "too many elements for the packed scatter map");
  if ((int64_t)n_nodes * 6 + 1 >= INT32_MAX) return set_err(nullptr, VEF_ERR_UNSUPPORTED, "too many nodes");
  int rc = check_params(nullptr, params);
  if (rc) return rc;
  for (int64_t i = 0; i < (int64_t)n_el * 9; i++)
    if (conn[i] < 0 || conn[i] >= n_nodes) return set_err(nullptr, VEF_ERR_MESH, "connectivity out of range");
  for (int64_t e = 0; e < n_el; e++)
    for (int a = 0; a < 9; a++)
      for (int b2 = a + 1; b2 < 9; b2++)
        if (conn[e * 9 + a] == conn[e * 9 + b2])
          return set_err(nullptr, VEF_ERR_MESH, "element with a repeated node");
  for (int i = 0; i < n_edges; i++) {
    if (edges[2 * i] < 0 || edges[2 * i] >= n_el || edges[2 * i + 1] < 1 || edges[2 * i + 1] > 4)
      return set_err(nullptr, VEF_ERR_MESH, "outlet edge out of range");
  }

  vef_handle* h = new vef_handle();
  h->n_nodes = n_nodes;
  h->n_el = n_el;
  h->n_edges = n_edges;
  h->N = (int64_t)n_nodes * 6 + 1;
  h->params = *params;
  h->dprm = make_dev_params(*params);

  // Node -> elements (ascending) incidence.
  std::vector<int64_t> ne_start(n_nodes + 1, 0);
  for (int64_t i = 0; i < (int64_t)n_el * 9; i++) ne_start[conn[i] + 1]++;
  for (int i = 0; i < n_nodes; i++) ne_start[i + 1] += ne_start[i];
  std::vector<int32_t> ne_list(ne_start[n_nodes]);
  {
    std::vector<int64_t> fill(ne_start.begin(), ne_start.end() - 1);
    for (int32_t e = 0; e < n_el; e++)
      for (int a = 0; a < 9; a++) ne_list[fill[conn[(int64_t)e * 9 + a]]++] = e;  // ascending e by construction
  }
  // Node pairs (r, c): neighbours of r sorted; contributions per pair.
  std::vector<int64_t> pair_start(n_nodes + 1, 0);
  std::vector<int32_t> pair_col;
  std::vector<int32_t> cstart, contrib;
  std::vector<int64_t> diag_pair(n_nodes);
  pair_col.reserve((size_t)n_el * 64);
  cstart.reserve((size_t)n_el * 64 + 1);
  contrib.reserve((size_t)n_el * 81);
  cstart.push_back(0);
  std::vector<std::pair<int32_t, int32_t>> tmp;  // (c, packed contribution)
  for (int32_t r = 0; r < n_nodes; r++) {
    tmp.clear();
    for (int64_t t = ne_start[r]; t < ne_start[r + 1]; t++) {
      const int32_t e = ne_list[t];
      int jn = 0;
      while (conn[(int64_t)e * 9 + jn] != r) jn++;
      for (int in = 0; in < 9; in++) tmp.push_back({conn[(int64_t)e * 9 + in], e * 81 + jn * 9 + in});
    }
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& b) {
                       return a.first < b.first;
                     });
    size_t i = 0;
    while (i < tmp.size()) {
      const int32_t c = tmp[i].first;
      if (c == r) diag_pair[r] = (int64_t)pair_col.size();
      pair_col.push_back(c);
      while (i < tmp.size() && tmp[i].first == c) contrib.push_back(tmp[i++].second);  // ascending e
      cstart.push_back((int32_t)contrib.size());
    }
    pair_start[r + 1] = (int64_t)pair_col.size();
  }
  const int64_t npairs = (int64_t)pair_col.size();
  std::vector<int32_t> pair_row(npairs);
  for (int32_t r = 0; r < n_nodes; r++)
    for (int64_t pp = pair_start[r]; pp < pair_start[r + 1]; pp++) pair_row[pp] = r;
  h->n_pairs = npairs;
  // Outlet-edge contributions: rows of the element's nodes, columns of its nodes.
  std::vector<int32_t> ecount(npairs + 1, 0);
  auto find_pair = [&](int32_t r, int32_t c) -> int64_t {
    const int32_t* b0 = pair_col.data() + pair_start[r];
    const int32_t* b1 = pair_col.data() + pair_start[r + 1];
    const int32_t* it = std::lower_bound(b0, b1, c);
    return (it != b1 && *it == c) ? (it - pair_col.data()) : -1;
  };
  for (int32_t k = 0; k < n_edges; k++) {
    const int32_t e = edges[2 * k];
    for (int jn = 0; jn < 9; jn++)
      for (int in = 0; in < 9; in++) ecount[find_pair(conn[(int64_t)e * 9 + jn], conn[(int64_t)e * 9 + in]) + 1]++;
  }
  std::vector<int32_t> estart(npairs + 1, 0);
  for (int64_t i = 0; i < npairs; i++) estart[i + 1] = estart[i] + ecount[i + 1];
  std::vector<int32_t> econtrib(std::max<int32_t>(estart[npairs], 1));
  {
    std::vector<int32_t> fill(estart.begin(), estart.end() - 1);
    for (int32_t k = 0; k < n_edges; k++) {
      const int32_t e = edges[2 * k];
      for (int jn = 0; jn < 9; jn++)
        for (int in = 0; in < 9; in++) {
          const int64_t pp = find_pair(conn[(int64_t)e * 9 + jn], conn[(int64_t)e * 9 + in]);
          econtrib[fill[pp]++] = k * 81 + jn * 9 + in;  // ascending edge index
        }
    }
  }
  // CSR pattern: row (r, q) = the 6 columns of each neighbour c, then N-1; row N-1 dense.
  h->h_rowptr.assign(h->N + 1, 0);
  for (int32_t r = 0; r < n_nodes; r++)
    for (int q = 0; q < 6; q++) h->h_rowptr[(int64_t)r * 6 + q + 1] = (pair_start[r + 1] - pair_start[r]) * 6 + 1;
  h->h_rowptr[h->N] = h->N;
  for (int64_t i = 0; i < h->N; i++) h->h_rowptr[i + 1] += h->h_rowptr[i];
  h->nnz = h->h_rowptr[h->N];
  h->h_col.resize(h->nnz);
  for (int32_t r = 0; r < n_nodes; r++)
    for (int q = 0; q < 6; q++) {
      int64_t p = h->h_rowptr[(int64_t)r * 6 + q];
      for (int64_t pp = pair_start[r]; pp < pair_start[r + 1]; pp++)
        for (int iq = 0; iq < 6; iq++) h->h_col[p++] = pair_col[pp] * 6 + iq;
      h->h_col[p++] = (int32_t)(h->N - 1);
    }
  {
    int64_t p = h->h_rowptr[h->N - 1];
    for (int64_t c = 0; c < h->N; c++) h->h_col[p++] = (int32_t)c;
  }

  // Element colours for the Jacobian passes: the two elements of a
  // two-contributor pair get different colours (breadth-first two-colouring of
  // that graph; greedy colouring in element order where it has an odd cycle).
  std::vector<int32_t> colour(n_el, -1);
  {
    std::vector<std::pair<int32_t, int32_t>> ce;
    for (int64_t pp = 0; pp < npairs; pp++)
      if (cstart[pp + 1] - cstart[pp] == 2) {
        const int32_t a = contrib[cstart[pp]] / 81, b = contrib[cstart[pp] + 1] / 81;
        ce.push_back({a, b});
        ce.push_back({b, a});
      }
    std::sort(ce.begin(), ce.end());
    ce.erase(std::unique(ce.begin(), ce.end()), ce.end());
    std::vector<int64_t> adj0(n_el + 1, 0);
    for (const auto& p : ce) adj0[p.first + 1]++;
    for (int32_t e = 0; e < n_el; e++) adj0[e + 1] += adj0[e];
    bool two = true;
    std::vector<int32_t> queue;
    for (int32_t s0 = 0; s0 < n_el && two; s0++) {
      if (colour[s0] >= 0) continue;
      colour[s0] = 0;
      queue.assign(1, s0);
      for (size_t qi = 0; qi < queue.size() && two; qi++) {
        const int32_t u = queue[qi];
        for (int64_t t = adj0[u]; t < adj0[u + 1]; t++) {
          const int32_t v = ce[t].second;
          if (colour[v] < 0) {
            colour[v] = 1 - colour[u];
            queue.push_back(v);
          } else if (colour[v] == colour[u]) {
            two = false;
            break;
          }
        }
      }
    }
    if (!two) {
      std::vector<char> used;
      for (int32_t e = 0; e < n_el; e++) colour[e] = -1;
      for (int32_t e = 0; e < n_el; e++) {
        used.assign(used.size(), 0);
        for (int64_t t = adj0[e]; t < adj0[e + 1]; t++) {
          const int32_t c = colour[ce[t].second];
          if (c >= 0) {
            if ((size_t)c >= used.size()) used.resize(c + 1, 0);
            used[c] = 1;
          }
        }
        int32_t c = 0;
        while ((size_t)c < used.size() && used[c]) c++;
        colour[e] = c;
      }
    }
    int32_t n_colours = 0;
    for (int32_t e = 0; e < n_el; e++) n_colours = std::max(n_colours, colour[e] + 1);
    h->colour_off.assign(n_colours + 1, 0);
    for (int32_t e = 0; e < n_el; e++) h->colour_off[colour[e] + 1]++;
    for (int32_t c = 0; c < n_colours; c++) h->colour_off[c + 1] += h->colour_off[c];
  }
  std::vector<int32_t> elist(std::max(n_el, 1));
  {
    std::vector<int32_t> fill(h->colour_off.begin(), h->colour_off.end());
    for (int32_t e = 0; e < n_el; e++) elist[fill[colour[e]]++] = e;
  }

  // Scatter codes (the Jacobian passes' and the stability assembly's), the
  // pairs summed from banks and their contribution lists.
  std::vector<int32_t> code((size_t)n_el * 81), code_fd((size_t)n_el * 81);
  std::vector<int64_t> sbase;
  std::vector<int32_t> slen, ustart(1, 0), ustart_m(1, 0);
  int64_t n_cidx = 0, n_cidx_m = 0;
  std::vector<int64_t> ep_pair, ebase;
  std::vector<int32_t> elen;
  for (int64_t pp = 0; pp < npairs; pp++) {
    const int32_t r = pair_row[pp];
    const int64_t pos = pp - pair_start[r];
    const int64_t rb = h->h_rowptr[(int64_t)r * 6];
    const int32_t rl = (int32_t)(h->h_rowptr[(int64_t)r * 6 + 1] - rb);
    const int32_t c0 = cstart[pp], c1 = cstart[pp + 1];
    auto at = [](int32_t w) { return (size_t)(w / 81) * 81 + w % 81; };
    if (c1 - c0 == 1) {
      code[at(contrib[c0])] = (int32_t)(pos << 1);
      code_fd[at(contrib[c0])] = (int32_t)(pos << 2) | kCodeStore;
    } else {
      h->h_sbase_s.push_back(rb - (int64_t)r * 6 + pos * 6);  // rows before r lose their arclength entry
      h->h_slen_s.push_back(rl - 1);
      for (int32_t c = c0; c < c1; c++) {
        code[at(contrib[c])] = (int32_t)((n_cidx << 1) | 1);
        n_cidx++;
      }
      ustart.push_back((int32_t)n_cidx);
      if (c1 - c0 == 2) {
        const int32_t wa = contrib[c0], wb = contrib[c0 + 1];
        const bool a_first = colour[wa / 81] < colour[wb / 81];
        code_fd[at(wa)] = (int32_t)(pos << 2) | (a_first ? kCodeStore : kCodeAdd);
        code_fd[at(wb)] = (int32_t)(pos << 2) | (a_first ? kCodeAdd : kCodeStore);
      } else {
        sbase.push_back(rb + pos * 6);
        slen.push_back(rl);
        for (int32_t c = c0; c < c1; c++) {
          code_fd[at(contrib[c])] = (int32_t)((n_cidx_m << 2) | kCodeBank);
          n_cidx_m++;
        }
        ustart_m.push_back((int32_t)n_cidx_m);
      }
    }
    if (estart[pp + 1] > estart[pp]) {
      ep_pair.push_back(pp);
      ebase.push_back(rb + pos * 6);
      elen.push_back(rl);
    }
  }
  h->n_shared = (int64_t)h->h_sbase_s.size();
  h->n_cidx = n_cidx;
  h->n_shared_m = (int64_t)sbase.size();
  h->n_cidx_m = n_cidx_m;
  h->n_epairs = (int64_t)ep_pair.size();

  // Device copies.
#define UP(dst, src, n)                         \
  do {                                          \
    int rc_ = dev_upload(h, &(dst), (src), (n)); \
    if (rc_) {                                  \
      g_create_err = h->err;                    \
      vef_destroy(h);                           \
      return rc_;                               \
    }                                           \
  } while (0)
  UP(h->d_coords, coords, (size_t)n_nodes * 2);
  UP(h->d_conn, conn, (size_t)n_el * 9);
  UP(h->d_edges, edges, (size_t)n_edges * 2);
  UP(h->d_rowptr, h->h_rowptr.data(), (size_t)h->N + 1);
  UP(h->d_col, h->h_col.data(), (size_t)h->nnz);
  UP(h->d_diag_pair, diag_pair.data(), (size_t)n_nodes);
  UP(h->d_cstart, cstart.data(), cstart.size());
  UP(h->d_contrib, contrib.data(), contrib.size());
  UP(h->d_estart, estart.data(), estart.size());
  UP(h->d_econtrib, econtrib.data(), econtrib.size());
  UP(h->d_code, code.data(), code.size());
  UP(h->d_ustart, ustart.data(), ustart.size());
  UP(h->d_code_fd, code_fd.data(), code_fd.size());
  UP(h->d_elist, elist.data(), elist.size());
  UP(h->d_sbase, sbase.data(), sbase.size());
  UP(h->d_slen, slen.data(), slen.size());
  UP(h->d_ustart_m, ustart_m.data(), ustart_m.size());
  UP(h->d_ep_pair, ep_pair.data(), ep_pair.size());
  UP(h->d_ebase, ebase.data(), ebase.size());
  UP(h->d_elen, elen.data(), elen.size());
  UP(h->d_r0bank, (const double*)nullptr, (size_t)n_el * 54);
  UP(h->d_geob, (const double*)nullptr, (size_t)n_el * 9 * kGeoWords);
  UP(h->d_acbank, (const double*)nullptr, (size_t)n_el * 54);
  UP(h->d_cbank, (const double*)nullptr, (size_t)std::max<int64_t>(h->n_cidx_m, 1) * 36);
  UP(h->d_ebank, (const double*)nullptr, (size_t)std::max(n_edges, 1) * 18 * kNeval);
  h->arc_blocks = (int)((n_nodes + kArcChunk - 1) / kArcChunk);
  UP(h->d_arc, (const double*)nullptr, (size_t)6 * h->arc_blocks);
  h->arc_scratch = scratch_address(&h->device);
#undef UP
  if (upload_tables(h)) {
    g_create_err = h->err;
    vef_destroy(h);
    return VEF_ERR_HIP;
  }
  *out = h;
  return VEF_OK;
}

int vef_assemble(vef_handle* h, const double* x, int32_t flag, double* b, double* values, void* stream) {
  if (!h) return set_err(nullptr, VEF_ERR_ARG, "handle is NULL");
  if (!x || !b) return set_err(h, VEF_ERR_ARG, "x and b are required");
  if (flag != VEF_RES_ONLY && flag != VEF_RES_JAC && flag != VEF_RES_JAC_FP32)
    return set_err(h, VEF_ERR_ARG, "flag must be VEF_RES_ONLY, VEF_RES_JAC or VEF_RES_JAC_FP32");
  if (flag != VEF_RES_ONLY && !values) return set_err(h, VEF_ERR_ARG, "values is required with a Jacobian flag");
  if (flag == VEF_RES_JAC_FP32 && h->dprm.dcs_on)
    return set_err(h, VEF_ERR_UNSUPPORTED, "the FP32 analytic Jacobian needs DfN > 1e-12 (discontinuity capturing off)");
  hipStream_t st = (hipStream_t)stream;
  const int64_t eu = h->N - 1;
  const int jac = flag != VEF_RES_ONLY;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  if (h->profiling) {
    for (int i = 0; i < 4; i++) VEF_HIP(h, hipEventCreate(&ev[i]));
    VEF_HIP(h, hipEventRecord(ev[0], st));
  }
  // 1. base residual rows of every element
  launch_bulk(h, 0, values, x, eu, st, jac ? h->d_geob : nullptr);
  VEF_HIP(h, hipGetLastError());
  // 2. outlet edges (residual and FD bank)
  if (h->n_edges > 0) {
    hipLaunchKernelGGL(vef_edge_kernel, dim3(h->n_edges), dim3(192), 0, st, h->d_coords, h->d_conn, h->d_edges, x,
                       eu, h->n_edges, h->dprm, h->d_ebank);
    VEF_HIP(h, hipGetLastError());
  }
  if (h->profiling) VEF_HIP(h, hipEventRecord(ev[1], st));
  // 3. element Jacobians (forward differences, or the FP32 analytic one):
  //    single-contributor entries straight into values
  if (flag == VEF_RES_JAC) {
    launch_bulk(h, 1, values, x, eu, st, h->d_geob);
    VEF_HIP(h, hipGetLastError());
  } else if (flag == VEF_RES_JAC_FP32) {
    FdOut fo{h->d_r0bank, h->d_code_fd, h->d_rowptr, values, h->d_cbank, h->d_acbank, h->d_geob, nullptr, 0};
    vefad::Params<float> prm{(float)h->params.ReN, (float)h->params.BvN, (float)h->params.VeN, (float)h->params.DfN,
                             (float)h->params.WiN0, (float)(1.0 - h->params.BvN), h->params.model};
    for (size_t c = 0; c + 1 < h->colour_off.size(); c++) {  // one launch per colour, in colour order
      fo.elist = h->d_elist + h->colour_off[c];
      fo.n_list = h->colour_off[c + 1] - h->colour_off[c];
      hipLaunchKernelGGL(vef_jac32_kernel, dim3((unsigned)fo.n_list), dim3(jac32::kThreads), 0, st, h->d_conn, x, eu,
                         h->n_el, prm, fo);
    }
    VEF_HIP(h, hipGetLastError());
  }
  if (h->profiling) VEF_HIP(h, hipEventRecord(ev[2], st));
  const int64_t node_rows = (int64_t)h->n_nodes * 6;
  if (jac) {
    // 4. entries of three or more elements in element order, 5. outlet-edge
    //    entries after them
    if (h->n_shared_m > 0) {
      hipLaunchKernelGGL(vef_shared_kernel, dim3((unsigned)((h->n_shared_m * 36 + 255) / 256)), dim3(256), 0, st,
                         h->n_shared_m, h->d_sbase, h->d_slen, h->d_ustart_m, h->d_cbank, values);
      VEF_HIP(h, hipGetLastError());
    }
    if (h->n_epairs > 0) {
      hipLaunchKernelGGL(vef_edge_add_kernel, dim3((unsigned)((h->n_epairs * 12 + 255) / 256)), dim3(256), 0, st,
                         h->n_epairs, h->d_ep_pair, h->d_ebase, h->d_elen, h->d_estart, h->d_econtrib, h->d_ebank,
                         values);
      VEF_HIP(h, hipGetLastError());
    }
    // Row N-1 (arclength row) is not touched by the bulk assembly: zero.
    const int64_t last0 = h->h_rowptr[h->N - 1];
    hipLaunchKernelGGL(vef_fill_kernel, dim3(1024), dim3(256), 0, st, values + last0, h->nnz - last0);
    VEF_HIP(h, hipGetLastError());
  }
  // 6. b and the dR/dEU column
  hipLaunchKernelGGL(vef_rhs_kernel, dim3((unsigned)((node_rows + 255) / 256)), dim3(256), 0, st, node_rows, jac,
                     h->d_rowptr, h->d_diag_pair, h->d_cstart, h->d_contrib, h->d_estart, h->d_econtrib,
                     h->d_r0bank, h->d_acbank, h->d_ebank, values, b);
  VEF_HIP(h, hipGetLastError());
  hipLaunchKernelGGL(vef_fill_kernel, dim3(1), dim3(64), 0, st, b + (h->N - 1), (int64_t)1);
  VEF_HIP(h, hipGetLastError());
  if (h->profiling) {
    VEF_HIP(h, hipEventRecord(ev[3], st));
    h->ev_bulk.push_back(ev[1]);
    h->ev_bulk.push_back(ev[2]);
    h->ev_scatter.push_back(ev[2]);
    h->ev_scatter.push_back(ev[3]);
    h->ev_other.push_back(ev[0]);
    h->ev_other.push_back(ev[1]);
  }
  h->last_flag = flag;
  return VEF_OK;
}

namespace {
void release_events(vef_handle* h) {
  // ev_other holds each assembly's first two events, ev_scatter its last two;
  // ev_bulk repeats events of both.
  for (hipEvent_t e : h->ev_other) (void)hipEventDestroy(e);
  for (hipEvent_t e : h->ev_scatter) (void)hipEventDestroy(e);
  h->ev_bulk.clear();
  h->ev_scatter.clear();
  h->ev_other.clear();
}
}  // namespace

int vef_profile(vef_handle* h, int32_t enable) {
  if (!h) return VEF_ERR_ARG;
  release_events(h);
  h->profiling = enable != 0;
  return VEF_OK;
}

int vef_profile_read(vef_handle* h, double* bulk_ms, double* scatter_ms, int64_t* launches) {
  if (!h || !bulk_ms || !scatter_ms || !launches) return VEF_ERR_ARG;
  double tb = 0.0, ts = 0.0;
  for (size_t i = 0; i + 1 < h->ev_bulk.size(); i += 2) {
    float ms = 0.0f;
    VEF_HIP(h, hipEventSynchronize(h->ev_bulk[i + 1]));
    VEF_HIP(h, hipEventElapsedTime(&ms, h->ev_bulk[i], h->ev_bulk[i + 1]));
    tb += ms;
  }
  for (size_t i = 0; i + 1 < h->ev_scatter.size(); i += 2) {
    float ms = 0.0f;
    VEF_HIP(h, hipEventSynchronize(h->ev_scatter[i + 1]));
    VEF_HIP(h, hipEventElapsedTime(&ms, h->ev_scatter[i], h->ev_scatter[i + 1]));
    ts += ms;
  }
  *bulk_ms = tb;
  *scatter_ms = ts;
  *launches = (int64_t)(h->ev_bulk.size() / 2);
  release_events(h);
  return VEF_OK;
}

static int arclength_impl(vef_handle* h, int32_t mode, const double* x, const double* x_old, const double* tangent,
                          int32_t increment, double LN0, double dArN, double dS, int32_t flag, double* b,
                          double* values, void* stream, double* partials) {
  if (!x || !b) return set_err(h, VEF_ERR_ARG, "x and b are required");
  if (flag != VEF_RES_ONLY && flag != VEF_RES_JAC && flag != VEF_RES_JAC_FP32)
    return set_err(h, VEF_ERR_ARG, "bad flag");
  if (flag != VEF_RES_ONLY && !values) return set_err(h, VEF_ERR_ARG, "values is required with a Jacobian flag");
  if (flag == VEF_RES_JAC_FP32) flag = VEF_RES_JAC;  // the row itself is the same for both Jacobians
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = h->N;
  if (mode == VEF_ARC_PIN) {
    hipLaunchKernelGGL(vef_arc_pin_rhs_kernel, dim3(1), dim3(64), 0, st, N, x, increment, LN0, dArN, b);
    VEF_HIP(h, hipGetLastError());
    if (flag == VEF_RES_JAC) {
      hipLaunchKernelGGL(vef_arc_pin_kernel, dim3(1024), dim3(256), 0, st, N, h->d_rowptr, values);
      VEF_HIP(h, hipGetLastError());
    }
    return VEF_OK;
  }
  if (mode != VEF_ARC_PSEUDO) return set_err(h, VEF_ERR_ARG, "mode must be VEF_ARC_PIN or VEF_ARC_PSEUDO");
  if (!x_old || !tangent) return set_err(h, VEF_ERR_ARG, "x_old and tangent are required for VEF_ARC_PSEUDO");
  hipLaunchKernelGGL(vef_arc_dot_kernel, dim3(h->arc_blocks), dim3(256), 0, st, (int64_t)h->n_nodes, x, x_old,
                     tangent, partials);
  VEF_HIP(h, hipGetLastError());
  hipLaunchKernelGGL(vef_arc_rhs_kernel, dim3(1), dim3(64), 0, st, N, h->arc_blocks, partials, x, x_old, tangent, dS,
                     b);
  VEF_HIP(h, hipGetLastError());
  if (flag == VEF_RES_JAC) {
    hipLaunchKernelGGL(vef_copy_kernel, dim3(1024), dim3(256), 0, st, values + h->h_rowptr[N - 1], tangent, N);
    VEF_HIP(h, hipGetLastError());
  }
  return VEF_OK;
}

int vef_arclength(vef_handle* h, int32_t mode, const double* x, const double* x_old, const double* tangent,
                  int32_t increment, double LN0, double dArN, double dS, int32_t flag, double* b, double* values,
                  void* stream) {
  if (!h) return set_err(nullptr, VEF_ERR_ARG, "handle is NULL");
  return arclength_impl(h, mode, x, x_old, tangent, increment, LN0, dArN, dS, flag, b, values, stream, h->d_arc);
}

int vef_arclength_noalloc(vef_handle* h, int32_t mode, const double* x, const double* x_old, const double* tangent,
                          int32_t increment, double LN0, double dArN, double dS, int32_t flag, double* b,
                          double* values, void* stream) {
  if (!h) return set_err(nullptr, VEF_ERR_ARG, "handle is NULL");
  if (h->arc_blocks > kArcScratchBlocks)
    return set_err(h, VEF_ERR_UNSUPPORTED, "vef_arclength_noalloc: more than 16,777,216 nodes (its scratch holds "
                                           "4096 node chunks of 4096)");
  if (!h->arc_scratch)
    return set_err(h, VEF_ERR_HIP, "vef_arclength_noalloc: the scratch array's device address is unknown");
  return arclength_impl(h, mode, x, x_old, tangent, increment, LN0, dArN, dS, flag, b, values, stream,
                        h->arc_scratch);
}

int vef_set_dirichlet(vef_handle* h, int64_t n_rows, const int64_t* rows, const double* values) {
  if (!h) return set_err(nullptr, VEF_ERR_ARG, "handle is NULL");
  if (n_rows < 0 || (n_rows > 0 && (!rows || !values))) return set_err(h, VEF_ERR_ARG, "bad Dirichlet list");
  // The reference inserts boundary by boundary, so a row listed twice keeps
  // its last value: the list's positions sorted by (row, position), the last
  // of each row kept.
  std::vector<int64_t> order((size_t)n_rows);
  for (int64_t i = 0; i < n_rows; i++) {
    if (rows[i] < 0 || rows[i] >= h->N) return set_err(h, VEF_ERR_ARG, "Dirichlet row out of range");
    order[(size_t)i] = i;
  }
  std::sort(order.begin(), order.end(),
            [rows](int64_t a, int64_t b) { return rows[a] != rows[b] ? rows[a] < rows[b] : a < b; });
  std::vector<int64_t> r;
  std::vector<double> v;
  r.reserve(order.size());
  v.reserve(order.size());
  for (size_t k = 0; k < order.size(); k++)
    if (k + 1 == order.size() || rows[order[k + 1]] != rows[order[k]]) {
      r.push_back(rows[order[k]]);
      v.push_back(values[order[k]]);
    }
  // the buffers are kept while the list fits them; work still queued may read
  // them, so the device is idle before they are written
  VEF_HIP(h, hipDeviceSynchronize());
  if ((int64_t)r.size() > h->bc_cap) {
    if (h->d_bc_rows) (void)hipFree(h->d_bc_rows);
    if (h->d_bc_values) (void)hipFree(h->d_bc_values);
    h->d_bc_rows = nullptr;
    h->d_bc_values = nullptr;
    h->bc_cap = 0;
    h->n_bc = 0;
    VEF_HIP(h, hipMalloc((void**)&h->d_bc_rows, sizeof(int64_t) * r.size()));
    VEF_HIP(h, hipMalloc((void**)&h->d_bc_values, sizeof(double) * r.size()));
    h->bc_cap = (int64_t)r.size();
  }
  h->n_bc = (int64_t)r.size();
  if (h->n_bc == 0) return VEF_OK;
  VEF_HIP(h, hipMemcpy(h->d_bc_rows, r.data(), sizeof(int64_t) * r.size(), hipMemcpyHostToDevice));
  VEF_HIP(h, hipMemcpy(h->d_bc_values, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
  return VEF_OK;
}

int vef_apply_dirichlet(vef_handle* h, const double* x, int32_t flag, double* b, double* values, void* stream) {
  if (!h) return set_err(nullptr, VEF_ERR_ARG, "handle is NULL");
  if (!x || !b) return set_err(h, VEF_ERR_ARG, "x and b are required");
  if (flag != VEF_RES_ONLY && flag != VEF_RES_JAC && flag != VEF_RES_JAC_FP32)
    return set_err(h, VEF_ERR_ARG, "bad flag");
  if (flag != VEF_RES_ONLY && !values) return set_err(h, VEF_ERR_ARG, "values is required with a Jacobian flag");
  if (flag == VEF_RES_JAC_FP32) flag = VEF_RES_JAC;
  if (h->n_bc == 0) return VEF_OK;
  const int64_t blocks = (h->n_bc + 3) / 4;
  hipLaunchKernelGGL(vef_dirichlet_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, h->n_bc,
                     h->d_bc_rows, h->d_bc_values, x, flag == VEF_RES_JAC ? 1 : 0, h->d_rowptr, h->d_col, b,
                     values);
  VEF_HIP(h, hipGetLastError());
  return VEF_OK;
}

int vef_set_drag_edges(vef_handle* h, int32_t n_edges, const int32_t* edges) {
  if (!h) return set_err(nullptr, VEF_ERR_ARG, "handle is NULL");
  if (n_edges < 0 || (n_edges > 0 && !edges)) return set_err(h, VEF_ERR_ARG, "bad edge list");
  for (int32_t k = 0; k < n_edges; k++) {
    if (edges[2 * k] < 0 || edges[2 * k] >= h->n_el) return set_err(h, VEF_ERR_MESH, "drag edge element out of range");
    if (edges[2 * k + 1] < 1 || edges[2 * k + 1] > 4) return set_err(h, VEF_ERR_MESH, "drag edge id must be 1..4");
  }
  if (h->d_drag_edges) (void)hipFree(h->d_drag_edges);
  if (h->d_drag_terms) (void)hipFree(h->d_drag_terms);
  h->d_drag_edges = nullptr;
  h->d_drag_terms = nullptr;
  h->n_drag_edges = n_edges;
  if (n_edges == 0) return VEF_OK;
  VEF_HIP(h, hipMalloc((void**)&h->d_drag_edges, sizeof(int32_t) * 2 * (size_t)n_edges));
  VEF_HIP(h, hipMalloc((void**)&h->d_drag_terms, sizeof(double) * (size_t)n_edges));
  VEF_HIP(h, hipMemcpy(h->d_drag_edges, edges, sizeof(int32_t) * 2 * (size_t)n_edges, hipMemcpyHostToDevice));
  return VEF_OK;
}

int vef_drag(vef_handle* h, const double* x, double* Fd, double* per_edge, void* stream) {
  if (!h) return set_err(nullptr, VEF_ERR_ARG, "handle is NULL");
  if (!x || !Fd) return set_err(h, VEF_ERR_ARG, "x and Fd are required");
  hipStream_t st = (hipStream_t)stream;
  const int32_t n = h->n_drag_edges;
  double* terms = per_edge ? per_edge : h->d_drag_terms;
  if (n > 0) {
    const int blocks = (n + kDragEdgesPerBlock - 1) / kDragEdgesPerBlock;
    hipLaunchKernelGGL(vef_drag_edge_kernel, dim3(blocks), dim3(64), 0, st, h->d_coords, h->d_conn,
                       h->d_drag_edges, x, h->N - 1, n, h->dprm, terms);
    VEF_HIP(h, hipGetLastError());
  }
  hipLaunchKernelGGL(vef_drag_sum_kernel, dim3(1), dim3(64), 0, st, n, terms, Fd);
  VEF_HIP(h, hipGetLastError());
  return VEF_OK;
}

int vef_nodal_stress(vef_handle* h, const double* x, double WiN, double* T, void* stream) {
  if (!h) return set_err(nullptr, VEF_ERR_ARG, "handle is NULL");
  if (!x || !T) return set_err(h, VEF_ERR_ARG, "x and T are required");
  const int64_t n = h->n_nodes;
  if (n > 0) {
    hipLaunchKernelGGL(vef_nodal_stress_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, n, x, h->dprm.omb, WiN, T);
    VEF_HIP(h, hipGetLastError());
  }
  return VEF_OK;
}

// ----------------------------------------------------------------------------
// Linear stability assembly
// ----------------------------------------------------------------------------
int64_t vef_stability_nnz(const vef_handle* h) {
  if (!h) return -1;
  return h->h_rowptr[h->N - 1] - (h->N - 1);
}

int vef_get_stability_pattern(const vef_handle* h, int64_t* rowptr, int32_t* colind) {
  if (!h || !rowptr || !colind) return VEF_ERR_ARG;
  const int64_t Ns = h->N - 1;
  for (int64_t r = 0; r <= Ns; r++) rowptr[r] = h->h_rowptr[r] - r;
  for (int64_t r = 0; r < Ns; r++)
    std::memcpy(colind + rowptr[r], h->h_col.data() + h->h_rowptr[r], sizeof(int32_t) * (rowptr[r + 1] - rowptr[r]));
  return VEF_OK;
}

int vef_set_stability_dirichlet(vef_handle* h, int64_t n_rows, const int64_t* rows) {
  if (!h) return set_err(nullptr, VEF_ERR_ARG, "handle is NULL");
  if (n_rows < 0 || (n_rows > 0 && !rows)) return set_err(h, VEF_ERR_ARG, "bad row list");
  const int64_t Ns = h->N - 1;
  std::vector<int64_t> r(rows, rows + n_rows);
  for (int64_t v : r)
    if (v < 0 || v >= Ns) return set_err(h, VEF_ERR_ARG, "stability Dirichlet row out of range");
  std::sort(r.begin(), r.end());
  r.erase(std::unique(r.begin(), r.end()), r.end());  // MatZeroRows is idempotent
  if (h->d_sbc_rows) {
    VEF_HIP(h, hipFree(h->d_sbc_rows));
    h->d_sbc_rows = nullptr;
  }
  h->n_sbc = (int64_t)r.size();
  if (h->n_sbc > 0) {
    int rc = dev_upload(h, &h->d_sbc_rows, r.data(), r.size());
    if (rc) return rc;
  }
  return VEF_OK;
}

static int ensure_stability(vef_handle* h) {
  if (h->stab_ready) return VEF_OK;
  const int64_t Ns = h->N - 1;
  std::vector<int64_t> rp(Ns + 1);
  for (int64_t r = 0; r <= Ns; r++) rp[r] = h->h_rowptr[r] - r;
  h->nnz_s = rp[Ns];
  int rc = dev_upload(h, &h->d_rowptr_s, rp.data(), rp.size());
  if (!rc) rc = dev_upload(h, &h->d_sbase_s, h->h_sbase_s.data(), h->h_sbase_s.size());
  if (!rc) rc = dev_upload(h, &h->d_slen_s, h->h_slen_s.data(), h->h_slen_s.size());
  if (!rc) rc = dev_upload(h, &h->d_cbank_a, (const double*)nullptr, (size_t)std::max<int64_t>(h->n_cidx, 1) * 36);
  if (!rc) rc = dev_upload(h, &h->d_cbank_b, (const double*)nullptr, (size_t)std::max<int64_t>(h->n_cidx, 1) * 36);
  if (!rc)
    rc = dev_upload(h, &h->d_stab_prep, (const double*)nullptr, (size_t)std::max<int64_t>(h->n_el, 1) * kPrepPerElement);
  if (rc) return rc;
  h->stab_ready = true;
  return VEF_OK;
}

int vef_stability_assemble(vef_handle* h, const double* x, double* A, double* B, void* stream) {
  if (!h) return set_err(nullptr, VEF_ERR_ARG, "handle is NULL");
  if (!x || !A || !B) return set_err(h, VEF_ERR_ARG, "x, A and B are required");
  if (!(h->params.DfN > 1.0e-12))
    return set_err(h, VEF_ERR_UNSUPPORTED,
                   "the stability assembly needs DfN > 1e-12 (discontinuity capturing off)");
  int rc = ensure_stability(h);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  vefstab::Params prm{h->params.ReN, h->params.BvN, h->params.VeN, h->params.DfN, h->params.WiN0,
                      1.0 - h->params.BvN, h->params.model};
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  if (h->profiling) {
    for (int i = 0; i < 3; i++) VEF_HIP(h, hipEventCreate(&ev[i]));
    VEF_HIP(h, hipEventRecord(ev[0], st));
  }
  hipLaunchKernelGGL(vef_stab_prep_kernel, dim3((unsigned)((h->n_el + 6) / 7)), dim3(64), 0, st, h->d_coords,
                     h->d_conn, x, h->N - 1, h->n_el, prm, h->d_stab_prep);
  VEF_HIP(h, hipGetLastError());
  hipLaunchKernelGGL(vef_stab_kernel, dim3((unsigned)((h->n_el + 7) / 8) * 64), dim3(64), 0, st, h->d_stab_prep,
                     h->d_conn, h->n_el, prm, h->d_code, h->d_rowptr_s, A, B, h->d_cbank_a, h->d_cbank_b);
  VEF_HIP(h, hipGetLastError());
  if (h->profiling) VEF_HIP(h, hipEventRecord(ev[1], st));
  if (h->n_shared > 0) {
    const unsigned blocks = (unsigned)((h->n_shared * 36 + 255) / 256);
    hipLaunchKernelGGL(vef_shared_kernel, dim3(blocks), dim3(256), 0, st, h->n_shared, h->d_sbase_s, h->d_slen_s,
                       h->d_ustart, h->d_cbank_a, A);
    VEF_HIP(h, hipGetLastError());
    hipLaunchKernelGGL(vef_shared_kernel, dim3(blocks), dim3(256), 0, st, h->n_shared, h->d_sbase_s, h->d_slen_s,
                       h->d_ustart, h->d_cbank_b, B);
    VEF_HIP(h, hipGetLastError());
  }
  if (h->n_sbc > 0) {
    hipLaunchKernelGGL(vef_stab_dirichlet_kernel, dim3((unsigned)((h->n_sbc + 3) / 4)), dim3(256), 0, st, h->n_sbc,
                       h->d_sbc_rows, h->d_rowptr_s, h->d_rowptr, h->d_col, A, B);
    VEF_HIP(h, hipGetLastError());
  }
  if (h->profiling) {
    VEF_HIP(h, hipEventRecord(ev[2], st));
    h->ev_bulk.push_back(ev[0]);
    h->ev_bulk.push_back(ev[1]);
    h->ev_scatter.push_back(ev[1]);
    h->ev_scatter.push_back(ev[2]);
    h->ev_other.push_back(ev[0]);  // release_events destroys ev_other and ev_scatter: each event once
  }
  return VEF_OK;
}

int vef_assemble_host(vef_handle* h, const double* x, int32_t flag, double* b, double* values) {
  if (!h) return set_err(nullptr, VEF_ERR_ARG, "handle is NULL");
  if (flag != VEF_RES_ONLY && flag != VEF_RES_JAC && flag != VEF_RES_JAC_FP32)
    return set_err(h, VEF_ERR_ARG, "bad flag");
  const bool jac = flag != VEF_RES_ONLY;
  if (!x || !b || (jac && !values)) return set_err(h, VEF_ERR_ARG, "missing array");
  if (!h->d_tmp_x) {
    VEF_HIP(h, hipMalloc((void**)&h->d_tmp_x, sizeof(double) * h->N));
    VEF_HIP(h, hipMalloc((void**)&h->d_tmp_b, sizeof(double) * h->N));
  }
  if (jac && !h->d_tmp_v) VEF_HIP(h, hipMalloc((void**)&h->d_tmp_v, sizeof(double) * h->nnz));
  VEF_HIP(h, hipMemcpy(h->d_tmp_x, x, sizeof(double) * h->N, hipMemcpyHostToDevice));
  int rc = vef_assemble(h, h->d_tmp_x, flag, h->d_tmp_b, jac ? h->d_tmp_v : nullptr, nullptr);
  if (rc) return rc;
  VEF_HIP(h, hipDeviceSynchronize());
  VEF_HIP(h, hipMemcpy(b, h->d_tmp_b, sizeof(double) * h->N, hipMemcpyDeviceToHost));
  if (jac) VEF_HIP(h, hipMemcpy(values, h->d_tmp_v, sizeof(double) * h->nnz, hipMemcpyDeviceToHost));
  return VEF_OK;
}

}  // extern "C"
