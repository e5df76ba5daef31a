// vef_stability_math.h — per-Gauss-point arithmetic of the linear stability
// assembly (Cylinder_Stability_2D, Stress_Reform = 'SQRT'), shared by the
// stability element kernel and its host-side check.
//
// The reference assembles the generalised eigenproblem A q = lambda B q by
// forward differences in a complex disturbance state d around the converged
// base state b (storeBulkStability, src/Stability/Equations_Stability_mod.F90:135-226).
// Per Gauss point it evaluates the base quantities (setFlowUnknownsStability,
// strongFormEquations, setStabilizationTerms), then the disturbance's
// (src/Stability/Element_Calculations_Stability_mod.F90:19-102,186-262,302-329;
// src/Stability/Constitutive_Models_SQRT_Stability_mod.F90:87-487), then the
// test-node terms (src/Stability/Residuals_Stability_mod.F90:125-153,209-313).
// On this path every imaginary part is zero: the spanwise terms are commented
// out (Element_Calculations_Stability_mod.F90:51-81) and eps_jac is the real
// part of perturbVariable's value, assigned to a PetscReal
// (Equations_Stability_mod.F90:155,179).  Divisions with a complex operand
// round as (a*c)/(c*c), real ones as a/c, every sum is the reference's
// left-to-right sum, and the code is compiled with -ffp-contract=off.
//
// The disturbance of a forward difference has one nonzero nodal value,
// eps_jac at (inod, ieq), so each interpolated disturbance field is
// eps_jac*(basis value of inod) or zero; the reference's nine-term sums give
// exactly that.  The 3x3 tensors of the reference have a zero third row and
// column in 2D, and terms multiplied by coefficients that are zero on this
// path (g = 1 and dg/dx = 0 of the base, g = dg/dx = 0 and f(3) = 0 of the
// disturbance, tdcs = 0 for DfN > 1e-12) add signed zeros; they are left out,
// which changes no nonzero value.  Entries that come out as zero may differ
// in sign from the reference's, and become +0 when added into the zeroed
// matrix, as in PETSc.
#ifndef VEF_STABILITY_MATH_H
#define VEF_STABILITY_MATH_H

#include <hip/hip_runtime.h>

#include <cmath>

namespace vefstab {

constexpr double kEpsJac = 1.0e-8;  // perturbVariable, src/Tools_mod.F90:844-855 (real part)

struct Params {
  double ReN, BvN, VeN, DfN, WiN0, omb;  // omb = 1 - BvN
  int model;                             // 0 L-PTT (VeN = epsilon), 1 Giesekus (VeN = alpha)
};

// Complex division with zero imaginary parts.
__host__ __device__ inline double cdiv(double a, double c) { return (a * c) / (c * c); }

// Base-state quantities of one Gauss point read by the stability terms.
// GUjk = GU(j,k) = dU_k/dx_j (the reference's index order).
struct Base {
  double U0, U1;
  double GU00, GU01, GU10, GU11;
  double S00, S01, S11;
  double dS00x, dS00y, dS01x, dS01y, dS11x, dS11y;
  double C00, C01, C11;     // C = S.S (C10 = C01)
  double SI00, SI01, SI11;  // S^-1 (SI10 = SI01)
  double f0, f1, f2, f_T;
  double W01;               // setOmegaTensor: omega from the traceless gradient
  double nb, db;            // omegaNominator2D(S, GUT), S11 + S22 of WTensor2DStability
  double WiN, tlsme, tlsic, tlsce;
};

// Base Gauss point: setFlowUnknownsStability on TEMP_TL_b (u[n*6 + q]),
// strongFormEquations (WiN = WiN0*EU, setFlowKinematics, strongFormConEQ_SQRT)
// and setStabilizationTerms (src/Element_Calculations_mod.F90:537-573,605-634,701-773;
// src/Constitutive_Models_SQRT_mod.F90:120-180,379-509,513-584).
__host__ __device__ inline void base_point(const double* u, const double* bf, const double* dFx, const double* dFy,
                                           double H2, double EU, const Params& prm, Base& o) {
  double Ux = u[0] * bf[0], Uy = u[1] * bf[0];
  double GU00 = u[0] * dFx[0], GU10 = u[0] * dFy[0], GU01 = u[1] * dFx[0], GU11 = u[1] * dFy[0];
  double S00 = u[3] * bf[0], S01 = u[4] * bf[0], S11 = u[5] * bf[0];
  double dS00x = u[3] * dFx[0], dS00y = u[3] * dFy[0], dS01x = u[4] * dFx[0], dS01y = u[4] * dFy[0];
  double dS11x = u[5] * dFx[0], dS11y = u[5] * dFy[0];
  for (int n = 1; n < 9; n++) {
    const double* v = u + 6 * n;
    Ux = Ux + v[0] * bf[n];
    Uy = Uy + v[1] * bf[n];
    GU00 = GU00 + v[0] * dFx[n];
    GU10 = GU10 + v[0] * dFy[n];
    GU01 = GU01 + v[1] * dFx[n];
    GU11 = GU11 + v[1] * dFy[n];
    S00 = S00 + v[3] * bf[n];
    S01 = S01 + v[4] * bf[n];
    S11 = S11 + v[5] * bf[n];
    dS00x = dS00x + v[3] * dFx[n];
    dS00y = dS00y + v[3] * dFy[n];
    dS01x = dS01x + v[4] * dFx[n];
    dS01y = dS01y + v[4] * dFy[n];
    dS11x = dS11x + v[5] * dFx[n];
    dS11y = dS11y + v[5] * dFy[n];
  }
  const double S10 = S01;
  const double WiN = prm.WiN0 * EU;
  // traceless gradient (N = 2) and G_dot_tr
  const double tr = GU00 + GU11;
  const double htr = (tr * 2.0) / (2.0 * 2.0);
  const double Gt00 = GU00 - htr, Gt01 = GU01, Gt10 = GU10, Gt11 = GU11 - htr;
  const double Gd00 = Gt00 + Gt00, Gd01 = Gt01 + Gt10, Gd10 = Gt10 + Gt01, Gd11 = Gt11 + Gt11;
  // C = S.S
  const double C00 = S00 * S00 + S01 * S10, C01 = S00 * S01 + S01 * S11;
  const double C10 = S10 * S00 + S11 * S10, C11 = S10 * S01 + S11 * S11;
  double f0, f1, f2, f_T;
  if (prm.model == 0) {
    f_T = 1.0 + prm.VeN * ((C00 + C11) - 2.0);
    f0 = f_T;
    f1 = f_T;
    f2 = 0.0;
  } else {
    f_T = 1.0 - 2.0 * prm.VeN;
    f0 = 1.0 - prm.VeN;
    f1 = 1.0 - 2.0 * prm.VeN;
    f2 = prm.VeN;
  }
  // T = ((1-BvN)/WiN)(C - I)
  const double a = prm.omb / WiN;
  const double T00 = a * (C00 - 1.0), T01 = a * (C01 - 0.0), T10 = a * (C10 - 0.0), T11 = a * (C11 - 1.0);
  const double T22 = a * (0.0 - 1.0);
  // setOmegaTensor (traceless transposed gradient) and WTensor2DStability's base numerator
  o.W01 = (-(S00 * Gt01) + S01 * (Gt00 - Gt11) + S11 * Gt10) / (S00 + S11);
  o.nb = -(S00 * GU01) + S01 * (GU00 - GU11) + S11 * GU10;
  o.db = S00 + S11;
  // setSInverse (inverseTensor2D)
  const double det = S00 * S11 - S01 * S10;
  o.SI00 = S11 / det;
  o.SI01 = -S01 / det;
  o.SI11 = S00 / det;
  // setStabilizationTerms (Helem**2 = H2; (ReN/dt)**2 = (WiN/dt)**2 = 0 with dt = +huge;
  // tdcs = 0 so the DCS part of tlsce is zero)
  double tsugn = fabs(Ux * dFx[0] + Uy * dFy[0]);
  for (int n = 1; n < 9; n++) tsugn = tsugn + fabs(Ux * dFx[n] + Uy * dFy[n]);
  const double ddT = T00 * T00 + T01 * T10 + T10 * T01 + T11 * T11 + T22 * T22;
  const double ddG = Gd00 * Gd00 + Gd01 * Gd10 + Gd10 * Gd01 + Gd11 * Gd11;
  double Ha = sqrt(a * a + 0.5 * ddT);
  Ha = cdiv(Ha, sqrt(1.0 + 0.5 * ddG));
  const double r2 = prm.ReN * tsugn, r3 = prm.BvN / H2, r4 = cdiv(Ha, H2);
  const double tlsme = cdiv(1.0, sqrt(r2 * r2 + r3 * r3 + r4 * r4));
  o.tlsme = tlsme;
  o.tlsic = cdiv(H2, tlsme);
  const double ddC = C00 * C00 + C01 * C10 + C10 * C01 + C11 * C11;
  const double t2 = WiN * tsugn, t3 = WiN * sqrt(0.5 * ddG), t4 = f1, t5 = f2 * sqrt(0.5 * ddC);
  const double t7 = prm.DfN / H2;
  o.tlsce = cdiv(1.0, sqrt(t2 * t2 + t3 * t3 + t4 * t4 + t5 * t5 + t7 * t7));

  o.U0 = Ux;
  o.U1 = Uy;
  o.GU00 = GU00;
  o.GU01 = GU01;
  o.GU10 = GU10;
  o.GU11 = GU11;
  o.S00 = S00;
  o.S01 = S01;
  o.S11 = S11;
  o.dS00x = dS00x;
  o.dS00y = dS00y;
  o.dS01x = dS01x;
  o.dS01y = dS01y;
  o.dS11x = dS11x;
  o.dS11y = dS11y;
  o.C00 = C00;
  o.C01 = C01;
  o.C11 = C11;
  o.f0 = f0;
  o.f1 = f1;
  o.f2 = f2;
  o.f_T = f_T;
  o.WiN = WiN;
}

// Disturbance quantities of one Gauss point that the test-node terms read.
struct Dist {
  double P00, P01, P10, P11;      // P_tot
  double CE00, CE01, CE10, CE11;  // Con_EQ
  double CT00, CT01, CT11;        // Con_EQ_time = WiN*S (symmetric)
  double Mom0, Mom1, Mt0, Mt1;    // Momentum, Momentum_time
  double RC0, RC1;                // ReN*Convection
  double Cont;
  double dS00x, dS00y, dS01x, dS01y, dS11x, dS11y;
};

// setFlowUnknownsStability .. strongFormContinuity for the disturbance whose
// only nonzero nodal value is eps_jac at (inod, ieq).
__host__ __device__ inline void disturbance(const Base& b, int inod, int ieq, const double* bf, const double* dFx,
                                            const double* dFy, const double* d2x, const double* d2y,
                                            const Params& prm, Dist& o) {
  const double vb = kEpsJac * bf[inod], vx = kEpsJac * dFx[inod], vy = kEpsJac * dFy[inod];
  const double v2x = kEpsJac * d2x[inod], v2y = kEpsJac * d2y[inod];
  auto pick = [ieq](int q, double v) { return ieq == q ? 0.0 + v : 0.0; };
  const double U0 = pick(0, vb), U1 = pick(1, vb);
  const double GU00 = pick(0, vx), GU10 = pick(0, vy), GU01 = pick(1, vx), GU11 = pick(1, vy);
  const double G2x0 = pick(0, v2x), G2y0 = pick(0, v2y), G2x1 = pick(1, v2x), G2y1 = pick(1, v2y);
  const double P = pick(2, vb), dPx = pick(2, vx), dPy = pick(2, vy);
  const double S00 = pick(3, vb), S01 = pick(4, vb), S11 = pick(5, vb), S10 = S01;
  const double dS00x = pick(3, vx), dS00y = pick(3, vy), dS01x = pick(4, vx), dS01y = pick(4, vy);
  const double dS11x = pick(5, vx), dS11y = pick(5, vy);
  const double dS10x = dS01x, dS10y = dS01y;
  const double bS10 = b.S01, bdS10x = b.dS01x, bdS10y = b.dS01y;
  const double WiN = b.WiN;

  // setFlowKinematics: G_dot = GU + GU^T
  const double Gd00 = GU00 + GU00, Gd01 = GU01 + GU10, Gd10 = GU10 + GU01, Gd11 = GU11 + GU11;

  // setCTensorStabilitySQRT: C_d = S_b.S_d + S_d.S_b
  const double M00 = b.S00 * S00 + b.S01 * S10, M01 = b.S00 * S01 + b.S01 * S11;
  const double M10 = bS10 * S00 + b.S11 * S10, M11 = bS10 * S01 + b.S11 * S11;
  const double C00 = M00 + (S00 * b.S00 + S01 * bS10);
  const double C01 = M01 + (S00 * b.S01 + S01 * b.S11);
  const double C11 = M11 + (S10 * b.S01 + S11 * b.S11);
  // dC_d/dx_k = dS_b,k.S_d + S_b.dS_d,k + dS_d,k.S_b + S_d.dS_b,k, the entries
  // (1,1,x), (2,1,y), (1,2,x), (2,2,y) that div(T) reads
  const double dC00x = (((b.dS00x * S00 + b.dS01x * S10) + (b.S00 * dS00x + b.S01 * dS10x)) +
                        (dS00x * b.S00 + dS01x * bS10)) +
                       (S00 * b.dS00x + S01 * bdS10x);
  const double dC10y = (((bdS10y * S00 + b.dS11y * S10) + (bS10 * dS00y + b.S11 * dS10y)) +
                        (dS10y * b.S00 + dS11y * bS10)) +
                       (S10 * b.dS00y + S11 * bdS10y);
  const double dC01x = (((b.dS00x * S01 + b.dS01x * S11) + (b.S00 * dS01x + b.S01 * dS11x)) +
                        (dS00x * b.S01 + dS01x * b.S11)) +
                       (S00 * b.dS01x + S01 * b.dS11x);
  const double dC11y = (((bdS10y * S01 + b.dS11y * S11) + (bS10 * dS01y + b.S11 * dS11y)) +
                        (dS10y * b.S01 + dS11y * b.S11)) +
                       (S10 * b.dS01y + S11 * b.dS11y);

  // conEQ_L_PTT_Stability_SQRT / conEQ_Giesekus_Stability_SQRT
  double f0 = 0.0, f1 = 0.0;
  if (prm.model == 0) {
    const double f_T = prm.VeN * (C00 + C11);
    f0 = f_T;
    f1 = f_T;
  }

  // setStressTensorStabilitySQRT: T_d = ((1-BvN)*C_d)/WiN, P_tot_d = T_d + BvN*G_dot - P I
  const double T00 = cdiv(prm.omb * C00, WiN), T01 = cdiv(prm.omb * C01, WiN), T11 = cdiv(prm.omb * C11, WiN);
  o.P00 = (T00 + prm.BvN * Gd00) - P;
  o.P01 = T01 + prm.BvN * Gd01;
  o.P10 = T01 + prm.BvN * Gd10;
  o.P11 = (T11 + prm.BvN * Gd11) - P;

  // strongFormMomentumStability
  const double divT0 = cdiv(prm.omb * dC00x, WiN) + cdiv(prm.omb * dC10y, WiN);
  const double divT1 = cdiv(prm.omb * dC01x, WiN) + cdiv(prm.omb * dC11y, WiN);
  const double dP0 = (-dPx + divT0) + prm.BvN * (G2x0 + G2y0);
  const double dP1 = (-dPy + divT1) + prm.BvN * (G2x1 + G2y1);
  const double Cv0 = (b.U0 * GU00 + b.U1 * GU10) + (U0 * b.GU00 + U1 * b.GU10);
  const double Cv1 = (b.U0 * GU01 + b.U1 * GU11) + (U0 * b.GU01 + U1 * b.GU11);
  o.RC0 = prm.ReN * Cv0;
  o.RC1 = prm.ReN * Cv1;
  o.Mom0 = o.RC0 - dP0;
  o.Mom1 = o.RC1 - dP1;
  o.Mt0 = prm.ReN * U0;
  o.Mt1 = prm.ReN * U1;
  o.Cont = GU00 + GU11;

  // setUCDStabilitySQRT
  double U00 = -(b.S00 * GU00 + b.S01 * GU10) - (S00 * b.GU00 + S01 * b.GU10);
  double U01 = -(b.S00 * GU01 + b.S01 * GU11) - (S00 * b.GU01 + S01 * b.GU11);
  double U10 = -(bS10 * GU00 + b.S11 * GU10) - (S10 * b.GU00 + S11 * b.GU10);
  double U11 = -(bS10 * GU01 + b.S11 * GU11) - (S10 * b.GU01 + S11 * b.GU11);
  U00 = (U00 + b.U0 * dS00x) + U0 * b.dS00x;
  U01 = (U01 + b.U0 * dS01x) + U0 * b.dS01x;
  U10 = (U10 + b.U0 * dS10x) + U0 * bdS10x;
  U11 = (U11 + b.U0 * dS11x) + U0 * b.dS11x;
  U00 = (U00 + b.U1 * dS00y) + U1 * b.dS00y;
  U01 = (U01 + b.U1 * dS01y) + U1 * b.dS01y;
  U10 = (U10 + b.U1 * dS10y) + U1 * bdS10y;
  U11 = (U11 + b.U1 * dS11y) + U1 * b.dS11y;
  // WTensor2DStability with GUT_b = GU_b^T and GUT_d = GU_d^T
  const double nd = (-(b.S00 * GU01) + b.S01 * (GU00 - GU11) + b.S11 * GU10) +
                    (-(S00 * b.GU01) + S01 * (b.GU00 - b.GU11) + S11 * b.GU10);
  const double dd = S00 + S11;
  const double Wd01 = cdiv(nd * b.db - b.nb * dd, b.db * b.db);
  const double Wd10 = -Wd01, Wb01 = b.W01, Wb10 = -b.W01;
  U00 = (U00 - Wb01 * S10) - Wd01 * bS10;
  U01 = (U01 - Wb01 * S11) - Wd01 * b.S11;
  U10 = (U10 - Wb10 * S00) - Wd10 * b.S00;
  U11 = (U11 - Wb10 * S01) - Wd10 * b.S01;

  // setSInverseStability: SI_d = -SI_b.(S_d.SI_b)
  const double bSI10 = b.SI01;
  const double N00 = S00 * b.SI00 + S01 * bSI10, N01 = S00 * b.SI01 + S01 * b.SI11;
  const double N10 = S10 * b.SI00 + S11 * bSI10, N11 = S10 * b.SI01 + S11 * b.SI11;
  const double SI00 = -(b.SI00 * N00 + b.SI01 * N10), SI01 = -(b.SI00 * N01 + b.SI01 * N11);
  const double SI10 = -(bSI10 * N00 + b.SI11 * N10), SI11 = -(bSI10 * N01 + b.SI11 * N11);
  // S3_d = (S_b.S_b).S_d + (S_b.S_d).S_b + S_d.(S_b.S_b)
  const double bC10 = b.C01;
  const double S3_00 = ((b.C00 * S00 + b.C01 * S10) + (M00 * b.S00 + M01 * bS10)) + (S00 * b.C00 + S01 * bC10);
  const double S3_01 = ((b.C00 * S01 + b.C01 * S11) + (M00 * b.S01 + M01 * b.S11)) + (S00 * b.C01 + S01 * b.C11);
  const double S3_10 = ((bC10 * S00 + b.C11 * S10) + (M10 * b.S00 + M11 * bS10)) + (S10 * b.C00 + S11 * bC10);
  const double S3_11 = ((bC10 * S01 + b.C11 * S11) + (M10 * b.S01 + M11 * b.S11)) + (S10 * b.C01 + S11 * b.C11);
  // term_S = f3_b S3_d + f2_d S_b + f2_b S_d - f1_b SI_d - f1_d SI_b
  const double tS00 = (((b.f2 * S3_00 + f1 * b.S00) + b.f1 * S00) - b.f0 * SI00) - f0 * b.SI00;
  const double tS01 = (((b.f2 * S3_01 + f1 * b.S01) + b.f1 * S01) - b.f0 * SI01) - f0 * b.SI01;
  const double tS10 = (((b.f2 * S3_10 + f1 * bS10) + b.f1 * S10) - b.f0 * SI10) - f0 * bSI10;
  const double tS11 = (((b.f2 * S3_11 + f1 * b.S11) + b.f1 * S11) - b.f0 * SI11) - f0 * b.SI11;
  o.CE00 = WiN * U00 + cdiv(tS00, 2.0);
  o.CE01 = WiN * U01 + cdiv(tS01, 2.0);
  o.CE10 = WiN * U10 + cdiv(tS10, 2.0);
  o.CE11 = WiN * U11 + cdiv(tS11, 2.0);
  o.CT00 = WiN * S00;
  o.CT01 = WiN * S01;
  o.CT11 = WiN * S11;
  o.dS00x = dS00x;
  o.dS00y = dS00y;
  o.dS01x = dS01x;
  o.dS01y = dS01y;
  o.dS11x = dS11x;
  o.dS11y = dS11y;
}

// Test-node coefficients that depend only on the base state, the test node
// and the parameters, not on the disturbance: the LSME momentum factor, the
// LSIC factors, the LSCE momentum operators of GUT_lsce rows 1 and 2 and the
// LSCE stress factor (src/Stability/Residuals_Stability_mod.F90:209-313).
// Every one of them is a leading subexpression of the reference's term, so
// computing it once per (Gauss point, test node) and reusing it for all 54
// disturbances gives the same bits.  R is double, or a counting type for the
// operation count.
template <class R>
struct NodeCoefT {
  R lsme_m;          // (tlsme*ReN)*(U.grad phi)
  R licx, licy;      // tlsic*fx, tlsic*fy
  R l1[4], l2[4];    // LSCEMomentumSQRT rows 1 and 2: l00, l01, l10, l11
  R lsce_s;          // tlsce*(WiN*(U.grad phi) + f_T*phi/2)
  R fx, fy, bifn;    // the test function and its gradient
  R pad;
};
using NodeCoef = NodeCoefT<double>;

template <class R, class B>
__host__ __device__ inline R cdiv_t(const R& a, const B& c) { return (a * c) / (c * c); }

template <class R>
__host__ __device__ inline void node_coef(const R& U0, const R& U1, const R& S00, const R& S01, const R& S11,
                                          const R& WiN, const R& tlsme, const R& tlsic, const R& tlsce,
                                          const R& f_T, double bifn, double fx, double fy, const Params& prm,
                                          NodeCoefT<R>& c) {
  const R Udf = U0 * fx + U1 * fy;
  c.lsme_m = (tlsme * prm.ReN) * Udf;  // LSMEMomentum(dFdXi, GsPt_b)
  const R bS10 = S01;
  const R rW = S00 + S11, WW = WiN * WiN;
  {
    // LSCEMomentumSQRT(1, dFdXi, GsPt_b): GUT_lsce row 1 = grad(phi)
    const R w = (S01 * fx + S11 * fy) / rW;
    const R m00 = S00 * fx + S01 * fy, m10 = bS10 * fx + S11 * fy;
    const R u00 = -m00 - w * bS10, u01 = -(w * S11);
    const R u10 = -m10 - (-w) * S00, u11 = -((-w) * S01);
    c.l1[0] = ((tlsce * u00) * prm.omb * WiN) / WW;
    c.l1[1] = ((tlsce * u01) * prm.omb * WiN) / WW;
    c.l1[2] = ((tlsce * u10) * prm.omb * WiN) / WW;
    c.l1[3] = ((tlsce * u11) * prm.omb * WiN) / WW;
    c.licx = tlsic * fx;
  }
  {
    // LSCEMomentumSQRT(2, dFdXi, GsPt_b): GUT_lsce row 2 = grad(phi)
    const R w = (-(S00 * fx) + S01 * (-fy)) / rW;
    const R m01 = S00 * fx + S01 * fy, m11 = bS10 * fx + S11 * fy;
    const R u00 = -(w * bS10), u01 = -m01 - w * S11;
    const R u10 = -((-w) * S00), u11 = -m11 - (-w) * S01;
    c.l2[0] = ((tlsce * u00) * prm.omb * WiN) / WW;
    c.l2[1] = ((tlsce * u01) * prm.omb * WiN) / WW;
    c.l2[2] = ((tlsce * u10) * prm.omb * WiN) / WW;
    c.l2[3] = ((tlsce * u11) * prm.omb * WiN) / WW;
    c.licy = tlsic * fy;
  }
  // LSCEStressSQRT(bifn, dFdXi, GsPt_b)
  c.lsce_s = tlsce * (WiN * Udf + cdiv_t(f_T * bifn, 2.0));
  c.fx = fx;
  c.fy = fy;
  c.bifn = bifn;
  c.pad = 0.0;
}

__host__ __device__ inline void node_coef(const Base& b, double bifn, double fx, double fy, const Params& prm,
                                          NodeCoef& c) {
  node_coef<double>(b.U0, b.U1, b.S00, b.S01, b.S11, b.WiN, b.tlsme, b.tlsic, b.tlsce, b.f_T, bifn, fx, fy, prm, c);
}

// Test-node rows of one Gauss point (momentumStabilitySQRT, continuityStability,
// stressStabilitySQRT), times WET, from the coefficients above and the
// disturbance: a[6] for A, bb[6] for B (negated, as
// equationsResidualsBulkStability does).
__host__ __device__ inline void node_terms_c(const NodeCoef& k, const Dist& d, double tlsme, double WiN, double WET,
                                             const Params& prm, double* a, double* bb) {
  const double bifn = k.bifn, fx = k.fx, fy = k.fy;
  double TA0, TA1, TB0, TB1;
  {
    const double lsic = k.licx * d.Cont;
    const double lce = k.l1[0] * d.CE00 + k.l1[1] * d.CE10 + k.l1[2] * d.CE01 + k.l1[3] * d.CE11;
    TA0 = (((d.RC0 * bifn + (d.P00 * fx + d.P01 * fy)) + k.lsme_m * d.Mom0) + lsic) + lce;
    const double lct = k.l1[0] * d.CT00 + k.l1[1] * d.CT01 + k.l1[2] * d.CT01 + k.l1[3] * d.CT11;
    TB0 = (d.Mt0 * bifn + k.lsme_m * d.Mt0) + lct;
  }
  {
    const double lsic = k.licy * d.Cont;
    const double lce = k.l2[0] * d.CE00 + k.l2[1] * d.CE10 + k.l2[2] * d.CE01 + k.l2[3] * d.CE11;
    TA1 = (((d.RC1 * bifn + (d.P10 * fx + d.P11 * fy)) + k.lsme_m * d.Mom1) + lsic) + lce;
    const double lct = k.l2[0] * d.CT00 + k.l2[1] * d.CT01 + k.l2[2] * d.CT01 + k.l2[3] * d.CT11;
    TB1 = (d.Mt1 * bifn + k.lsme_m * d.Mt1) + lct;
  }
  const double TA2 = d.Cont * bifn + tlsme * (fx * d.Mom0 + fy * d.Mom1);
  const double TB2 = tlsme * (fx * d.Mt0 + fy * d.Mt1);
  const double d00 = d.dS00x * fx + d.dS00y * fy;
  const double d01 = d.dS01x * fx + d.dS01y * fy;
  const double d11 = d.dS11x * fx + d.dS11y * fy;
  // LSMEStressSQRT with Momentum_d and with Momentum_time_d
  const double l3 = (tlsme * (-(fx * d.Mom0))) * WiN;
  const double l4 = (tlsme * (-(fy * d.Mom0) - fx * d.Mom1)) * WiN;
  const double l5 = (tlsme * (-(fy * d.Mom1))) * WiN;
  const double k3 = (tlsme * (-(fx * d.Mt0))) * WiN;
  const double k4 = (tlsme * (-(fy * d.Mt0) - fx * d.Mt1)) * WiN;
  const double k5 = (tlsme * (-(fy * d.Mt1))) * WiN;
  const double TA3 = ((d.CE00 * bifn + prm.DfN * d00) + l3) + k.lsce_s * d.CE00;
  const double TA4 = ((d.CE01 * bifn + prm.DfN * d01) + l4) + k.lsce_s * d.CE01;
  const double TA5 = ((d.CE11 * bifn + prm.DfN * d11) + l5) + k.lsce_s * d.CE11;
  const double TB3 = (d.CT00 * bifn + k3) + k.lsce_s * d.CT00;
  const double TB4 = (d.CT01 * bifn + k4) + k.lsce_s * d.CT01;
  const double TB5 = (d.CT11 * bifn + k5) + k.lsce_s * d.CT11;
  a[0] = TA0 * WET;
  a[1] = TA1 * WET;
  a[2] = TA2 * WET;
  a[3] = TA3 * WET;
  a[4] = TA4 * WET;
  a[5] = TA5 * WET;
  bb[0] = (-TB0) * WET;
  bb[1] = (-TB1) * WET;
  bb[2] = (-TB2) * WET;
  bb[3] = (-TB3) * WET;
  bb[4] = (-TB4) * WET;
  bb[5] = (-TB5) * WET;
}

// Both parts for one (Gauss point, test node, disturbance).
__host__ __device__ inline void node_terms(const Base& b, const Dist& d, double bifn, double fx, double fy,
                                           double WET, const Params& prm, double* a, double* bb) {
  NodeCoef k;
  node_coef(b, bifn, fx, fy, prm, k);
  node_terms_c(k, d, b.tlsme, b.WiN, WET, prm, a, bb);
}

// (dRes - Res)/eps_jac with Res = 0 (the zero disturbance gives zero residuals).
__host__ __device__ inline double fd_quotient(double r) { return cdiv(r - 0.0, kEpsJac); }

}  // namespace vefstab

#endif
