// =============================================================================
// TEST INFRASTRUCTURE ONLY — the CPU oracle.
//
// A plain C++ restatement of the reference's element-level residual and
// finite-difference Jacobian assembly for the 2D square-root-conformation
// viscoelastic cylinder problem (Cylinder_2D, 9-node quadratic quadrangles,
// Stress_Reform='SQRT', Continuation_Method='Arclength').  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
// as the checker.  The product (the HIP library) never calls into this file.
//
// It follows the reference line by line, keeping its data shapes (3x3 tensors
// whose third row/column is zero in 2D, Fortran's sequential dot_product /
// matmul / traceTensor / doubleDotProduct summation order, left-to-right
// evaluation) so that it reproduces the reference's FP64 operation sequence:
//   src/GaussIntegration/Gauss_Quadrangle_mod.F90:54-101,133-198,243-372,405-478
//   src/Element_Calculations_mod.F90:106-142,308-361,449-533,537-634,658-773,852-924
//   src/Constitutive_Models_SQRT_mod.F90:120-180,285-305,379-530,534-584
//   src/Residuals_mod.F90:161-187,223-234,295-490,599-624
//   src/Equations_mod.F90:21-144,226-313,317-554
//   src/Storage_mod.F90:10-115, src/Tools_mod.F90:129-146,206-340,735-855
//   src/Petsc_mod.F90:629-743 (matrix pattern)
//
// Scalar semantics fixed here (the reference is compiled with PetscScalar =
// complex(8); all imaginary parts are zero on this path):
//   * complex add/sub/mul/sqrt with zero imaginary parts equal the real ops;
//   * every division with a complex operand rounds as (a*c)/(c*c) (checked for
//     flang 22 on 2e5 random operands for complex/complex, complex/real and
//     real/complex), written cdiv() below; real/real divisions stay a/c;
//   * no FMA contraction (build with -ffp-contract=off), IEEE div and sqrt.
// =============================================================================
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef VEF_COUNT_FLOPS
#include "flop_real.hpp"
typedef FlopReal vreal;
#else
typedef double vreal;
static inline double vsqrt(double a) { return std::sqrt(a); }
static inline double fabsr(double a) { return std::fabs(a); }
static inline double val(double a) { return a; }
#endif

static inline vreal Z() { return vreal(); }  // zero (structural zero when counting)

// Division where either operand is PetscScalar (complex(8)) with zero
// imaginary parts: flang evaluates re((a+0i)/(c+0i)) as (a*c)/(c*c).
static inline vreal cdiv(vreal a, vreal c) { return (a * c) / (c * c); }

extern "C" {

// Physical parameters (Parameters_mod.F90:98-111, Physical_mod.F90:327-351).
typedef struct {
  double ReN, BvN, VeN, DfN, WiN0, dt;
  int32_t model;  // 0: L-PTT (VeN = epsilon; VeN=0 is Oldroyd-B), 1: Giesekus (VeN = alpha)
  int32_t pad;
} vo_params;

}  // extern "C"

// ----------------------------------------------------------------------------
// a1: Gauss tables, Gauss_Quadrangle_mod.F90
// ----------------------------------------------------------------------------
struct Tables {
  double P1[3], W1[3];
  double P2[9][2], W2[9];
  // [gauss point][node]
  double bfn[9][9], dfdc[9][9], dfde[9][9], d2fdc2[9][9], d2fde2[9][9], d2fdce[9][9];
  // [edge 0..3][gauss point][node]
  double bfnE[4][3][9], dfdcE[4][3][9], dfdeE[4][3][9], d2fdc2E[4][3][9], d2fde2E[4][3][9],
      d2fdceE[4][3][9];
};

// setQuadraticQuadrangle, Gauss_Quadrangle_mod.F90:405-478
static void quadratic_quadrangle(double c, double e, double* bfn, double* dfdc, double* dfde,
                                 double* d2fdc2, double* d2fde2, double* d2fdce) {
  bfn[0] = +(1.0 - c) * c * (1.0 - e) * e / 4.0;
  bfn[1] = -((1.0 + c) * c * (1.0 - e) * e / 4.0);
  bfn[2] = +(1.0 + c) * c * (1.0 + e) * e / 4.0;
  bfn[3] = -((1.0 - c) * c * (1.0 + e) * e / 4.0);
  bfn[4] = -((1.0 + c) * (1.0 - c) * (1.0 - e) * e / 2.0);
  bfn[5] = +(1.0 + c) * c * (1.0 + e) * (1.0 - e) / 2.0;
  bfn[6] = +(1.0 + c) * (1.0 - c) * (1.0 + e) * e / 2.0;
  bfn[7] = -((1.0 - c) * c * (1.0 + e) * (1.0 - e) / 2.0);
  bfn[8] = +(1.0 + c) * (1.0 - c) * (1.0 + e) * (1.0 - e);

  dfdc[0] = +(1.0 - 2.0 * c) * (1.0 - e) * e / 4.0;
  dfdc[1] = -((1.0 + 2.0 * c) * (1.0 - e) * e / 4.0);
  dfdc[2] = +(1.0 + 2.0 * c) * (1.0 + e) * e / 4.0;
  dfdc[3] = -((1.0 - 2.0 * c) * (1.0 + e) * e / 4.0);
  dfdc[4] = -((-2.0 * c) * (1.0 - e) * e / 2.0);
  dfdc[5] = +(1.0 + 2.0 * c) * (1.0 + e) * (1.0 - e) / 2.0;
  dfdc[6] = +(-2.0 * c) * (1.0 + e) * e / 2.0;
  dfdc[7] = -((1.0 - 2.0 * c) * (1.0 + e) * (1.0 - e) / 2.0);
  dfdc[8] = +(-2.0 * c) * (1.0 + e) * (1.0 - e);

  d2fdc2[0] = +(-2.0) * (1.0 - e) * e / 4.0;
  d2fdc2[1] = -((+2.0) * (1.0 - e) * e / 4.0);
  d2fdc2[2] = +(+2.0) * (1.0 + e) * e / 4.0;
  d2fdc2[3] = -((-2.0) * (1.0 + e) * e / 4.0);
  d2fdc2[4] = -((-2.0) * (1.0 - e) * e / 2.0);
  d2fdc2[5] = +(+2.0) * (1.0 + e) * (1.0 - e) / 2.0;
  d2fdc2[6] = +(-2.0) * (1.0 + e) * e / 2.0;
  d2fdc2[7] = -((-2.0) * (1.0 + e) * (1.0 - e) / 2.0);
  d2fdc2[8] = +(-2.0) * (1.0 + e) * (1.0 - e);

  d2fdce[0] = +(1.0 - 2.0 * c) * (1.0 - 2.0 * e) / 4.0;
  d2fdce[1] = -((1.0 + 2.0 * c) * (1.0 - 2.0 * e) / 4.0);
  d2fdce[2] = +(1.0 + 2.0 * c) * (1.0 + 2.0 * e) / 4.0;
  d2fdce[3] = -((1.0 - 2.0 * c) * (1.0 + 2.0 * e) / 4.0);
  d2fdce[4] = -((-2.0 * c) * (1.0 - 2.0 * e) / 2.0);
  d2fdce[5] = +(1.0 + 2.0 * c) * (-2.0 * e) / 2.0;
  d2fdce[6] = +(-2.0 * c) * (1.0 + 2.0 * e) / 2.0;
  d2fdce[7] = -((1.0 - 2.0 * c) * (-2.0 * e) / 2.0);
  d2fdce[8] = +(-2.0 * c) * (-2.0 * e);

  dfde[0] = +(1.0 - c) * c * (1.0 - 2.0 * e) / 4.0;
  dfde[1] = -((1.0 + c) * c * (1.0 - 2.0 * e) / 4.0);
  dfde[2] = +(1.0 + c) * c * (1.0 + 2.0 * e) / 4.0;
  dfde[3] = -((1.0 - c) * c * (1.0 + 2.0 * e) / 4.0);
  dfde[4] = -((1.0 + c) * (1.0 - c) * (1.0 - 2.0 * e) / 2.0);
  dfde[5] = +(1.0 + c) * c * (-2.0 * e) / 2.0;
  dfde[6] = +(1.0 + c) * (1.0 - c) * (1.0 + 2.0 * e) / 2.0;
  dfde[7] = -((1.0 - c) * c * (-2.0 * e) / 2.0);
  dfde[8] = +(1.0 + c) * (1.0 - c) * (-2.0 * e);

  d2fde2[0] = +(1.0 - c) * c * (-2.0) / 4.0;
  d2fde2[1] = -((1.0 + c) * c * (-2.0) / 4.0);
  d2fde2[2] = +(1.0 + c) * c * (+2.0) / 4.0;
  d2fde2[3] = -((1.0 - c) * c * (+2.0) / 4.0);
  d2fde2[4] = -((1.0 + c) * (1.0 - c) * (-2.0) / 2.0);
  d2fde2[5] = +(1.0 + c) * c * (-2.0) / 2.0;
  d2fde2[6] = +(1.0 + c) * (1.0 - c) * (+2.0) / 2.0;
  d2fde2[7] = -((1.0 - c) * c * (-2.0) / 2.0);
  d2fde2[8] = +(1.0 + c) * (1.0 - c) * (-2.0);
}

static void build_tables(Tables& T) {
  // setGaussPointsQuadrangle1D (NGauss_1D = 3), :69-72
  T.P1[0] = 0.0 + 1.0e-8;
  T.P1[1] = +std::sqrt(3.0 / 5.0);
  T.P1[2] = -std::sqrt(3.0 / 5.0);
  // setGaussWeightsQuadrangle1D, :166-169
  T.W1[0] = 8.0 / 9.0;
  T.W1[1] = 5.0 / 9.0;
  T.W1[2] = 5.0 / 9.0;
  // 2D points / weights, :92-99 and :189-196 (i: xi outer, j: eta inner)
  int ng = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      T.P2[ng][0] = T.P1[i];
      T.P2[ng][1] = T.P1[j];
      T.W2[ng] = T.W1[i] * T.W1[j];
      ng++;
    }
  for (int g = 0; g < 9; g++)
    quadratic_quadrangle(T.P2[g][0], T.P2[g][1], T.bfn[g], T.dfdc[g], T.dfde[g], T.d2fdc2[g],
                         T.d2fde2[g], T.d2fdce[g]);
  // Edge tables, :261-312 (ied 1: eta=-1, 2: xi=+1, 3: eta=+1, 4: xi=-1)
  for (int g = 0; g < 3; g++)
    for (int ed = 0; ed < 4; ed++) {
      double c = 0, e = 0;
      switch (ed) {
        case 0: c = T.P1[g]; e = -1.0; break;
        case 1: c = +1.0; e = T.P1[g]; break;
        case 2: c = T.P1[g]; e = +1.0; break;
        default: c = -1.0; e = T.P1[g]; break;
      }
      quadratic_quadrangle(c, e, T.bfnE[ed][g], T.dfdcE[ed][g], T.dfdeE[ed][g], T.d2fdc2E[ed][g],
                           T.d2fde2E[ed][g], T.d2fdceE[ed][g]);
    }
}

static Tables make_tables() {
  Tables T;
  build_tables(T);
  return T;
}
static const Tables& tables() {
  static const Tables T = make_tables();  // thread-safe initialisation
  return T;
}

// ----------------------------------------------------------------------------
// Element state (NodeArrays, Global_Variables_mod.F90:86-110) and Gauss-point
// quantities (GaussPointQuantities, :56-84), real parts only.
// ----------------------------------------------------------------------------
struct Elem {
  vreal TL[9][6];  // TEMP_TL(inod, ieq): Ux Uy P Sxx Sxy Syy
  vreal EU;        // TEMP_EU(Nex)
  double X[9][3];  // Xi_loc(inod, 1:3), z = 0
  vreal dFdXi[3][9], d2FdXi2[3][9];
  const double *bfn, *dfdc, *dfde, *d2fdc2, *d2fde2, *d2fdce;
};

struct GsPt {
  vreal GU[3][3], GUT[3][3], GU_tr[3][3], GUT_tr[3][3], G_dot_tr[3][3], P_tot[3][3], G2U[3][3];
  vreal U_f[3], dPdXi[3], Momentum[3], Convection[3];
  vreal Continuity, P;
  vreal dTdXi[3][3][3], dCdXi[3][3][3], dSdXi[3][3][3];
  vreal S[3][3], Con_EQ[3][3], T_ve[3][3], C[3][3];
  vreal f[3], g[2], dgdxi[2][3], f_T;
  vreal normal[3], dL;
  vreal dXidci[3][3];
  vreal detJ, WET;
  vreal mag_re, mag_im;  // mag_con_EQ (complex sqrt of a real; purely imaginary if negative)
  vreal tlsic, tlsme, tlsce, tdcs;
  vreal H2, tsugn;       // Helem**2 and the streamline sum of setStabilizationTerms
  vreal WiN;             // module global WiN, rewritten per Gauss point
};

struct Phys {
  vreal ReN, BvN, VeN, DfN, WiN0, dt;
  int model;
  bool dcs_off;  // DfN > 1e-12 switches tdcs off (Element_Calculations_mod.F90:758-760)
};

static Phys make_phys(const vo_params* p) {
  Phys q;
  q.ReN = (p->ReN == 0.0) ? Z() : vreal(p->ReN);
  q.BvN = vreal(p->BvN);
  q.VeN = (p->VeN == 0.0) ? Z() : vreal(p->VeN);
  q.DfN = (p->DfN == 0.0) ? Z() : vreal(p->DfN);
  q.WiN0 = vreal(p->WiN0);
  q.dt = vreal(p->dt);
  q.model = p->model;
  q.dcs_off = p->DfN > 1.0e-12;
  return q;
}

static void zero33(vreal A[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) A[i][j] = Z();
}

// Fortran matmul of 3x3 arrays: C(i,j) = sum_k A(i,k)*B(k,j), k sequential from 0.
static void matmul33(const vreal A[3][3], const vreal B[3][3], vreal Cm[3][3]) {
  vreal R[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      vreal s = Z();
      for (int k = 0; k < 3; k++) s = s + A[i][k] * B[k][j];
      R[i][j] = s;
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Cm[i][j] = R[i][j];
}

// traceTensor, Tools_mod.F90:206-238
static vreal trace33(const vreal A[3][3]) {
  vreal s = Z();
  for (int i = 0; i < 3; i++) s = s + A[i][i];
  return s;
}

// doubleDotProduct, Tools_mod.F90:262-340: sum_i sum_j A(i,j)*B(j,i), i outer.
static vreal ddot33(const vreal A[3][3], const vreal B[3][3]) {
  vreal s = Z();
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) s = s + A[i][j] * B[j][i];
  return s;
}

// dot_product of two length-3 vectors, sequential from 0.
static vreal dot3(const vreal* a, const vreal* b) {
  vreal s = Z();
  for (int k = 0; k < 3; k++) s = s + a[k] * b[k];
  return s;
}

// dot_product(TEMP_TL(:,i), v(:)) over the 9 nodes.
static vreal dot9_col(const Elem& E, int ieq, const vreal* v) {
  vreal s = Z();
  for (int n = 0; n < 9; n++) s = s + E.TL[n][ieq] * v[n];
  return s;
}
static vreal dot9_col_d(const Elem& E, int ieq, const double* v) {
  vreal s = Z();
  for (int n = 0; n < 9; n++) s = s + E.TL[n][ieq] * vreal(v[n]);
  return s;
}

// ----------------------------------------------------------------------------
// a3: basis2D, Element_Calculations_mod.F90:308-361 (idir = [1,2]).
// GsPt%Xi (the physical position) is not used by the residual and is skipped.
// ----------------------------------------------------------------------------
static void basis2D(Elem& E, GsPt& G) {
  zero33(G.dXidci);
  for (int n = 0; n < 9; n++)
    for (int j = 0; j < 2; j++) {  // j = 3 (z) multiplies the zero z-coordinate
      G.dXidci[0][j] = G.dXidci[0][j] + vreal(E.dfdc[n]) * vreal(E.X[n][j]);
      G.dXidci[1][j] = G.dXidci[1][j] + vreal(E.dfde[n]) * vreal(E.X[n][j]);
    }
  vreal dxdc = G.dXidci[0][0], dydc = G.dXidci[0][1];
  vreal dxde = G.dXidci[1][0], dyde = G.dXidci[1][1];
  G.detJ = dxdc * dyde - dxde * dydc;
  vreal dcdx = dyde / G.detJ;
  vreal dedx = -(dydc / G.detJ);
  vreal dcdy = -(dxde / G.detJ);
  vreal dedy = dxdc / G.detJ;
  for (int r = 0; r < 3; r++)
    for (int n = 0; n < 9; n++) {
      E.dFdXi[r][n] = Z();
      E.d2FdXi2[r][n] = Z();
    }
  vreal two = vreal(2.0);
  for (int n = 0; n < 9; n++) {
    vreal fc = vreal(E.dfdc[n]), fe = vreal(E.dfde[n]);
    vreal fcc = vreal(E.d2fdc2[n]), fee = vreal(E.d2fde2[n]), fce = vreal(E.d2fdce[n]);
    E.dFdXi[0][n] = dcdx * fc + dedx * fe;
    E.dFdXi[1][n] = dcdy * fc + dedy * fe;
    E.d2FdXi2[0][n] = (dcdx * dcdx) * fcc + (dedx * dedx) * fee + two * dcdx * dedx * fce;
    E.d2FdXi2[1][n] = (dcdy * dcdy) * fcc + (dedy * dedy) * fee + two * dcdy * dedy * fce;
  }
}

// ----------------------------------------------------------------------------
// a4: setFlowUnknowns, Element_Calculations_mod.F90:449-533 (Neq_f=2, Neq_s=3)
// ----------------------------------------------------------------------------
static void setFlowUnknowns(const Elem& E, GsPt& G) {
  for (int i = 0; i < 3; i++) G.U_f[i] = Z();
  zero33(G.GU);
  zero33(G.G2U);
  G.P = Z();
  for (int i = 0; i < 3; i++) G.dPdXi[i] = Z();
  for (int i = 0; i < 2; i++) {
    G.U_f[i] = dot9_col_d(E, i, E.bfn);
    for (int j = 0; j < 3; j++) {
      G.GU[j][i] = dot9_col(E, i, E.dFdXi[j]);
      G.G2U[j][i] = dot9_col(E, i, E.d2FdXi2[j]);
    }
    G.dPdXi[i] = dot9_col(E, 2, E.dFdXi[i]);
  }
  G.P = dot9_col_d(E, 2, E.bfn);
  zero33(G.S);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      for (int c = 0; c < 3; c++) G.dSdXi[a][b][c] = Z();
  // getStressComponent order: (1,1), (1,2), (2,2), Tools_mod.F90:769-793
  const int comp[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  for (int s = 0; s < 3; s++) {
    int j = comp[s][0], k = comp[s][1];
    G.S[j][k] = dot9_col_d(E, 3 + s, E.bfn);
    G.S[k][j] = G.S[j][k];
    for (int d = 0; d < 3; d++) G.dSdXi[j][k][d] = dot9_col(E, 3 + s, E.dFdXi[d]);
    for (int d = 0; d < 3; d++) G.dSdXi[k][j][d] = G.dSdXi[j][k][d];
  }
}

// setFlowKinematics, Element_Calculations_mod.F90:605-634 (N = 2)
static void setFlowKinematics(GsPt& G, const vreal I[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) G.GUT[i][j] = G.GU[j][i];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) G.GU_tr[i][j] = G.GU[i][j];
  vreal tr = trace33(G.GU);
  vreal Nc = vreal(2.0);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) G.GU_tr[i][j] = G.GU[i][j] - cdiv(tr * I[i][j], Nc);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) G.GUT_tr[i][j] = G.GU_tr[j][i];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) G.G_dot_tr[i][j] = G.GU_tr[i][j] + G.GUT_tr[i][j];
}

// WTensor2D + omegaNominator2DReal/omegaDenominator2DReal,
// Constitutive_Models_SQRT_mod.F90:513-584.  S and GUT are the 2x2 leading blocks.
static vreal omega12(const vreal S[3][3], const vreal GUT[3][3]) {
  vreal n = -(S[0][0] * GUT[1][0]) + S[0][1] * (GUT[0][0] - GUT[1][1]) + S[1][1] * GUT[0][1];
  vreal d = S[0][0] + S[1][1];
  return n / d;
}

// ----------------------------------------------------------------------------
// a6: strongFormConEQ_SQRT and helpers, Constitutive_Models_SQRT_mod.F90:120-509
// ----------------------------------------------------------------------------
static void strongFormConEQ_SQRT(const Phys& ph, GsPt& G, const vreal I[3][3], vreal UCDout[3][3]) {
  // setCTensorSQRT :379-393
  matmul33(G.S, G.S, G.C);
  for (int k = 0; k < 3; k++) {
    vreal dS[3][3], A[3][3], B[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) dS[i][j] = G.dSdXi[i][j][k];
    matmul33(dS, G.S, A);
    matmul33(G.S, dS, B);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) G.dCdXi[i][j][k] = A[i][j] + B[i][j];
  }
  // conEQ_p
  if (ph.model == 0) {
    // conEQ_L_PTT_SQRT :153-180 (N = Problem_Main%Ndim = 2)
    vreal trC = trace33(G.C);
    G.f_T = vreal(1.0) + ph.VeN * (trC - vreal(2.0));
    G.f[0] = G.f_T;
    G.f[1] = G.f[0];
    G.f[2] = Z();
  } else {
    // conEQ_Giesekus_SQRT :285-305
    vreal aG = ph.VeN;
    G.f_T = vreal(1.0) - vreal(2.0) * aG;
    G.f[0] = vreal(1.0) - aG;
    G.f[1] = vreal(1.0) - vreal(2.0) * aG;
    G.f[2] = aG;
  }
  G.g[0] = vreal(1.0);
  G.g[1] = vreal(1.0);
  for (int d = 0; d < 3; d++) {
    G.dgdxi[0][d] = Z();
    G.dgdxi[1][d] = Z();
  }
  // setStressTensorSQRT :483-509
  vreal a = (vreal(1.0) - ph.BvN) / G.WiN;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) G.T_ve[i][j] = G.g[1] * G.C[i][j] - G.g[0] * I[i][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) G.T_ve[i][j] = a * G.T_ve[i][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      G.P_tot[i][j] = G.T_ve[i][j] + ph.BvN * G.G_dot_tr[i][j] - G.P * I[i][j];
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        G.dTdXi[i][j][k] = G.dgdxi[1][k] * G.C[i][j] + G.g[1] * G.dCdXi[i][j][k] - G.dgdxi[0][k] * I[i][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) G.dTdXi[i][j][k] = a * G.dTdXi[i][j][k];

  // setUpperConvectedDerivativeSQRT :397-416 (dSdt = 0 off the transient path)
  vreal UCD[3][3];
  zero33(UCD);
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) UCD[i][j] = UCD[i][j] + G.U_f[k] * G.dSdXi[i][j][k];
  // setOmegaTensor :420-455 with the traceless transposed gradient
  vreal W[3][3];
  zero33(W);
  W[0][1] = omega12(G.S, G.GUT_tr);
  W[1][0] = -W[0][1];
  vreal M1[3][3], M2[3][3];
  matmul33(G.S, G.GU_tr, M1);
  matmul33(W, G.S, M2);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) UCD[i][j] = UCD[i][j] - M1[i][j] - M2[i][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) UCDout[i][j] = UCD[i][j];

  // setSInverse :459-479 -> inverseTensor2D (Tools_mod.F90:129-146)
  vreal SI[3][3];
  zero33(SI);
  {
    vreal detA = G.S[0][0] * G.S[1][1] - G.S[0][1] * G.S[1][0];
    SI[0][0] = G.S[1][1] / detA;
    SI[0][1] = -G.S[0][1] / detA;
    SI[1][0] = -G.S[1][0] / detA;
    SI[1][1] = G.S[0][0] / detA;
  }
  // S3 = matmul(S, matmul(S,S)); the inner product has the operands of C.
  // Multiplied by f(3), which is zero for L-PTT.
  vreal S3[3][3];
  if (ph.model == 0) zero33(S3);
  else matmul33(G.S, G.C, S3);
  vreal term_S[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      term_S[i][j] = G.f[2] * S3[i][j] + G.f[1] * G.S[i][j] - G.f[0] * SI[i][j];
  vreal two = vreal(2.0);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) G.Con_EQ[i][j] = G.WiN * UCD[i][j] + term_S[i][j] / two;
}

// ----------------------------------------------------------------------------
// a5: strongFormMomentum :658-683, strongFormContinuity :687-697
// ----------------------------------------------------------------------------
static void strongFormMomentum(const Phys& ph, GsPt& G) {
  vreal div_Tve[3];
  for (int i = 0; i < 3; i++) {
    div_Tve[i] = Z();
    for (int j = 0; j < 3; j++) div_Tve[i] = div_Tve[i] + G.dTdXi[j][i][j];
  }
  vreal div_Ptot[3];
  for (int i = 0; i < 3; i++)
    div_Ptot[i] = -G.dPdXi[i] + div_Tve[i] + ph.BvN * (G.G2U[0][i] + G.G2U[1][i] + G.G2U[2][i]);
  // Convection = dUfdt (zero) + matmul(U_f, GU)
  for (int j = 0; j < 3; j++) {
    vreal m = Z();
    for (int k = 0; k < 3; k++) m = m + G.U_f[k] * G.GU[k][j];
    G.Convection[j] = Z() + m;
  }
  for (int i = 0; i < 3; i++) G.Momentum[i] = ph.ReN * G.Convection[i] - div_Ptot[i];
}

// ----------------------------------------------------------------------------
// a7: setStabilizationTerms, Element_Calculations_mod.F90:701-773
// ----------------------------------------------------------------------------
// Planted mistakes in the tau coefficients, for the tests that must catch
// them (0: none; vo_tau_plant lists them).
static int g_tau_plant = 0;

// Helem**2 of setStabilizationTerms (Element_Calculations_mod.F90:718-727):
// geometry only.
static vreal elementH2(const Elem& E) {
  vreal Helem = Z();
  for (int n = 0; n < 9; n++)
    Helem = Helem + E.dFdXi[0][n] * E.dFdXi[0][n] + E.dFdXi[1][n] * E.dFdXi[1][n] +
            E.dFdXi[2][n] * E.dFdXi[2][n];
  Helem = vreal(1.0) / (g_tau_plant == 1 ? Helem : vsqrt(Helem));
  return Helem * Helem;
}

static void setStabilizationTerms(const Phys& ph, const Elem& E, GsPt& G) {
  vreal tsugn = Z();
  for (int n = 0; n < 9; n++) {
    vreal dF[3] = {E.dFdXi[0][n], E.dFdXi[1][n], E.dFdXi[2][n]};
    vreal un = dot3(G.U_f, dF);
    tsugn = tsugn + (g_tau_plant == 2 ? un : fabsr(un));
  }
  vreal H2 = elementH2(E);
  G.H2 = H2;
  G.tsugn = tsugn;

  vreal ddT = ddot33(G.T_ve, G.T_ve);
  vreal ddG = ddot33(G.G_dot_tr, G.G_dot_tr);
  vreal x = (vreal(1.0) - ph.BvN) / G.WiN;
  vreal Ha = vsqrt(x * x + (g_tau_plant == 4 ? ddT : vreal(0.5) * ddT));
  if (g_tau_plant != 5) Ha = cdiv(Ha, vsqrt(vreal(1.0) + vreal(0.5) * ddG));

  vreal r1 = (g_tau_plant == 13 ? vreal(1.0) : ph.ReN) / ph.dt;
  vreal r2 = ph.ReN * tsugn;
  vreal r3 = ph.BvN / (g_tau_plant == 3 ? vsqrt(H2) : H2);
  vreal r4 = cdiv(Ha, H2);
  G.tlsme = r1 * r1 + r2 * r2 + r3 * r3 + r4 * r4;
  G.tlsme = cdiv(vreal(1.0), vsqrt(G.tlsme));
  G.tlsic = g_tau_plant == 6 ? H2 * G.tlsme : cdiv(H2, G.tlsme);

  // tdcs = (Helem**2)/sqrt(0.5*S:S), zeroed when DfN > 1e-12 (:746-760).  Work
  // whose result the reference overwrites or multiplies by a zero coefficient
  // is skipped; the values that reach the residual are unchanged.
  if (ph.dcs_off) {
    G.tdcs = Z();
  } else {
    vreal ddS = ddot33(G.S, G.S);
    G.tdcs = vreal(0.5) * ddS;
    G.tdcs = cdiv(H2, g_tau_plant == 11 ? G.tdcs : vsqrt(G.tdcs));
  }
  vreal ddC = (ph.model == 0)      ? Z()
              : g_tau_plant == 10 ? ddot33(G.T_ve, G.T_ve)
                                  : ddot33(G.C, G.C);  // multiplied by f(3) = 0 for L-PTT

  // mag_con_EQ = sqrt(0.5*Con_EQ:Con_EQ); complex sqrt of a real argument.  Only
  // used multiplied by tdcs.
  G.mag_re = Z();
  G.mag_im = Z();
  if (!ph.dcs_off) {
    vreal mc = vreal(0.5) * ddot33(G.Con_EQ, G.Con_EQ);
    if (val(mc) >= 0.0) G.mag_re = vsqrt(mc);
    else G.mag_im = vsqrt(-mc);
  }

  vreal t1 = (g_tau_plant == 14 ? vreal(1.0) : G.WiN) / ph.dt;
  vreal t2 = g_tau_plant == 7 ? tsugn : G.WiN * tsugn;
  vreal t3 = G.WiN * vsqrt(g_tau_plant == 8 ? ddG : vreal(0.5) * ddG);
  vreal t4 = g_tau_plant == 9 ? G.f[0] : G.f[1];
  vreal t5 = G.f[2] * vsqrt(vreal(0.5) * ddC);
  // ((tdcs*mag)/H2)**2 with mag complex: exactly one of mag_re, mag_im is nonzero,
  // so the square is real: +re^2 or -im^2.
  vreal z6r = cdiv(G.tdcs * G.mag_re, H2);
  vreal z6i = cdiv(G.tdcs * G.mag_im, H2);
  vreal t6 = z6r * z6r - z6i * z6i;
  vreal t7 = ph.DfN / (g_tau_plant == 12 ? vsqrt(H2) : H2);
  G.tlsce = t1 * t1 + t2 * t2 + t3 * t3 + t4 * t4 + t5 * t5 + t6 + t7 * t7;
  G.tlsce = cdiv(vreal(1.0), vsqrt(G.tlsce));
}

// setFlowQuantities :247-262 (+ setContinuationParameter :577-601)
static void setFlowQuantities(const Phys& ph, const Elem& E, GsPt& G, const vreal I[3][3]) {
  setFlowUnknowns(E, G);
  G.WiN = ph.WiN0 * E.EU;  // WiN = WiN0*EU(Nex)
  setFlowKinematics(G, I);
  vreal UCD[3][3];
  strongFormConEQ_SQRT(ph, G, I, UCD);
  strongFormMomentum(ph, G);
  G.Continuity = trace33(G.GU);
  setStabilizationTerms(ph, E, G);
}

static void identity(vreal I[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) I[i][j] = (i == j) ? vreal(1.0) : Z();
}

// ----------------------------------------------------------------------------
// a8: test-node residual terms, Residuals_mod.F90 + equationsResidualsBulkMain
// (Equations_mod.F90:102-144)
// ----------------------------------------------------------------------------
// The summands of TERM_RES for one test node at one Gauss point, in the order
// the reference adds them.
struct NodeTerms {
  vreal mom[2][5];  // momentumMainSQRT :331-335: ReN*Convection(i)*bifn, P_tot(i,:).dFdXi, LSME, LSIC, LSCE
  vreal cont[2];    // continuityBulkMain :90-110: Continuity*bifn, LSMEContinuity
  vreal str[3][5];  // stressesBulkMainSQRT :375-380: Con_EQ(j,k)*bifn, DfN*dSdXi(j,k,:).dFdXi, LSME, LSCE, DCS
};

// Planted mistakes in the test-node terms, for the tests that must catch them
// (0: none; vo_term_plant lists them).
static int g_term_plant = 0;

static void nodeTerms(const Phys& ph, const Elem& E, const GsPt& G, int n, NodeTerms& t) {
  const int P = g_term_plant;
  vreal bifn = vreal(E.bfn[n]);
  vreal dF[3] = {E.dFdXi[0][n], E.dFdXi[1][n], E.dFdXi[2][n]};
  vreal one = vreal(1.0);
  // Values the reference recomputes with identical operands for each i or each
  // stress component are computed once here (bitwise the same).
  vreal omb = P == 8 ? one : one - ph.BvN;
  vreal Udf = dot3(G.U_f, dF);
  vreal lsme_m = G.tlsme * (P == 3 ? one : ph.ReN) * Udf;  // LSMEMomentum :161-173
  // momentumMainSQRT :295-337
  for (int i = 0; i < 2; i++) {
    vreal lsme = lsme_m * G.Momentum[P == 4 ? 1 - i : i];
    vreal lsic = G.tlsic * dF[P == 5 ? 1 - i : i];  // LSICMomentum :177-187
    lsic = lsic * G.Continuity;
    // LSCEMomentumSQRT :388-433
    vreal GUT_l[3][3], GU_l[3][3];
    zero33(GUT_l);
    for (int k = 0; k < 3; k++) GUT_l[i][k] = dF[k];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) GU_l[a][b] = P == 7 ? GUT_l[a][b] : GUT_l[b][a];
    vreal W_l[3][3];
    zero33(W_l);
    W_l[0][1] = omega12(G.S, GUT_l);
    W_l[1][0] = -W_l[0][1];
    vreal M1[3][3], M2[3][3], lsce[3][3];
    matmul33(G.S, GU_l, M1);
    matmul33(W_l, G.S, M2);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        vreal U = -M1[a][b] - M2[a][b];
        lsce[a][b] = P == 6 ? G.tlsce * U * omb * G.WiN : cdiv(G.tlsce * U * omb, G.WiN);
      }
    vreal Prow[3] = {G.P_tot[i][0], G.P_tot[i][1], G.P_tot[i][2]};
    t.mom[i][0] = (P == 1 ? one : ph.ReN) * G.Convection[i] * bifn;
    t.mom[i][1] = P == 2 ? -dot3(Prow, dF) : dot3(Prow, dF);
    t.mom[i][2] = lsme;
    t.mom[i][3] = lsic;
    t.mom[i][4] = ddot33(lsce, G.Con_EQ);
  }
  // continuityBulkMain :90-110 with LSMEContinuity :223-234
  t.cont[0] = (P == 9 ? -G.Continuity : G.Continuity) * bifn;
  t.cont[1] = G.tlsme * dot3(dF, G.Momentum);
  if (P == 10) t.cont[1] = -t.cont[1];
  // stressesBulkMainSQRT :341-384
  // LSCEStressSQRT :459-475 (independent of the stress component)
  vreal lsce_s = (P == 16 ? one : G.WiN) * Udf;
  lsce_s = lsce_s + (P == 15 ? G.f_T * bifn : cdiv(G.f_T * bifn, vreal(2.0)));
  lsce_s = G.tlsce * lsce_s;
  const int comp[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  for (int s = 0; s < 3; s++) {
    int j = comp[s][0], k = comp[s][1];
    // LSMEStressSQRT :437-455
    vreal lsme = P == 18 ? dF[k] * G.Momentum[j] : -(dF[k] * G.Momentum[j]);
    if (j != k && P != 13) lsme = lsme - dF[j] * G.Momentum[k];
    lsme = P == 14 ? cdiv(G.tlsme * lsme, G.WiN) : G.tlsme * lsme * G.WiN;
    vreal lsce = lsce_s * G.Con_EQ[j][k];
    // DCSStressSQRT :479-490 times mag_con_EQ (real part of a complex product)
    vreal dS[3] = {G.dSdXi[j][k][0], G.dSdXi[j][k][1], G.dSdXi[j][k][2]};
    vreal dcs = G.tdcs * dot3(dS, dF);
    if (P != 17) dcs = dcs * G.mag_re;
    t.str[s][0] = G.Con_EQ[j][P == 11 ? j : k] * bifn;
    t.str[s][1] = (P == 12 ? -ph.DfN : ph.DfN) * dot3(dS, dF);
    t.str[s][2] = lsme;
    t.str[s][3] = lsce;
    t.str[s][4] = dcs;
  }
}

static void nodeResidual(const Phys& ph, const Elem& E, const GsPt& G, int n, vreal TERM[6]) {
  NodeTerms t;
  nodeTerms(ph, E, G, n, t);
  for (int i = 0; i < 2; i++) TERM[i] = t.mom[i][0] + t.mom[i][1] + t.mom[i][2] + t.mom[i][3] + t.mom[i][4];
  TERM[2] = t.cont[0] + t.cont[1];
  for (int s = 0; s < 3; s++)
    TERM[3 + s] = t.str[s][0] + t.str[s][1] + t.str[s][2] + t.str[s][3] + t.str[s][4];
}

// residualBulk, Equations_mod.F90:59-98
static void residualBulk(const Phys& ph, Elem& E, vreal Res[9][6]) {
  const Tables& T = tables();
  vreal I[3][3];
  identity(I);
  for (int n = 0; n < 9; n++)
    for (int q = 0; q < 6; q++) Res[n][q] = Z();
  for (int g = 0; g < 9; g++) {
    E.bfn = T.bfn[g];
    E.dfdc = T.dfdc[g];
    E.dfde = T.dfde[g];
    E.d2fdc2 = T.d2fdc2[g];
    E.d2fde2 = T.d2fde2[g];
    E.d2fdce = T.d2fdce[g];
    GsPt G;
    basis2D(E, G);
    setFlowQuantities(ph, E, G, I);
    G.WET = vreal(T.W2[g]) * fabsr(G.detJ);
    for (int n = 0; n < 9; n++) {
      vreal TERM[6];
      nodeResidual(ph, E, G, n, TERM);
      for (int q = 0; q < 6; q++) Res[n][q] = Res[n][q] + TERM[q] * G.WET;
    }
  }
}

// ----------------------------------------------------------------------------
// a12: edge path, residualEdge (Equations_mod.F90:387-424), setEdgeQuantities
// (Element_Calculations_mod.F90:852-924), openBC (Residuals_mod.F90:599-624)
// ----------------------------------------------------------------------------
static void residualEdge(const Phys& ph, Elem& E, int ied /*1..4*/, vreal Res[9][6]) {
  const Tables& T = tables();
  vreal I[3][3];
  identity(I);
  for (int n = 0; n < 9; n++)
    for (int q = 0; q < 6; q++) Res[n][q] = Z();
  int ed = ied - 1;
  for (int g = 0; g < 3; g++) {
    E.bfn = T.bfnE[ed][g];
    E.dfdc = T.dfdcE[ed][g];
    E.dfde = T.dfdeE[ed][g];
    E.d2fdc2 = T.d2fdc2E[ed][g];
    E.d2fde2 = T.d2fde2E[ed][g];
    E.d2fdce = T.d2fdceE[ed][g];
    GsPt G;
    basis2D(E, G);
    setFlowQuantities(ph, E, G, I);
    vreal dxdc = G.dXidci[0][0], dydc = G.dXidci[0][1];
    vreal dxde = G.dXidci[1][0], dyde = G.dXidci[1][1];
    G.normal[2] = Z();
    switch (ied) {
      case 1: G.normal[0] = dydc; G.normal[1] = -dxdc; break;
      case 2: G.normal[0] = dyde; G.normal[1] = -dxde; break;
      case 3: G.normal[0] = -dydc; G.normal[1] = dxdc; break;
      default: G.normal[0] = -dyde; G.normal[1] = dxde; break;
    }
    G.dL = vsqrt(G.normal[0] * G.normal[0] + G.normal[1] * G.normal[1]);
    G.normal[0] = G.normal[0] / G.dL;
    G.normal[1] = G.normal[1] / G.dL;
    G.WET = vreal(T.W1[g]) * G.dL;
    for (int n = 0; n < 9; n++) {
      vreal TERM[6];
      for (int q = 0; q < 6; q++) TERM[q] = Z();
      vreal bifn = vreal(E.bfn[n]);
      for (int i = 0; i < 2; i++) {
        vreal v[3];
        for (int k = 0; k < 3; k++) v[k] = G.P_tot[k][i] + G.P * I[k][i];
        TERM[i] = -(dot3(G.normal, v) * bifn);
      }
      for (int q = 0; q < 6; q++) Res[n][q] = Res[n][q] + TERM[q] * G.WET;
    }
  }
}

// ----------------------------------------------------------------------------
// a10: storeBulk / storeEdge FD Jacobian, Equations_mod.F90:226-313, 465-554;
// perturbVariable (Tools_mod.F90:831-840): eps_jac = 1e-8.
// A[(inod,ieq)][(jnod,jeq)] = TEMP_A_f(inod,ieq,jnod,jeq) = dR(jnod,jeq)/du(inod,ieq);
// Ac[(jnod,jeq)] = TEMP_Ac_f(jnod,jeq,1).  The perturbation is applied and
// removed in place, so entries perturbed earlier keep fl(fl(u+h)-h).
// ----------------------------------------------------------------------------
static void elementFD(const Phys& ph, Elem& E, int ied, bool jac, vreal Res[9][6], vreal* A, vreal* Ac) {
  auto resid = [&](vreal R[9][6]) {
    if (ied == 0) residualBulk(ph, E, R);
    else residualEdge(ph, E, ied, R);
  };
  resid(Res);
  if (!jac) return;
  const vreal h = vreal(1.0e-8);
  vreal dRes[9][6];
  for (int inod = 0; inod < 9; inod++)
    for (int ieq = 0; ieq < 6; ieq++) {
      E.TL[inod][ieq] = E.TL[inod][ieq] + h;
      resid(dRes);
      E.TL[inod][ieq] = E.TL[inod][ieq] - h;
      int col = inod * 6 + ieq;
      for (int jnod = 0; jnod < 9; jnod++)
        for (int jeq = 0; jeq < 6; jeq++)
          A[col * 54 + jnod * 6 + jeq] = cdiv(dRes[jnod][jeq] - Res[jnod][jeq], h);
    }
  E.EU = E.EU + h;
  resid(dRes);
  E.EU = E.EU - h;
  for (int jnod = 0; jnod < 9; jnod++)
    for (int jeq = 0; jeq < 6; jeq++) Ac[jnod * 6 + jeq] = cdiv(dRes[jnod][jeq] - Res[jnod][jeq], h);
}

// copyToLocal, Element_Calculations_mod.F90:106-142 (global node numbering)
static void copyToLocal(Elem& E, const double* coords, const int32_t* conn, int e, const double* TL,
                        double EU) {
  for (int n = 0; n < 9; n++) {
    int gn = conn[e * 9 + n];
    for (int q = 0; q < 6; q++) E.TL[n][q] = vreal(TL[(size_t)gn * 6 + q]);
    E.X[n][0] = coords[(size_t)gn * 2 + 0];
    E.X[n][1] = coords[(size_t)gn * 2 + 1];
    E.X[n][2] = 0.0;
  }
  E.EU = vreal(EU);
}

// ----------------------------------------------------------------------------
// a11/a13: matrix pattern (preAllocatePetsc, Petsc_mod.F90:629-743) and the
// ADD_VALUES scatter (Storage_mod.F90:10-115), rows/cols 0-based
// (gnod*6 + ieq), the arclength unknown last (N-1).
// ----------------------------------------------------------------------------
struct CSR {
  int64_t n;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> col;
};

static void build_pattern(int n_nodes, int n_el, const int32_t* conn, CSR& P) {
  int64_t N = (int64_t)n_nodes * 6 + 1;
  P.n = N;
  std::vector<std::vector<int32_t>> nb(n_nodes);
  for (int e = 0; e < n_el; e++)
    for (int a = 0; a < 9; a++)
      for (int b = 0; b < 9; b++) nb[conn[e * 9 + a]].push_back(conn[e * 9 + b]);
  for (auto& v : nb) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  P.rowptr.assign(N + 1, 0);
  for (int r = 0; r < n_nodes; r++)
    for (int q = 0; q < 6; q++) P.rowptr[(int64_t)r * 6 + q + 1] = (int64_t)nb[r].size() * 6 + 1;
  P.rowptr[N] = N;  // dense arclength row
  for (int64_t i = 0; i < N; i++) P.rowptr[i + 1] += P.rowptr[i];
  P.col.resize(P.rowptr[N]);
  for (int r = 0; r < n_nodes; r++)
    for (int q = 0; q < 6; q++) {
      int64_t p = P.rowptr[(int64_t)r * 6 + q];
      for (int32_t c : nb[r])
        for (int iq = 0; iq < 6; iq++) P.col[p++] = c * 6 + iq;
      P.col[p++] = (int32_t)(N - 1);
    }
  int64_t p = P.rowptr[N - 1];
  for (int64_t c = 0; c < N; c++) P.col[p++] = (int32_t)c;
}

static inline int64_t find_slot(const int64_t* rowptr, const int32_t* col, int64_t row, int32_t c) {
  const int32_t* b = col + rowptr[row];
  const int32_t* en = col + rowptr[row + 1];
  const int32_t* it = std::lower_bound(b, en, c);
  if (it == en || *it != c) return -1;
  return it - col;
}

// Adds one element's local residual / Jacobian in PETSc's ADD_VALUES order.
static int scatter_element(const int32_t* nodes, const double* Res54, const double* A, const double* Ac,
                           int64_t N, const int64_t* rowptr, const int32_t* col, double* b, double* vals) {
  for (int n = 0; n < 9; n++)
    for (int q = 0; q < 6; q++) b[(int64_t)nodes[n] * 6 + q] += Res54[n * 6 + q];
  if (!A) return 0;
  for (int jn = 0; jn < 9; jn++)
    for (int jq = 0; jq < 6; jq++) {
      int64_t row = (int64_t)nodes[jn] * 6 + jq;
      for (int in = 0; in < 9; in++)
        for (int iq = 0; iq < 6; iq++) {
          int64_t s = find_slot(rowptr, col, row, nodes[in] * 6 + iq);
          if (s < 0) return -1;
          vals[s] += A[(in * 6 + iq) * 54 + jn * 6 + jq];
        }
    }
  for (int jn = 0; jn < 9; jn++)
    for (int jq = 0; jq < 6; jq++) {
      int64_t row = (int64_t)nodes[jn] * 6 + jq;
      int64_t s = find_slot(rowptr, col, row, (int32_t)(N - 1));
      if (s < 0) return -1;
      vals[s] += Ac[jn * 6 + jq];
    }
  return 0;
}

// ----------------------------------------------------------------------------
// f3: linear stability assembly of Cylinder_Stability_2D (Neq = 6, Nex = 0)
// with Stress_Reform = 'SQRT'.
//   assembleSlepcObjects      src/Stability/Stability_mod.F90:97-199
//   setEquationsBulkStability src/Stability/Equations_Stability_mod.F90:20-57
//   residualBulkStability     :61-96, equationsResidualsBulkStability :98-131
//   storeBulkStability        :135-226
//   setFlowQuantitiesStability, setFlowUnknownsStability,
//   strongFormMomentumStability src/Stability/Element_Calculations_Stability_mod.F90:19-102,186-262,302-329
//   strongFormConEQ_Stability_SQRT and helpers
//                             src/Stability/Constitutive_Models_SQRT_Stability_mod.F90:87-487
//   momentumStabilitySQRT, continuityStability, stressStabilitySQRT
//                             src/Stability/Residuals_Stability_mod.F90:125-153,209-313
// The base state b is TEMP_TL_b = TL (real), the disturbance d is TEMP_TL_d
// (complex; zero in storeBulkStability's base evaluation, Stability_mod.F90:54-58).
// Every disturbance value has a zero imaginary part on this path (the
// spanwise terms are commented out, Element_Calculations_Stability_mod.F90:51-81,
// and eps_jac is the real part of perturbVariable's 1e-8+1e-8i, assigned to a
// PetscReal, Equations_Stability_mod.F90:155,179), so the real restatement
// with the division rule cdiv() reproduces the reference's values.
// A (Res_a) and B (Res_b, negated) are forward differences in the
// disturbance.  The discontinuity-capturing terms multiply tdcs = 0 (DfN >
// 1e-12 is required here), so they add signed zeros only and are skipped.
// ----------------------------------------------------------------------------
// Planted mistakes for the tests that show the linearisation check catches
// them (0: none; 1: sign of n_b*d_d in W_d; 2: d.U_f.grad(S_b) left out of UCD_d).
static int g_stab_plant = 0;

struct GsPtD {
  vreal U_f[3], GU[3][3], GUT[3][3], G_dot[3][3], G2U[3][3], dPdXi[3], P;
  vreal S[3][3], dSdXi[3][3][3];
  vreal C[3][3], dCdXi[3][3][3], f[3], g[2], dgdxi[2][3], f_T;
  vreal T_ve[3][3], P_tot[3][3], dTdXi[3][3][3];
  vreal Con_EQ[3][3], Con_EQ_time[3][3];
  vreal Convection[3], Momentum[3], Momentum_time[3], Continuity;
};

// setFlowUnknownsStabilityComplex (Element_Calculations_Stability_mod.F90:186-262)
// for the disturbance nodal values TLd[9][6].
static void setFlowUnknownsD(const Elem& E, const vreal TLd[9][6], GsPtD& d) {
  auto dot9 = [&](int ieq, const vreal* v) {
    vreal s = Z();
    for (int n = 0; n < 9; n++) s = s + TLd[n][ieq] * v[n];
    return s;
  };
  auto dot9d = [&](int ieq, const double* v) {
    vreal s = Z();
    for (int n = 0; n < 9; n++) s = s + TLd[n][ieq] * vreal(v[n]);
    return s;
  };
  for (int i = 0; i < 3; i++) d.U_f[i] = Z(), d.dPdXi[i] = Z();
  zero33(d.GU);
  zero33(d.G2U);
  for (int i = 0; i < 2; i++) {
    d.U_f[i] = dot9d(i, E.bfn);
    for (int j = 0; j < 3; j++) {
      d.GU[j][i] = dot9(i, E.dFdXi[j]);
      d.G2U[j][i] = dot9(i, E.d2FdXi2[j]);
    }
    d.dPdXi[i] = dot9(2, E.dFdXi[i]);
  }
  d.P = dot9d(2, E.bfn);
  zero33(d.S);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      for (int c = 0; c < 3; c++) d.dSdXi[a][b][c] = Z();
  const int comp[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  for (int s = 0; s < 3; s++) {
    int j = comp[s][0], k = comp[s][1];
    d.S[j][k] = dot9d(3 + s, E.bfn);
    d.S[k][j] = d.S[j][k];
    for (int c = 0; c < 3; c++) d.dSdXi[j][k][c] = dot9(3 + s, E.dFdXi[c]);
    for (int c = 0; c < 3; c++) d.dSdXi[k][j][c] = d.dSdXi[j][k][c];
  }
}

// setFlowKinematics (Element_Calculations_mod.F90:605-634) for the disturbance:
// GUT and G_dot are what the stability terms read.
static void setFlowKinematicsD(GsPtD& d) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) d.GUT[i][j] = d.GU[j][i];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) d.G_dot[i][j] = d.GU[i][j] + d.GUT[i][j];
}

// omegaNominator2D (Constitutive_Models_SQRT_mod.F90:540-584) on the leading
// 2x2 blocks of S and GUT.
static vreal omegaNom(const vreal S[3][3], const vreal GUT[3][3]) {
  return -(S[0][0] * GUT[1][0]) + S[0][1] * (GUT[0][0] - GUT[1][1]) + S[1][1] * GUT[0][1];
}

// strongFormConEQ_Stability_SQRT (Constitutive_Models_SQRT_Stability_mod.F90:87-125)
// with setCTensorStabilitySQRT :129-147, conEQ_L_PTT_Stability_SQRT :151-175,
// conEQ_Giesekus_Stability_SQRT :276-297, setStressTensorStabilitySQRT :431-459,
// setUCDStabilitySQRT :376-413, WTensor2DStability :463-486,
// setSInverseStability :417-427.
static void strongFormConEQStability(const Phys& ph, const GsPt& b, GsPtD& d, const vreal I[3][3]) {
  vreal M1[3][3], M2[3][3], M3[3][3], M4[3][3];
  matmul33(b.S, d.S, M1);
  matmul33(d.S, b.S, M2);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) d.C[i][j] = M1[i][j] + M2[i][j];
  for (int k = 0; k < 3; k++) {
    vreal bdS[3][3], ddS[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) bdS[i][j] = b.dSdXi[i][j][k], ddS[i][j] = d.dSdXi[i][j][k];
    matmul33(bdS, d.S, M1);
    matmul33(b.S, ddS, M2);
    matmul33(ddS, b.S, M3);
    matmul33(d.S, bdS, M4);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) d.dCdXi[i][j][k] = M1[i][j] + M2[i][j] + M3[i][j] + M4[i][j];
  }
  if (ph.model == 0) {
    vreal trC_d = trace33(d.C);
    d.f_T = ph.VeN * trC_d;
    d.f[0] = d.f_T;
    d.f[1] = d.f[0];
    d.f[2] = Z();
  } else {
    d.f_T = Z();
    d.f[0] = Z();
    d.f[1] = Z();
    d.f[2] = Z();
  }
  d.g[0] = Z();
  d.g[1] = Z();
  for (int c = 0; c < 3; c++) d.dgdxi[0][c] = Z(), d.dgdxi[1][c] = Z();

  const vreal omb = vreal(1.0) - ph.BvN;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) d.T_ve[i][j] = b.g[1] * d.C[i][j] + d.g[1] * b.C[i][j] - d.g[0] * I[i][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) d.T_ve[i][j] = cdiv(omb * d.T_ve[i][j], b.WiN);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) d.P_tot[i][j] = d.T_ve[i][j] + ph.BvN * d.G_dot[i][j] - d.P * I[i][j];
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        d.dTdXi[i][j][k] = b.dgdxi[1][k] * d.C[i][j] + d.dgdxi[1][k] * b.C[i][j] + b.g[1] * d.dCdXi[i][j][k] +
                           d.g[1] * b.dCdXi[i][j][k] - d.dgdxi[0][k] * I[i][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) d.dTdXi[i][j][k] = cdiv(omb * d.dTdXi[i][j][k], b.WiN);

  // UCD_d
  vreal UCD[3][3];
  matmul33(b.S, d.GU, M1);
  matmul33(d.S, b.GU, M2);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) UCD[i][j] = -M1[i][j] - M2[i][j];
  for (int k = 0; k < 3; k++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        UCD[i][j] = g_stab_plant == 2 ? UCD[i][j] + b.U_f[k] * d.dSdXi[i][j][k]
                                      : UCD[i][j] + b.U_f[k] * d.dSdXi[i][j][k] + d.U_f[k] * b.dSdXi[i][j][k];
  vreal W_b[3][3], W_d[3][3];
  zero33(W_b);
  W_b[0][1] = omega12(b.S, b.GUT_tr);  // setOmegaTensor(GsPt_b)
  W_b[1][0] = -W_b[0][1];
  {
    vreal n_b = omegaNom(b.S, b.GUT);
    vreal d_b = b.S[0][0] + b.S[1][1];
    vreal n_d = omegaNom(b.S, d.GUT) + omegaNom(d.S, b.GUT);
    vreal d_d = d.S[0][0] + d.S[1][1];
    zero33(W_d);
    W_d[0][1] = cdiv(g_stab_plant == 1 ? n_d * d_b + n_b * d_d : n_d * d_b - n_b * d_d, d_b * d_b);
    W_d[1][0] = -W_d[0][1];
  }
  matmul33(W_b, d.S, M1);
  matmul33(W_d, b.S, M2);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) UCD[i][j] = UCD[i][j] - M1[i][j] - M2[i][j];

  // SI_b = setSInverse(real(S_b)) (inverseTensor2D, real divisions); SI_d
  vreal SI_b[3][3], SI_d[3][3];
  zero33(SI_b);
  {
    vreal detA = b.S[0][0] * b.S[1][1] - b.S[0][1] * b.S[1][0];
    SI_b[0][0] = b.S[1][1] / detA;
    SI_b[0][1] = -b.S[0][1] / detA;
    SI_b[1][0] = -b.S[1][0] / detA;
    SI_b[1][1] = b.S[0][0] / detA;
  }
  matmul33(d.S, SI_b, M1);
  matmul33(SI_b, M1, M2);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) SI_d[i][j] = -M2[i][j];
  // S3_b = S.(S.S); S3_d = (S.S).S_d + (S.S_d).S + S_d.(S.S)
  vreal SS[3][3], S3_b[3][3], S3_d[3][3];
  matmul33(b.S, b.S, SS);
  matmul33(b.S, SS, S3_b);
  matmul33(SS, d.S, M1);
  matmul33(b.S, d.S, M3);
  matmul33(M3, b.S, M2);
  matmul33(d.S, SS, M4);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) S3_d[i][j] = M1[i][j] + M2[i][j] + M4[i][j];
  const vreal two = vreal(2.0);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      vreal tS = b.f[2] * S3_d[i][j] + d.f[2] * S3_b[i][j] + d.f[1] * b.S[i][j] + b.f[1] * d.S[i][j] -
                 b.f[0] * SI_d[i][j] - d.f[0] * SI_b[i][j];
      d.Con_EQ[i][j] = b.WiN * UCD[i][j] + cdiv(tS, two);
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) d.Con_EQ_time[i][j] = b.WiN * d.S[i][j];
}

// strongFormMomentumStability (Element_Calculations_Stability_mod.F90:302-329)
// and strongFormContinuity (Element_Calculations_mod.F90:687-697).
static void strongFormMomentumStability(const Phys& ph, const GsPt& b, GsPtD& d) {
  vreal div_Tve[3], div_Ptot[3];
  for (int i = 0; i < 3; i++) {
    div_Tve[i] = Z();
    for (int j = 0; j < 3; j++) div_Tve[i] = div_Tve[i] + d.dTdXi[j][i][j];
  }
  for (int i = 0; i < 3; i++)
    div_Ptot[i] = -d.dPdXi[i] + div_Tve[i] + ph.BvN * (d.G2U[0][i] + d.G2U[1][i] + d.G2U[2][i]);
  for (int j = 0; j < 3; j++) {
    vreal m1 = Z(), m2 = Z();
    for (int k = 0; k < 3; k++) m1 = m1 + b.U_f[k] * d.GU[k][j];
    for (int k = 0; k < 3; k++) m2 = m2 + d.U_f[k] * b.GU[k][j];
    d.Convection[j] = m1 + m2;
  }
  for (int i = 0; i < 3; i++) d.Momentum[i] = ph.ReN * d.Convection[i] - div_Ptot[i];
  for (int i = 0; i < 3; i++) d.Momentum_time[i] = ph.ReN * d.U_f[i];
  d.Continuity = trace33(d.GU);
}

// setFlowQuantitiesStability's disturbance part for a given base Gauss point.
static void setFlowQuantitiesD(const Phys& ph, const Elem& E, const vreal TLd[9][6], const GsPt& b, GsPtD& d,
                               const vreal I[3][3]) {
  setFlowUnknownsD(E, TLd, d);
  setFlowKinematicsD(d);
  strongFormConEQStability(ph, b, d, I);
  strongFormMomentumStability(ph, b, d);
}

// LSCEMomentumSQRT (Residuals_mod.F90:388-433) of a Gauss point's quantities.
static void lsceMomentum(const Phys& ph, const GsPt& G, int i, const vreal dF[3], vreal lsce[3][3]) {
  vreal GUT_l[3][3], GU_l[3][3];
  zero33(GUT_l);
  for (int k = 0; k < 3; k++) GUT_l[i][k] = dF[k];
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 3; c++) GU_l[a][c] = GUT_l[c][a];
  vreal W_l[3][3];
  zero33(W_l);
  W_l[0][1] = omega12(G.S, GUT_l);
  W_l[1][0] = -W_l[0][1];
  vreal M1[3][3], M2[3][3];
  matmul33(G.S, GU_l, M1);
  matmul33(W_l, G.S, M2);
  const vreal omb = vreal(1.0) - ph.BvN;
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 3; c++) {
      vreal U = -M1[a][c] - M2[a][c];
      lsce[a][c] = cdiv(G.tlsce * U * omb, G.WiN);
    }
}

// momentumStabilitySQRT, continuityStability, stressStabilitySQRT
// (Residuals_Stability_mod.F90:125-153,209-313) at test node n; TB already
// negated (equationsResidualsBulkStability, Equations_Stability_mod.F90:124).
static void nodeResidualStability(const Phys& ph, const Elem& E, const GsPt& b, const GsPtD& d, int n, vreal TA[6],
                                  vreal TB[6]) {
  vreal bifn = vreal(E.bfn[n]);
  vreal dF[3] = {E.dFdXi[0][n], E.dFdXi[1][n], E.dFdXi[2][n]};
  for (int i = 0; i < 2; i++) {
    vreal lsme_term = b.tlsme * ph.ReN * dot3(b.U_f, dF);  // LSMEMomentum(dFdXi, GsPt_b)
    vreal lsic_term = b.tlsic * dF[i];
    lsic_term = lsic_term * d.Continuity;
    vreal lsce[3][3];
    lsceMomentum(ph, b, i, dF, lsce);
    vreal lsce_term = ddot33(lsce, d.Con_EQ);
    vreal Prow[3] = {d.P_tot[i][0], d.P_tot[i][1], d.P_tot[i][2]};
    TA[i] = ph.ReN * d.Convection[i] * bifn + dot3(Prow, dF) + lsme_term * d.Momentum[i] + lsic_term + lsce_term;
    lsce_term = ddot33(lsce, d.Con_EQ_time);
    TB[i] = d.Momentum_time[i] * bifn + lsme_term * d.Momentum_time[i] + lsce_term;
  }
  TA[2] = d.Continuity * bifn + b.tlsme * dot3(dF, d.Momentum);
  TB[2] = b.tlsme * dot3(dF, d.Momentum_time);
  const int comp[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  vreal lsce_s = b.WiN * dot3(b.U_f, dF);  // LSCEStressSQRT(bifn, dFdXi, GsPt_b)
  lsce_s = lsce_s + cdiv(b.f_T * bifn, vreal(2.0));
  lsce_s = b.tlsce * lsce_s;
  for (int s = 0; s < 3; s++) {
    int j = comp[s][0], k = comp[s][1];
    vreal lsme = -(dF[k] * d.Momentum[j]);  // LSMEStressSQRT(j,k,dFdXi,tlsme_b,Momentum_d)
    if (j != k) lsme = lsme - dF[j] * d.Momentum[k];
    lsme = b.tlsme * lsme * b.WiN;
    vreal dS[3] = {d.dSdXi[j][k][0], d.dSdXi[j][k][1], d.dSdXi[j][k][2]};
    TA[3 + s] = d.Con_EQ[j][k] * bifn + ph.DfN * dot3(dS, dF) + lsme + lsce_s * d.Con_EQ[j][k];
    vreal lsme_t = -(dF[k] * d.Momentum_time[j]);
    if (j != k) lsme_t = lsme_t - dF[j] * d.Momentum_time[k];
    lsme_t = b.tlsme * lsme_t * b.WiN;
    TB[3 + s] = d.Con_EQ_time[j][k] * bifn + lsme_t + lsce_s * d.Con_EQ_time[j][k];
  }
  for (int q = 0; q < 6; q++) TB[q] = -TB[q];
}

static void setGaussPoint(Elem& E, const Tables& T, int g) {
  E.bfn = T.bfn[g];
  E.dfdc = T.dfdc[g];
  E.dfde = T.dfde[g];
  E.d2fdc2 = T.d2fdc2[g];
  E.d2fde2 = T.d2fde2[g];
  E.d2fdce = T.d2fdce[g];
}

// residualBulkStability (Equations_Stability_mod.F90:61-96); E.TL is the base.
static void residualBulkStability(const Phys& ph, Elem& E, const vreal TLd[9][6], vreal ResA[9][6], vreal ResB[9][6]) {
  const Tables& T = tables();
  vreal I[3][3];
  identity(I);
  for (int n = 0; n < 9; n++)
    for (int q = 0; q < 6; q++) ResA[n][q] = Z(), ResB[n][q] = Z();
  for (int g = 0; g < 9; g++) {
    setGaussPoint(E, T, g);
    GsPt b;
    basis2D(E, b);
    setFlowQuantities(ph, E, b, I);
    GsPtD d;
    setFlowQuantitiesD(ph, E, TLd, b, d, I);
    b.WET = vreal(T.W2[g]) * fabsr(b.detJ);
    for (int n = 0; n < 9; n++) {
      vreal TA[6], TB[6];
      nodeResidualStability(ph, E, b, d, n, TA, TB);
      for (int q = 0; q < 6; q++) {
        ResA[n][q] = ResA[n][q] + TA[q] * b.WET;
        ResB[n][q] = ResB[n][q] + TB[q] * b.WET;
      }
    }
  }
}

// storeBulkStability (Equations_Stability_mod.F90:135-226) before the
// scatter: A[(inod,ieq)*54 + (jnod,jeq)] = TEMP_A_f(inod,ieq,jnod,jeq), the
// quotient (dRes_a - Res_a)/eps_jac; B likewise.  TEMP_TL_d starts at 0.
static void elementStability(const Phys& ph, Elem& E, vreal* A, vreal* B) {
  vreal TLd[9][6];
  for (int n = 0; n < 9; n++)
    for (int q = 0; q < 6; q++) TLd[n][q] = Z();
  vreal ResA[9][6], ResB[9][6], dA[9][6], dB[9][6];
  residualBulkStability(ph, E, TLd, ResA, ResB);
  const vreal h = vreal(1.0e-8);
  for (int inod = 0; inod < 9; inod++)
    for (int ieq = 0; ieq < 6; ieq++) {
      TLd[inod][ieq] = TLd[inod][ieq] + h;
      residualBulkStability(ph, E, TLd, dA, dB);
      TLd[inod][ieq] = TLd[inod][ieq] - h;
      int c = inod * 6 + ieq;
      for (int jnod = 0; jnod < 9; jnod++)
        for (int jeq = 0; jeq < 6; jeq++) {
          A[c * 54 + jnod * 6 + jeq] = cdiv(dA[jnod][jeq] - ResA[jnod][jeq], h);
          B[c * 54 + jnod * 6 + jeq] = cdiv(dB[jnod][jeq] - ResB[jnod][jeq], h);
        }
    }
}

// The base residual's test-node terms with the test-function coefficients
// (tau's, LSME/LSCE operators) taken from Gauss point c and the strong-form
// quantities from Gauss point s — the frozen-coefficient form whose
// derivative in the state is the stability operator A.  Used by the tests
// only; with c = s it equals nodeResidual.
static void nodeResidualFrozen(const Phys& ph, const Elem& E, const GsPt& c, const GsPt& s, int n, vreal TERM[6]) {
  vreal bifn = vreal(E.bfn[n]);
  vreal dF[3] = {E.dFdXi[0][n], E.dFdXi[1][n], E.dFdXi[2][n]};
  for (int i = 0; i < 2; i++) {
    vreal lsme = c.tlsme * ph.ReN * dot3(c.U_f, dF) * s.Momentum[i];
    vreal lsic = c.tlsic * dF[i] * s.Continuity;
    vreal lsce[3][3];
    lsceMomentum(ph, c, i, dF, lsce);
    vreal Prow[3] = {s.P_tot[i][0], s.P_tot[i][1], s.P_tot[i][2]};
    TERM[i] = ph.ReN * s.Convection[i] * bifn + dot3(Prow, dF) + lsme + lsic + ddot33(lsce, s.Con_EQ);
  }
  TERM[2] = s.Continuity * bifn + c.tlsme * dot3(dF, s.Momentum);
  vreal lsce_s = c.tlsce * (c.WiN * dot3(c.U_f, dF) + cdiv(c.f_T * bifn, vreal(2.0)));
  const int comp[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  for (int q = 0; q < 3; q++) {
    int j = comp[q][0], k = comp[q][1];
    vreal lsme = -(dF[k] * s.Momentum[j]);
    if (j != k) lsme = lsme - dF[j] * s.Momentum[k];
    lsme = c.tlsme * lsme * c.WiN;
    vreal dS[3] = {s.dSdXi[j][k][0], s.dSdXi[j][k][1], s.dSdXi[j][k][2]};
    TERM[3 + q] = s.Con_EQ[j][k] * bifn + ph.DfN * dot3(dS, dF) + lsme + lsce_s * s.Con_EQ[j][k];
  }
}

#ifndef VEF_COUNT_FLOPS
static void to_double(const vreal* a, double* b, int n) {
  for (int i = 0; i < n; i++) b[i] = val(a[i]);
}
#endif

extern "C" {

#ifndef VEF_COUNT_FLOPS

// Gauss tables: out[0:3] P1, [3:6] W1, [6:24] P2, [24:33] W2, then 6 bulk
// tables [9 gp][9 node] each, then 6 edge tables [4][3][9] each.
int vo_tables(double* out) {
  const Tables& T = tables();
  double* o = out;
  for (int i = 0; i < 3; i++) *o++ = T.P1[i];
  for (int i = 0; i < 3; i++) *o++ = T.W1[i];
  for (int g = 0; g < 9; g++) { *o++ = T.P2[g][0]; *o++ = T.P2[g][1]; }
  for (int g = 0; g < 9; g++) *o++ = T.W2[g];
  const double(*bt[6])[9] = {T.bfn, T.dfdc, T.dfde, T.d2fdc2, T.d2fde2, T.d2fdce};
  for (int t = 0; t < 6; t++)
    for (int g = 0; g < 9; g++)
      for (int n = 0; n < 9; n++) *o++ = bt[t][g][n];
  const double(*et[6])[3][9] = {T.bfnE, T.dfdcE, T.dfdeE, T.d2fdc2E, T.d2fde2E, T.d2fdceE};
  for (int t = 0; t < 6; t++)
    for (int ed = 0; ed < 4; ed++)
      for (int g = 0; g < 3; g++)
        for (int n = 0; n < 9; n++) *o++ = et[t][ed][g][n];
  return (int)(o - out);
}

// Geometry at one bulk Gauss point: out = [detJ, WET, dFdXi(2x9), d2FdXi2(2x9)].
int vo_basis2d(const double* xy /*[9][2]*/, int gp, int edge /*0 bulk, 1..4*/, double* out) {
  const Tables& T = tables();
  Elem E;
  for (int n = 0; n < 9; n++) {
    E.X[n][0] = xy[n * 2];
    E.X[n][1] = xy[n * 2 + 1];
    E.X[n][2] = 0;
  }
  if (edge == 0) {
    E.bfn = T.bfn[gp]; E.dfdc = T.dfdc[gp]; E.dfde = T.dfde[gp];
    E.d2fdc2 = T.d2fdc2[gp]; E.d2fde2 = T.d2fde2[gp]; E.d2fdce = T.d2fdce[gp];
  } else {
    int ed = edge - 1;
    E.bfn = T.bfnE[ed][gp]; E.dfdc = T.dfdcE[ed][gp]; E.dfde = T.dfdeE[ed][gp];
    E.d2fdc2 = T.d2fdc2E[ed][gp]; E.d2fde2 = T.d2fde2E[ed][gp]; E.d2fdce = T.d2fdceE[ed][gp];
  }
  GsPt G;
  basis2D(E, G);
  out[0] = G.detJ;
  out[1] = (edge == 0) ? T.W2[gp] * std::fabs(G.detJ) : 0.0;
  for (int r = 0; r < 2; r++)
    for (int n = 0; n < 9; n++) {
      out[2 + r * 9 + n] = E.dFdXi[r][n];
      out[20 + r * 9 + n] = E.d2FdXi2[r][n];
    }
  return 0;
}

// Interpolated fields at bulk Gauss point gp (setFlowUnknowns):
// out = U[2], GU[2][2] (GU(j,i) = dU_i/dx_j), G2U[2][2], P, dPdXi[2],
//       S[3] (xx,xy,yy), dSdXi[3][2]  -> 22 values.
int vo_gauss_fields(const double* xy, const double* TL, int gp, double* out) {
  const Tables& T = tables();
  Elem E;
  for (int n = 0; n < 9; n++) {
    for (int q = 0; q < 6; q++) E.TL[n][q] = TL[n * 6 + q];
    E.X[n][0] = xy[n * 2];
    E.X[n][1] = xy[n * 2 + 1];
    E.X[n][2] = 0;
  }
  E.bfn = T.bfn[gp]; E.dfdc = T.dfdc[gp]; E.dfde = T.dfde[gp];
  E.d2fdc2 = T.d2fdc2[gp]; E.d2fde2 = T.d2fde2[gp]; E.d2fdce = T.d2fdce[gp];
  GsPt G;
  basis2D(E, G);
  setFlowUnknowns(E, G);
  double* o = out;
  *o++ = G.U_f[0];
  *o++ = G.U_f[1];
  for (int j = 0; j < 2; j++)
    for (int i = 0; i < 2; i++) *o++ = G.GU[j][i];
  for (int j = 0; j < 2; j++)
    for (int i = 0; i < 2; i++) *o++ = G.G2U[j][i];
  *o++ = G.P;
  *o++ = G.dPdXi[0];
  *o++ = G.dPdXi[1];
  *o++ = G.S[0][0];
  *o++ = G.S[0][1];
  *o++ = G.S[1][1];
  const int comp[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  for (int s2 = 0; s2 < 3; s2++)
    for (int d = 0; d < 2; d++) *o++ = G.dSdXi[comp[s2][0]][comp[s2][1]][d];
  return (int)(o - out);
}

// Pointwise strong form at one point from explicit field values, for the
// closed-form checks.  in: U[2], GU[2][2] (GU(j,i) = dU_i/dx_j), G2U[2][2]
// (G2U(j,i) = d2U_i/dx_j2), P, dPdXi[2], S[3] (xx,xy,yy), dS[3][2], EU.
// out: Con_EQ[2][2], Momentum[2], Continuity, T_ve[3][3], P_tot[3][3], W12,
//      UCD[2][2], f_T, C[2][2] -> 4+2+1+9+9+1+4+1+4 = 35 values.
int vo_point_strong_form(const vo_params* p, const double* U, const double* GU, const double* G2U, double P,
                         const double* dP, const double* S, const double* dS, double EU, double* out) {
  Phys ph = make_phys(p);
  vreal I[3][3];
  identity(I);
  GsPt G;
  for (int i = 0; i < 3; i++) G.U_f[i] = 0.0, G.dPdXi[i] = 0.0;
  zero33(G.GU);
  zero33(G.G2U);
  for (int i = 0; i < 2; i++) {
    G.U_f[i] = U[i];
    G.dPdXi[i] = dP[i];
    for (int j = 0; j < 2; j++) {
      G.GU[j][i] = GU[j * 2 + i];
      G.G2U[j][i] = G2U[j * 2 + i];
    }
  }
  G.P = P;
  zero33(G.S);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      for (int c = 0; c < 3; c++) G.dSdXi[a][b][c] = 0.0;
  const int comp[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  for (int s = 0; s < 3; s++) {
    int j = comp[s][0], k = comp[s][1];
    G.S[j][k] = S[s];
    G.S[k][j] = S[s];
    for (int d = 0; d < 2; d++) {
      G.dSdXi[j][k][d] = dS[s * 2 + d];
      G.dSdXi[k][j][d] = dS[s * 2 + d];
    }
  }
  G.WiN = ph.WiN0 * EU;
  setFlowKinematics(G, I);
  vreal UCD[3][3];
  strongFormConEQ_SQRT(ph, G, I, UCD);
  strongFormMomentum(ph, G);
  G.Continuity = trace33(G.GU);
  double* o = out;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) *o++ = G.Con_EQ[i][j];
  *o++ = G.Momentum[0];
  *o++ = G.Momentum[1];
  *o++ = G.Continuity;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) *o++ = G.T_ve[i][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) *o++ = G.P_tot[i][j];
  *o++ = omega12(G.S, G.GUT_tr);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) *o++ = UCD[i][j];
  *o++ = G.f_T;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) *o++ = G.C[i][j];
  return (int)(o - out);
}

// One element: residual (and FD Jacobian).  ied = 0 bulk, 1..4 outlet edge.
// Res[9][6]; A[54*54] (column-block layout of TEMP_A_f); Ac[54].
int vo_element(const vo_params* p, const double* xy, const double* TL /*[9][6]*/, double EU, int ied, int jac,
               double* Res, double* A, double* Ac) {
  Phys ph = make_phys(p);
  Elem E;
  for (int n = 0; n < 9; n++) {
    for (int q = 0; q < 6; q++) E.TL[n][q] = TL[n * 6 + q];
    E.X[n][0] = xy[n * 2];
    E.X[n][1] = xy[n * 2 + 1];
    E.X[n][2] = 0.0;
  }
  E.EU = EU;
  vreal R[9][6];
  std::vector<vreal> Av(jac ? 54 * 54 : 0), Acv(jac ? 54 : 0);
  elementFD(ph, E, ied, jac != 0, R, jac ? Av.data() : nullptr, jac ? Acv.data() : nullptr);
  to_double(&R[0][0], Res, 54);
  if (jac) {
    to_double(Av.data(), A, 54 * 54);
    to_double(Acv.data(), Ac, 54);
  }
  return 0;
}

// Pattern size and arrays (preAllocatePetsc).
int64_t vo_pattern_nnz(int n_nodes, int n_el, const int32_t* conn) {
  CSR P;
  build_pattern(n_nodes, n_el, conn, P);
  return P.rowptr[P.n];
}
int vo_pattern(int n_nodes, int n_el, const int32_t* conn, int64_t* rowptr, int32_t* col) {
  CSR P;
  build_pattern(n_nodes, n_el, conn, P);
  std::memcpy(rowptr, P.rowptr.data(), sizeof(int64_t) * (P.n + 1));
  std::memcpy(col, P.col.data(), sizeof(int32_t) * P.rowptr[P.n]);
  return 0;
}

// Full assembly in the reference's order: initializeLSObjects (zero b, and the
// matrix values if jac) then setEquationsBulk over all elements, then
// setEquationsEdge over the outlet edges (Newton_Raphson_mod.F90:332-357).
// Element-local work of a block of elements runs on `nthreads` OpenMP threads;
// the ADD_VALUES scatter is applied serially in element order.
int vo_assemble(const vo_params* p, int n_nodes, int n_el, const double* coords, const int32_t* conn,
                int n_edges, const int32_t* edges /*[n][2] (elem, ied)*/, const double* TL, double EU, int jac,
                const int64_t* rowptr, const int32_t* col, double* b, double* vals, int el_begin, int el_end,
                int nthreads) {
  Phys ph = make_phys(p);
  int64_t N = (int64_t)n_nodes * 6 + 1;
  int64_t nnz = rowptr[N];
  std::fill(b, b + N, 0.0);
  if (jac) std::fill(vals, vals + nnz, 0.0);
  if (el_end < 0 || el_end > n_el) el_end = n_el;
  const int BLK = 256;
  std::vector<double> Rbuf(BLK * 54), Abuf(jac ? (size_t)BLK * 2916 : 0), Acbuf(jac ? BLK * 54 : 0);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  for (int e0 = el_begin; e0 < el_end; e0 += BLK) {
    int e1 = std::min(el_end, e0 + BLK);
#pragma omp parallel for schedule(static)
    for (int e = e0; e < e1; e++) {
      Elem E;
      copyToLocal(E, coords, conn, e, TL, EU);
      vreal R[9][6];
      int l = e - e0;
      elementFD(ph, E, 0, jac != 0, R, jac ? &Abuf[(size_t)l * 2916] : nullptr, jac ? &Acbuf[(size_t)l * 54] : nullptr);
      to_double(&R[0][0], &Rbuf[(size_t)l * 54], 54);
    }
    for (int e = e0; e < e1; e++) {
      int l = e - e0;
      if (scatter_element(conn + (size_t)e * 9, &Rbuf[(size_t)l * 54], jac ? &Abuf[(size_t)l * 2916] : nullptr,
                          jac ? &Acbuf[(size_t)l * 54] : nullptr, N, rowptr, col, b, vals))
        return -1;
    }
  }
  if (el_begin == 0 && el_end == n_el) {
    for (int k = 0; k < n_edges; k++) {
      int e = edges[2 * k], ied = edges[2 * k + 1];
      Elem E;
      copyToLocal(E, coords, conn, e, TL, EU);
      vreal R[9][6];
      double Rd[54];
      std::vector<double> A(jac ? 2916 : 0), Ac(jac ? 54 : 0);
      elementFD(ph, E, ied, jac != 0, R, jac ? A.data() : nullptr, jac ? Ac.data() : nullptr);
      to_double(&R[0][0], Rd, 54);
      if (scatter_element(conn + (size_t)e * 9, Rd, jac ? A.data() : nullptr, jac ? Ac.data() : nullptr, N, rowptr,
                          col, b, vals))
        return -1;
    }
  }
  return 0;
}

// a14: the arclength row after the bulk assembly (arclength / pseudoArclength,
// src/Extra_Equations_mod.F90:698-828).  mode 0 (Increment <= 2): b(N-1) =
// EU - (LN0 + (Increment-1)*dArN) and, with the Jacobian, MatZeroRowsColumns
// of row/column N-1 with diagonal 1.  mode 1: b(N-1) = sum_ieq
// dot(dTLodS(:,ieq), TL(:,ieq)-TLo(:,ieq)) + dot(dEUodS, EU-EUo) - dSo and row
// N-1 = [dTLodS (node-major), dEUodS].  t is the tangent in the global
// ordering (t[6g+q] = dTLodS(g,q), t[N-1] = dEUodS).
int vo_arclength(int mode, int n_nodes, const double* x, const double* xo, const double* t, int increment,
                 double LN0, double dArN, double dSo, int jac, const int64_t* rowptr, const int32_t* col,
                 double* vals, double* b) {
  const int64_t N = (int64_t)n_nodes * 6 + 1;
  if (mode == 0) {
    b[N - 1] = x[N - 1] - (LN0 + (double)(increment - 1) * dArN);
    if (jac) {
      for (int64_t r = 0; r < N; r++)
        for (int64_t p = rowptr[r]; p < rowptr[r + 1]; p++)
          if (r == N - 1 || col[p] == N - 1) vals[p] = (r == N - 1 && col[p] == N - 1) ? 1.0 : 0.0;
    }
    return 0;
  }
  double value = 0.0;
  for (int q = 0; q < 6; q++) {
    double d = 0.0;
    for (int64_t g = 0; g < n_nodes; g++) d = d + t[g * 6 + q] * (x[g * 6 + q] - xo[g * 6 + q]);
    value = value + d;
  }
  value = value + t[N - 1] * (x[N - 1] - xo[N - 1]) - dSo;
  b[N - 1] = value;
  if (jac)
    for (int64_t p = rowptr[N - 1]; p < rowptr[N]; p++) vals[p] = t[col[p]];
  return 0;
}

// f4: drag on a boundary (calculateEdgeQuantities with ieq_ed = 1,
// src/Post_Process_mod.F90:1796-1890): term = sum over the boundary's edges,
// in (element, edge) order, of term_el = sum over the 3 edge Gauss points of
// dot_product(-normal, P_tot(:,1))*WET, with the Gauss-point quantities of
// setFlowQuantities and setEdgeQuantities as in the outlet residual.
// per_edge (optional) receives each term_el.
int vo_drag(const vo_params* p, int n_nodes, int n_el, const double* coords, const int32_t* conn, const double* x,
            int n_edges, const int32_t* edges, double* Fd, double* per_edge) {
  (void)n_el;
  Phys ph = make_phys(p);
  const Tables& T = tables();
  vreal I[3][3];
  identity(I);
  const double EU = x[(size_t)n_nodes * 6];
  vreal term = Z();
  for (int k = 0; k < n_edges; k++) {
    const int e = edges[2 * k], ied = edges[2 * k + 1];
    if (ied < 1 || ied > 4) return -1;
    Elem E;
    copyToLocal(E, coords, conn, e, x, EU);
    const int ed = ied - 1;
    vreal term_el = Z();
    for (int g = 0; g < 3; g++) {
      E.bfn = T.bfnE[ed][g];
      E.dfdc = T.dfdcE[ed][g];
      E.dfde = T.dfdeE[ed][g];
      E.d2fdc2 = T.d2fdc2E[ed][g];
      E.d2fde2 = T.d2fde2E[ed][g];
      E.d2fdce = T.d2fdceE[ed][g];
      GsPt G;
      basis2D(E, G);
      setFlowQuantities(ph, E, G, I);
      vreal dxdc = G.dXidci[0][0], dydc = G.dXidci[0][1];
      vreal dxde = G.dXidci[1][0], dyde = G.dXidci[1][1];
      G.normal[2] = Z();
      switch (ied) {
        case 1: G.normal[0] = dydc; G.normal[1] = -dxdc; break;
        case 2: G.normal[0] = dyde; G.normal[1] = -dxde; break;
        case 3: G.normal[0] = -dydc; G.normal[1] = dxdc; break;
        default: G.normal[0] = -dyde; G.normal[1] = dxde; break;
      }
      G.dL = vsqrt(G.normal[0] * G.normal[0] + G.normal[1] * G.normal[1]);
      G.normal[0] = G.normal[0] / G.dL;
      G.normal[1] = G.normal[1] / G.dL;
      G.WET = vreal(T.W1[g]) * G.dL;
      vreal mn[3] = {-G.normal[0], -G.normal[1], -G.normal[2]};
      vreal col[3] = {G.P_tot[0][0], G.P_tot[1][0], G.P_tot[2][0]};
      term_el = term_el + dot3(mn, col) * G.WET;
    }
    if (per_edge) per_edge[k] = val(term_el);
    term = term + term_el;
  }
  *Fd = val(term);
  return 0;
}

// f4: nodal stress tensor of the SQRT formulation (nodalStressTensorSQRT,
// src/Post_Process_mod.F90:2176-2236) for the model names whose g1 = g2 = 1
// (every name but m-L-PTT, m-e-PTT and the FENE models, so L-PTT and
// GIESEKUS): S from TL(4:6) (getStressComponent, Tools_mod.F90:769-793),
// C = matmul(S,S), T = ((1-BvN)/WiN)*(g2*C - g1*I).  WiN is the module value
// the caller holds.  out[g][4] = T(1,1), T(1,2), T(2,2), T(3,3).
int vo_nodal_stress(double BvN, double WiN, int n_nodes, const double* x, double* out) {
  for (int g = 0; g < n_nodes; g++) {
    double S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    S[0][0] = x[(size_t)g * 6 + 3];
    S[0][1] = x[(size_t)g * 6 + 4];
    S[1][0] = S[0][1];
    S[1][1] = x[(size_t)g * 6 + 5];
    double C[3][3], Tt[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s = s + S[i][k] * S[k][j];
        C[i][j] = s;
      }
    const double g1 = 1.0, g2 = 1.0;
    const double a = (1.0 - BvN) / WiN;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) Tt[i][j] = a * (g2 * C[i][j] - g1 * (i == j ? 1.0 : 0.0));
    out[(size_t)g * 4 + 0] = Tt[0][0];
    out[(size_t)g * 4 + 1] = Tt[0][1];
    out[(size_t)g * 4 + 2] = Tt[1][1];
    out[(size_t)g * 4 + 3] = Tt[2][2];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// f3: stability assembly entry points.
// ---------------------------------------------------------------------------
static void load_elem(Elem& E, const double* xy, const double* TL, double EU) {
  for (int n = 0; n < 9; n++) {
    for (int q = 0; q < 6; q++) E.TL[n][q] = TL[n * 6 + q];
    E.X[n][0] = xy[n * 2];
    E.X[n][1] = xy[n * 2 + 1];
    E.X[n][2] = 0.0;
  }
  E.EU = EU;
}

int vo_stab_plant(int which) {
  g_stab_plant = which;
  return 0;
}

// Plants a mistake in setStabilizationTerms (0 removes it):
//  1 Helem = 1/sum (no square root)      2 tsugn without abs()
//  3 BvN/Helem in tau_LSME               4 T:T without the 1/2 in Ha
//  5 Ha not divided by sqrt(1+G:G/2)     6 tau_LSIC = Helem**2 * tau_LSME
//  7 tau_LSCE's WiN*tsugn without WiN    8 G:G without the 1/2 in tau_LSCE
//  9 f(1) in place of f(2)               10 T:T in place of C:C
// 11 tdcs without the square root        12 DfN/Helem in tau_LSCE
// 13 1/dt in place of ReN/dt             14 1/dt in place of WiN/dt
int vo_tau_plant(int which) {
  if (which < 0 || which > 14) return -1;
  g_tau_plant = which;
  return 0;
}

// One bulk Gauss point of an element: the quantities setStabilizationTerms
// forms (Element_Calculations_mod.F90:701-773) from the element's nodal values.
// out: Helem**2, tsugn, tlsme, tlsic, tlsce, tdcs.
int vo_gauss_tau(const vo_params* p, const double* xy, const double* TL, double EU, int gp, double* out) {
  if (gp < 0 || gp > 8) return -1;
  Phys ph = make_phys(p);
  const Tables& T = tables();
  Elem E;
  load_elem(E, xy, TL, EU);
  E.bfn = T.bfn[gp]; E.dfdc = T.dfdc[gp]; E.dfde = T.dfde[gp];
  E.d2fdc2 = T.d2fdc2[gp]; E.d2fde2 = T.d2fde2[gp]; E.d2fdce = T.d2fdce[gp];
  vreal I[3][3];
  identity(I);
  GsPt G;
  basis2D(E, G);
  setFlowQuantities(ph, E, G, I);
  out[0] = G.H2;
  out[1] = G.tsugn;
  out[2] = G.tlsme;
  out[3] = G.tlsic;
  out[4] = G.tlsce;
  out[5] = G.tdcs;
  return 0;
}

// Plants a mistake in the test-node terms (0 removes it):
//  1 ReN*Convection*bifn without ReN         2 P_tot(i,:).dFdXi with the wrong sign
//  3 LSMEMomentum without ReN                4 LSMEMomentum times the other Momentum component
//  5 LSICMomentum with the other dFdXi       6 LSCEMomentumSQRT times WiN instead of over WiN
//  7 LSCEMomentumSQRT with GU_lsce untransposed   8 LSCEMomentumSQRT without (1-BvN)
//  9 Continuity*bifn with the wrong sign     10 LSMEContinuity with the wrong sign
// 11 Con_EQ(j,j) for Con_EQ(j,k)             12 DfN term with the wrong sign
// 13 LSMEStressSQRT without its j /= k term  14 LSMEStressSQRT over WiN instead of times WiN
// 15 LSCEStressSQRT f_T*bifn without the 1/2 16 LSCEStressSQRT without WiN on U.dFdXi
// 17 DCS term without mag_con_EQ             18 LSMEStressSQRT with the wrong sign
int vo_term_plant(int which) {
  if (which < 0 || which > 18) return -1;
  g_term_plant = which;
  return 0;
}

// The summands of TERM_RES at bulk Gauss point gp for each test node, as
// nodeResidual adds them: out[n*27 + c] with c = 5*i + t for the momentum
// rows (t: ReN*Convection*bifn, P_tot.dFdXi, LSME, LSIC, LSCE), 10 + t for
// the continuity row (Continuity*bifn, LSME), 12 + 5*s + t for the stress
// rows s = xx, xy, yy (Con_EQ*bifn, DfN term, LSME, LSCE, DCS); out[243] =
// WET, out[244] = tdcs * mag_con_EQ (real part).
int vo_node_terms(const vo_params* p, const double* xy, const double* TL, double EU, int gp, double* out) {
  if (gp < 0 || gp > 8) return -1;
  Phys ph = make_phys(p);
  const Tables& T = tables();
  Elem E;
  load_elem(E, xy, TL, EU);
  E.bfn = T.bfn[gp]; E.dfdc = T.dfdc[gp]; E.dfde = T.dfde[gp];
  E.d2fdc2 = T.d2fdc2[gp]; E.d2fde2 = T.d2fde2[gp]; E.d2fdce = T.d2fdce[gp];
  vreal I[3][3];
  identity(I);
  GsPt G;
  basis2D(E, G);
  setFlowQuantities(ph, E, G, I);
  G.WET = vreal(T.W2[gp]) * fabsr(G.detJ);
  for (int n = 0; n < 9; n++) {
    NodeTerms t;
    nodeTerms(ph, E, G, n, t);
    double* o = out + n * 27;
    for (int i = 0; i < 2; i++)
      for (int k = 0; k < 5; k++) o[5 * i + k] = val(t.mom[i][k]);
    o[10] = val(t.cont[0]);
    o[11] = val(t.cont[1]);
    for (int s = 0; s < 3; s++)
      for (int k = 0; k < 5; k++) o[12 + 5 * s + k] = val(t.str[s][k]);
  }
  out[243] = val(G.WET);
  out[244] = val(G.tdcs * G.mag_re);
  return 0;
}

// Element matrices of storeBulkStability: A, B [54 columns (inod,ieq)][54 rows (jnod,jeq)].
int vo_stab_element(const vo_params* p, const double* xy, const double* TLb, double EU, double* A, double* B) {
  Phys ph = make_phys(p);
  if (!ph.dcs_off) return -2;
  Elem E;
  load_elem(E, xy, TLb, EU);
  std::vector<vreal> Av(2916), Bv(2916);
  elementStability(ph, E, Av.data(), Bv.data());
  to_double(Av.data(), A, 2916);
  to_double(Bv.data(), B, 2916);
  return 0;
}

// Res_a, Res_b of residualBulkStability for a given disturbance TLd[9][6].
int vo_stab_residual(const vo_params* p, const double* xy, const double* TLb, double EU, const double* TLd,
                     double* ResA, double* ResB) {
  Phys ph = make_phys(p);
  if (!ph.dcs_off) return -2;
  Elem E;
  load_elem(E, xy, TLb, EU);
  vreal D[9][6], RA[9][6], RB[9][6];
  for (int n = 0; n < 9; n++)
    for (int q = 0; q < 6; q++) D[n][q] = TLd[n * 6 + q];
  residualBulkStability(ph, E, D, RA, RB);
  to_double(&RA[0][0], ResA, 54);
  to_double(&RB[0][0], ResB, 54);
  return 0;
}

// At bulk Gauss point gp: stab[9][6] = the stability test-node terms
// TERM_RES_a for disturbance TLd, and fd[9][6] = the central difference
// (T(b + h d) - T(b - h d))/(2h) of the base test-node terms with the
// test-function coefficients frozen at the base state b.
int vo_stab_linearization(const vo_params* p, const double* xy, const double* TLb, double EU, const double* TLd,
                          int gp, double h, double* stab, double* fd) {
  Phys ph = make_phys(p);
  if (!ph.dcs_off) return -2;
  const Tables& T = tables();
  vreal I[3][3];
  identity(I);
  Elem E;
  load_elem(E, xy, TLb, EU);
  setGaussPoint(E, T, gp);
  GsPt b;
  basis2D(E, b);
  setFlowQuantities(ph, E, b, I);
  vreal D[9][6];
  for (int n = 0; n < 9; n++)
    for (int q = 0; q < 6; q++) D[n][q] = TLd[n * 6 + q];
  GsPtD d;
  setFlowQuantitiesD(ph, E, D, b, d, I);
  GsPt sp, sm;
  Elem Ep = E, Em = E;
  for (int n = 0; n < 9; n++)
    for (int q = 0; q < 6; q++) {
      Ep.TL[n][q] = E.TL[n][q] + vreal(h) * D[n][q];
      Em.TL[n][q] = E.TL[n][q] - vreal(h) * D[n][q];
    }
  basis2D(Ep, sp);
  setFlowQuantities(ph, Ep, sp, I);
  basis2D(Em, sm);
  setFlowQuantities(ph, Em, sm, I);
  for (int n = 0; n < 9; n++) {
    vreal TA[6], TB[6], Tp[6], Tm[6];
    nodeResidualStability(ph, E, b, d, n, TA, TB);
    nodeResidualFrozen(ph, E, b, sp, n, Tp);
    nodeResidualFrozen(ph, E, b, sm, n, Tm);
    for (int q = 0; q < 6; q++) {
      stab[n * 6 + q] = val(TA[q]);
      fd[n * 6 + q] = (val(Tp[q]) - val(Tm[q])) / (2.0 * h);
    }
  }
  return 0;
}

// Pattern of the stability matrices (preAllocatePetsc with Nex = 0): the node
// blocks only, rows/cols 6*gnod + ieq.
static void build_pattern_stability(int n_nodes, int n_el, const int32_t* conn, CSR& P) {
  CSR Q;
  build_pattern(n_nodes, n_el, conn, Q);  // + the arclength column (last in each row) and row
  int64_t N = (int64_t)n_nodes * 6;
  P.n = N;
  P.rowptr.assign(N + 1, 0);
  for (int64_t r = 0; r < N; r++) P.rowptr[r + 1] = P.rowptr[r] + (Q.rowptr[r + 1] - Q.rowptr[r] - 1);
  P.col.resize(P.rowptr[N]);
  for (int64_t r = 0; r < N; r++)
    std::memcpy(&P.col[P.rowptr[r]], &Q.col[Q.rowptr[r]], sizeof(int32_t) * (P.rowptr[r + 1] - P.rowptr[r]));
}
int64_t vo_stab_pattern_nnz(int n_nodes, int n_el, const int32_t* conn) {
  CSR P;
  build_pattern_stability(n_nodes, n_el, conn, P);
  return P.rowptr[P.n];
}
int vo_stab_pattern(int n_nodes, int n_el, const int32_t* conn, int64_t* rowptr, int32_t* col) {
  CSR P;
  build_pattern_stability(n_nodes, n_el, conn, P);
  std::memcpy(rowptr, P.rowptr.data(), sizeof(int64_t) * (P.n + 1));
  std::memcpy(col, P.col.data(), sizeof(int32_t) * P.rowptr[P.n]);
  return 0;
}

// setEquationsBulkStability over all elements (MatZeroEntries first): element
// matrices on `nthreads` OpenMP threads, the ADD_VALUES scatter serially in
// element order.  x[6*n_nodes + 1] is the base state (TL node-major, EU last).
int vo_stab_assemble(const vo_params* p, int n_nodes, int n_el, const double* coords, const int32_t* conn,
                     const double* x, const int64_t* rowptr, const int32_t* col, double* valsA, double* valsB,
                     int el_begin, int el_end, int nthreads) {
  Phys ph = make_phys(p);
  if (!ph.dcs_off) return -2;
  const int64_t N = (int64_t)n_nodes * 6;
  const double EU = x[N];
  std::fill(valsA, valsA + rowptr[N], 0.0);
  std::fill(valsB, valsB + rowptr[N], 0.0);
  if (el_end < 0 || el_end > n_el) el_end = n_el;
  const int BLK = 256;
  std::vector<double> Abuf((size_t)BLK * 2916), Bbuf((size_t)BLK * 2916);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  for (int e0 = el_begin; e0 < el_end; e0 += BLK) {
    int e1 = std::min(el_end, e0 + BLK);
#pragma omp parallel for schedule(static)
    for (int e = e0; e < e1; e++) {
      Elem E;
      copyToLocal(E, coords, conn, e, x, EU);
      std::vector<vreal> Av(2916), Bv(2916);
      elementStability(ph, E, Av.data(), Bv.data());
      to_double(Av.data(), &Abuf[(size_t)(e - e0) * 2916], 2916);
      to_double(Bv.data(), &Bbuf[(size_t)(e - e0) * 2916], 2916);
    }
    for (int e = e0; e < e1; e++) {
      const int32_t* nodes = conn + (size_t)e * 9;
      const double* A = &Abuf[(size_t)(e - e0) * 2916];
      const double* B = &Bbuf[(size_t)(e - e0) * 2916];
      for (int jn = 0; jn < 9; jn++)
        for (int jq = 0; jq < 6; jq++) {
          int64_t row = (int64_t)nodes[jn] * 6 + jq;
          for (int in = 0; in < 9; in++)
            for (int iq = 0; iq < 6; iq++) {
              int64_t s = find_slot(rowptr, col, row, nodes[in] * 6 + iq);
              if (s < 0) return -1;
              valsA[s] += A[(in * 6 + iq) * 54 + jn * 6 + jq];
              valsB[s] += B[(in * 6 + iq) * 54 + jn * 6 + jq];
            }
        }
    }
  }
  return 0;
}

#else  // VEF_COUNT_FLOPS

// Operation counts of one element-residual evaluation (bulk or edge) at a given
// state: out = [add, mul, div, sqrt, total].
int vo_count_flops(const vo_params* p, const double* xy, const double* TL, double EU, int ied, uint64_t* out) {
  Phys ph = make_phys(p);
  Elem E;
  for (int n = 0; n < 9; n++) {
    for (int q = 0; q < 6; q++) E.TL[n][q] = vreal(TL[n * 6 + q]);
    E.X[n][0] = xy[n * 2];
    E.X[n][1] = xy[n * 2 + 1];
    E.X[n][2] = 0.0;
  }
  E.EU = vreal(EU);
  vreal R[9][6];
  FlopCounter::reset();
  if (ied == 0) residualBulk(ph, E, R);
  else residualEdge(ph, E, ied, R);
  out[0] = FlopCounter::add;
  out[1] = FlopCounter::mul;
  out[2] = FlopCounter::div;
  out[3] = FlopCounter::sqrt;
  out[4] = FlopCounter::total();
  return 0;
}

// Operation counts of the geometry-only part of one residualBulk evaluation,
// all 9 Gauss points: basis2D, Helem**2 and the two parameter-over-Helem**2
// quotients of setStabilizationTerms (BvN/H2, DfN/H2).  out = add, mul, div,
// sqrt, total.
int vo_count_flops_geometry(const vo_params* p, const double* xy, uint64_t* out) {
  Phys ph = make_phys(p);
  Elem E;
  for (int n = 0; n < 9; n++) {
    for (int q = 0; q < 6; q++) E.TL[n][q] = vreal(0.0);
    E.X[n][0] = xy[n * 2];
    E.X[n][1] = xy[n * 2 + 1];
    E.X[n][2] = 0.0;
  }
  const Tables& T = tables();
  FlopCounter::reset();
  for (int g = 0; g < 9; g++) {
    setGaussPoint(E, T, g);
    GsPt b;
    basis2D(E, b);
    vreal H2 = elementH2(E);
    vreal r3 = ph.BvN / H2;
    vreal t7 = ph.DfN / H2;
    (void)r3;
    (void)t7;
  }
  out[0] = FlopCounter::add;
  out[1] = FlopCounter::mul;
  out[2] = FlopCounter::div;
  out[3] = FlopCounter::sqrt;
  out[4] = FlopCounter::total();
  return 0;
}

// Operation counts of the stability element: out[0..4] one residualBulkStability
// evaluation (add, mul, div, sqrt, total), out[5..9] its base part (basis2D and
// the base state's setFlowQuantities at the 9 Gauss points).
int vo_count_flops_stability(const vo_params* p, const double* xy, const double* TLb, double EU, const double* TLd,
                             uint64_t* out) {
  Phys ph = make_phys(p);
  Elem E;
  for (int n = 0; n < 9; n++) {
    for (int q = 0; q < 6; q++) E.TL[n][q] = vreal(TLb[n * 6 + q]);
    E.X[n][0] = xy[n * 2];
    E.X[n][1] = xy[n * 2 + 1];
    E.X[n][2] = 0.0;
  }
  E.EU = vreal(EU);
  vreal D[9][6], RA[9][6], RB[9][6];
  for (int n = 0; n < 9; n++)
    for (int q = 0; q < 6; q++) D[n][q] = vreal(TLd[n * 6 + q]);
  FlopCounter::reset();
  residualBulkStability(ph, E, D, RA, RB);
  out[0] = FlopCounter::add;
  out[1] = FlopCounter::mul;
  out[2] = FlopCounter::div;
  out[3] = FlopCounter::sqrt;
  out[4] = FlopCounter::total();
  const Tables& T = tables();
  vreal I[3][3];
  identity(I);
  FlopCounter::reset();
  for (int g = 0; g < 9; g++) {
    setGaussPoint(E, T, g);
    GsPt b;
    basis2D(E, b);
    setFlowQuantities(ph, E, b, I);
  }
  out[5] = FlopCounter::add;
  out[6] = FlopCounter::mul;
  out[7] = FlopCounter::div;
  out[8] = FlopCounter::sqrt;
  out[9] = FlopCounter::total();
  return 0;
}

#endif

}  // extern "C"
