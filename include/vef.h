/*
 * vef.h — C ABI of the MI355X residual/Jacobian assembler for the 2D
 * square-root-conformation viscoelastic flow past a cylinder (9-node Q2
 * quadrangles, unknowns Ux Uy P Sxx Sxy Syy per node plus one arclength
 * unknown).
 *
 * The assembler replaces, behind one call, the reference's per-Newton-iteration
 * assembly sequence
 *     initializeLSObjects        (src/Petsc_mod.F90:747-763: zero b, and A if NRP)
 *     setEquationsBulk           (src/Equations_mod.F90:21-55, called at
 *                                 src/Newton_Raphson_mod.F90:332-333)
 *     setEquationsEdge (Outlet)  (src/Equations_mod.F90:317-342, called at
 *                                 src/Newton_Raphson_mod.F90:349-350)
 * i.e. the element residuals, the forward-difference element Jacobians
 * (eps_jac = 1e-8, src/Equations_mod.F90:264-306) and their ADD_VALUES scatter
 * (src/Storage_mod.F90:10-115) into the AIJ pattern that
 * preAllocatePetsc builds (src/Petsc_mod.F90:629-743).  The caller keeps the
 * rest of assemblePetscObjects (arclength row, Dirichlet rows) and the solve.
 *
 * Conventions
 *   - All indices are 0-based.  Global unknown of node g, equation q is
 *     row/column 6*g + q (q: 0 Ux, 1 Uy, 2 P, 3 Sxx, 4 Sxy, 5 Syy); the
 *     arclength unknown is the last one, N-1, with N = 6*n_nodes + 1
 *     (getRows, src/Tools_mod.F90:735-752; storeJacobian_Ac jcol = Mcols-1).
 *   - The state is the solution vector x[N] in that ordering: x[6g+q] is
 *     TL(g,q) and x[N-1] is EU(Nex), the arclength parameter (WiN = WiN0*EU).
 *   - The matrix is CSR with int64 row pointers and int32 column indices,
 *     rows sorted; values are FP64.
 *   - Functions return 0 on success and a nonzero VEF_ERR_* code otherwise;
 *     vef_last_error() gives the message.  Nothing aborts the process.
 *   - One handle is used by one host thread at a time.  Work is ordered on the
 *     HIP stream passed in (NULL = the default stream).
 */
#ifndef VEF_H
#define VEF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vef_handle vef_handle;

enum { VEF_MODEL_LPTT = 0, VEF_MODEL_GIESEKUS = 1 };
/* Flag_NR: 'MNR' residual only, 'NRP' residual and Jacobian (src/Equations_mod.F90:264).
 * VEF_RES_JAC_FP32 has no counterpart in the reference: the residual as with
 * VEF_RES_JAC (FP64, the reference's arithmetic) and an analytic Jacobian of
 * the element residual computed in FP32 (forward-mode differentiation of the
 * Gauss-point terms, contraction with the basis functions) in place of the
 * forward differences, for an inexact Newton; values are stored as FP64.
 * It needs DfN > 1e-12. */
enum { VEF_RES_ONLY = 0, VEF_RES_JAC = 1, VEF_RES_JAC_FP32 = 2 };
enum {
  VEF_OK = 0,
  VEF_ERR_ARG = 1,
  VEF_ERR_MESH = 2,
  VEF_ERR_HIP = 3,
  VEF_ERR_UNSUPPORTED = 4,
  VEF_ERR_NOMEM = 5
};

/* Dimensionless numbers (src/Parameters_mod.F90:98-111, src/Physical_mod.F90:327-351).
 * model VEF_MODEL_LPTT: VeN is the PTT epsilon (VeN = 0 is Oldroyd-B);
 * model VEF_MODEL_GIESEKUS: VeN is the mobility alpha.  DfN > 1e-12 switches
 * the discontinuity-capturing term off and adds DfN*grad(S):grad(phi). */
typedef struct {
  double ReN, BvN, VeN, DfN, WiN0;
  int32_t model;
  int32_t reserved; /* must be 0 */
} vef_params;

/* Mesh: coords[n_nodes][2] (x, y); connectivity[n_elements][9] in the
 * reference's local order (corners (-1,-1) (1,-1) (1,1) (-1,1), mid-edges
 * bottom right top left, centre; src/GaussIntegration/Gauss_Quadrangle_mod.F90:413-421);
 * outlet_edges[n_outlet_edges][2] = (element, edge id 1..4) ascending, the
 * edge ids of setEdgeQuantities (src/Element_Calculations_mod.F90:887-905).
 * Builds the CSR pattern and the device-side scatter map; copies the mesh.
 * Replaces: createPetscMatrix + preAllocatePetsc (src/Petsc_mod.F90:290-339,629-743)
 * and the set-up of the Gauss tables (src/GaussIntegration/Gauss_Quadrangle_mod.F90:14-34). */
int vef_create(int32_t n_nodes, int32_t n_elements, const double* coords, const int32_t* connectivity,
               int32_t n_outlet_edges, const int32_t* outlet_edges, const vef_params* params,
               vef_handle** out);

/* Replaces changing the module globals ReN, BvN, VeN, DfN and WiN0 between
 * continuation steps (src/Physical_mod.F90:327-351). */
int vef_set_params(vef_handle* h, const vef_params* params);

/* N = 6*n_nodes + 1 and the number of stored entries of the pattern. */
int64_t vef_num_rows(const vef_handle* h);
int64_t vef_nnz(const vef_handle* h);

/* Copies the CSR pattern to host arrays rowptr[N+1], colind[nnz]
 * (MatGetRowIJ of the preallocated matrix). */
int vef_get_pattern(const vef_handle* h, int64_t* rowptr, int32_t* colind);

/* Device copies of the same pattern, owned by the handle. */
const int64_t* vef_device_rowptr(const vef_handle* h);
const int32_t* vef_device_colind(const vef_handle* h);

/* One assembly on the GPU, all pointers device pointers.
 *   x[N]       state (TL node-major, EU last)
 *   b[N]       residual: zeroed, then element and outlet-edge contributions
 *              added in the reference's ADD_VALUES order
 *   values[nnz] matrix values (written with VEF_RES_JAC or VEF_RES_JAC_FP32,
 *              may be NULL with VEF_RES_ONLY): zeroed, then the element
 *              Jacobians (forward differences, or the FP32 analytic ones)
 *              added in the same order; column N-1 holds dR/dEU.  Row N-1
 *              stays zero (the caller's arclength row goes there).
 * Replaces initializeLSObjects + setEquationsBulk + setEquationsEdge for the
 * Cylinder_2D outlet (src/Newton_Raphson_mod.F90:101,332-357). */
int vef_assemble(vef_handle* h, const double* x, int32_t flag, double* b, double* values, void* hip_stream);

/* The arclength row N-1, applied after vef_assemble (arclength and
 * pseudoArclength, src/Extra_Equations_mod.F90:698-828).
 *   flag: VEF_RES_JAC and VEF_RES_JAC_FP32 write the row of values alike.
 *   VEF_ARC_PIN (continuation steps 1-2): b[N-1] = x[N-1] - (LN0 + (increment-1)*dArN);
 *     with VEF_RES_JAC, row and column N-1 are zeroed and the diagonal set to 1
 *     (MatZeroRowsColumns).  x_old and tangent may be NULL.
 *   VEF_ARC_PSEUDO (steps >= 3): b[N-1] = tangent . (x - x_old) - dS, summed per
 *     equation over the nodes then the arclength entry, as the reference does;
 *     with VEF_RES_JAC, row N-1 = tangent.  The sum uses a fixed blocked order
 *     (deterministic; within 1e-13 of the sequential sum relative to the sum of
 *     |terms|).  tangent[N] holds dTLodS node-major and dEUodS last. */
enum { VEF_ARC_PIN = 0, VEF_ARC_PSEUDO = 1 };
int vef_arclength(vef_handle* h, int32_t mode, const double* x, const double* x_old, const double* tangent,
                  int32_t increment, double LN0, double dArN, double dS, int32_t flag, double* b, double* values,
                  void* hip_stream);

/* As vef_arclength, allocating nothing and using no scratch of the handle:
 * no counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The partial
 * sums of the pseudo-arclength dot product go to a file-scope device array of
 * 6 x 4096 doubles (192 KiB per GPU, kept until the process ends), so meshes
 * of at most 16,777,216 nodes (larger ones: VEF_ERR_UNSUPPORTED).  Every call
 * on one GPU shares that array: calls that overlap on one GPU (other streams,
 * other threads) race.  Results equal vef_arclength's bit for bit. */
int vef_arclength_noalloc(vef_handle* h, int32_t mode, const double* x, const double* x_old, const double* tangent,
                          int32_t increment, double LN0, double dArN, double dS, int32_t flag, double* b,
                          double* values, void* hip_stream);

/* Dirichlet rows, applied after the assembly and the arclength row
 * (applyDirichletBCFullyDeveloped and applyDirichletBC in assemblePetscObjects,
 * src/Dirichlet_mod.F90:104-190, src/Newton_Raphson_mod.F90:384-394).
 * vef_set_dirichlet stores rows (host int64[n]: global rows 6*gnod + ieq) with
 * their prescribed values (host double[n]); a row listed more than once keeps
 * its last value, as the reference's boundary-by-boundary INSERT_VALUES does.
 * vef_apply_dirichlet sets b[row] = x[row] - value and, with VEF_RES_JAC (or
 * VEF_RES_JAC_FP32), zeroes the row's stored values and puts 1 on its diagonal
 * (MatZeroRows with diag = 1). */
int vef_set_dirichlet(vef_handle* h, int64_t n_rows, const int64_t* rows, const double* values);
int vef_apply_dirichlet(vef_handle* h, const double* x, int32_t flag, double* b, double* values, void* hip_stream);

/* Drag on a boundary, the Fd the reference prints after each converged step
 * (calculateEdgeQuantities with ieq_ed = 1, src/Post_Process_mod.F90:1796-1890;
 * the edges of the Cylinder boundary from setBoundaryEdges,
 * src/Boundary_mod.F90:477-522).  vef_set_drag_edges stores the boundary's
 * edges (host int32[n][2]: element 0-based, edge id 1..4), in the order the
 * reference lists them (ascending element, then edge id).  vef_drag computes
 *   Fd = sum over the edges of term_el,
 *   term_el = sum over the 3 edge Gauss points of dot(-normal, P_tot(:,1))*WET,
 * both sums sequential, with P_tot at WiN = WiN0*x[N-1].  Fd: device double[1];
 * per_edge: device double[n] for the term_el values, or NULL. */
int vef_set_drag_edges(vef_handle* h, int32_t n_edges, const int32_t* edges);
int vef_drag(vef_handle* h, const double* x, double* Fd, double* per_edge, void* hip_stream);

/* Nodal stress tensor of the square-root formulation at every node
 * (nodalStressTensorSQRT, src/Post_Process_mod.F90:2176-2236, the wall-stress
 * output of calculateNodalQuantities, :1992-2168):
 *   T = ((1-BvN)/WiN)*(C - I), C = S.S,
 * for the L-PTT and Giesekus model names (g1 = g2 = 1 there).  WiN is the value
 * of the reference's module variable at post-processing time, passed by the
 * caller.  T: device double[n_nodes][4] = (Txx, Txy, Tyy, Tzz). */
int vef_nodal_stress(vef_handle* h, const double* x, double WiN, double* T, void* hip_stream);

/* Linear stability assembly of Cylinder_Stability_2D (6 unknowns per node,
 * no arclength unknown) for the unchanged SLEPc shift-invert solve.
 * Replaces assembleSlepcObjects (src/Stability/Stability_mod.F90:97-199) for
 * this problem: MatZeroEntries of A_f and B_f, setEquationsBulkStability
 * (src/Stability/Equations_Stability_mod.F90:20-57: element residuals
 * residualBulkStability and the forward-difference matrices of
 * storeBulkStability, :61-226, in the disturbance around the base state, with
 * eps_jac = 1e-8, and their ADD_VALUES scatter), and applyDirichletBCStability
 * (src/Dirichlet_mod.F90:287-316).  No edge equations apply to the names of
 * this problem (setBoundaryEQsStability, src/Boundary_mod.F90:319-439), and
 * no periodic boundaries are configured (src/Parameters_mod.F90:152).
 *   - The base state is x[N] as for vef_assemble (TL node-major, EU last;
 *     WiN = WiN0*EU).  The reference's values are complex with zero imaginary
 *     parts on this path; A and B hold the real parts (FP64).
 *   - Pattern: the node blocks of vef_get_pattern without the arclength
 *     column and row: rowptr[6*n_nodes + 1], colind[vef_stability_nnz].
 *   - Requires DfN > 1e-12 (the discontinuity-capturing coefficient is then
 *     zero); otherwise VEF_ERR_UNSUPPORTED.
 *   - vef_set_stability_dirichlet stores the rows (host int64[n], 6*gnod +
 *     ieq) of the Inlet and Outlet (all unknowns) and Wall and Cylinder
 *     (velocity) boundaries; vef_stability_assemble then sets them to the rows
 *     of the identity in A and to zero in B (MatZeroRows with diag 1 and 0).
 *   - A, B: device double[vef_stability_nnz], fully overwritten.
 * While profiling is enabled (vef_profile), the element kernel counts as
 * bulk_ms and the ordered sums and Dirichlet rows as scatter_ms. */
int64_t vef_stability_nnz(const vef_handle* h);
int vef_get_stability_pattern(const vef_handle* h, int64_t* rowptr, int32_t* colind);
int vef_set_stability_dirichlet(vef_handle* h, int64_t n_rows, const int64_t* rows);
int vef_stability_assemble(vef_handle* h, const double* x, double* A, double* B, void* hip_stream);

/* The same with host arrays (copies in, assembles, copies out). */
int vef_assemble_host(vef_handle* h, const double* x, int32_t flag, double* b, double* values);

/* Kernel timing, the counterpart of the reference's "Bulk EQs time" and
 * "Assembly time" prints (src/Newton_Raphson_mod.F90:337,398).  While enabled,
 * each vef_assemble records HIP events on its stream around the
 * forward-difference element kernel (bulk_ms) and around the kernels after it
 * (ordered sums of shared entries, outlet-edge entries, b; scatter_ms).  vef_profile_read waits for
 * the recorded events, returns the summed milliseconds and the number of
 * assemblies since the last read, and clears them.  (Re-)enabling clears too. */
int vef_profile(vef_handle* h, int32_t enable);
int vef_profile_read(vef_handle* h, double* bulk_ms, double* scatter_ms, int64_t* launches);

/* Last error message of the handle (or of the last failed vef_create when h
 * is NULL). */
const char* vef_last_error(const vef_handle* h);

void vef_destroy(vef_handle* h);

#ifdef __cplusplus
}
#endif

#endif /* VEF_H */
