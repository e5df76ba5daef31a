/* vef_mf.h — multifrontal LU of the node block of the Newton matrix on the
 * GPU, in a nested-dissection order worked out by the caller.
 *
 * Replaces the direct factorisation and solve the reference's Newton
 * iteration hands to PETSc/MUMPS (solveLinearSystem's KSPSolve,
 * src/Petsc_mod.F90:807-871, with KSPPREONLY, PCLU and MATSOLVERMUMPS set in
 * src/Petsc_mod.F90:127-141) for the node unknowns; the caller handles the
 * dense arclength row and column as a border (Schur complement of one
 * scalar) and keeps the symbolic analysis (the elimination tree and the
 * fronts' row lists).
 *
 * The plan: n_fronts fronts in n_buckets buckets.  Buckets are listed by
 * nondecreasing height (a front's height exceeds its children's), each with
 * count fronts padded to s_pad pivot rows and u_pad update rows; fronts are
 * numbered bucket by bucket.  Front f has front_s[f] <= s_pad pivot rows and
 * front_u[f] <= u_pad update rows; rows[rows_off[f] .. rows_off[f+1]) are the
 * unknowns (0 <= row < n) of its pivot rows, then of its update rows.  Its
 * parent front is parent[f] (-1 at a root) and it is child number slot[f]
 * (0 or 1) of it; ea_map[ea_off[f] + i] is the physical row, in the parent's
 * padded front, of update row i (a pivot row p of the parent is row p, an
 * update row j is row s_pad(parent) + j).
 *
 * Storage: the fronts of bucket b are consecutive column-major m x m blocks,
 * m = s_pad + u_pad, the buckets consecutive in order (vef_mf_layout gives
 * each front's element offset).  Row/column i < s_pad of a front is pivot i,
 * s_pad + j update j.  All calls are ordered on the given HIP stream; one
 * host thread per handle.
 *
 * Precision: bits = 64 stores and factorises the fronts in FP64; bits = 32
 * in FP32 (values rounded on entry; FP32 getrf/getrs/gemm), the solve's
 * global vectors staying FP64 (the single-precision factorisation of an
 * inexact Newton iteration, refined by its caller).
 */
#ifndef VEF_MF_H
#define VEF_MF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vef_mf vef_mf;

typedef struct {
  int64_t n;                      /* unknowns of the node block */
  int32_t n_buckets;
  const int32_t* bucket_height;   /* [n_buckets] */
  const int32_t* bucket_count;    /* [n_buckets] */
  const int32_t* bucket_s;        /* [n_buckets] padded pivot rows */
  const int32_t* bucket_u;        /* [n_buckets] padded update rows */
  int32_t n_fronts;
  const int32_t* front_s;         /* [n_fronts] */
  const int32_t* front_u;         /* [n_fronts] */
  const int64_t* rows_off;        /* [n_fronts + 1] */
  const int32_t* rows;            /* [rows_off[n_fronts]] */
  const int32_t* parent;          /* [n_fronts] */
  const int32_t* slot;            /* [n_fronts] */
  const int64_t* ea_off;          /* [n_fronts + 1] */
  const int32_t* ea_map;          /* [ea_off[n_fronts]] */
} vef_mf_plan;

/* Checks the plan (host arrays), copies what the device needs and allocates
 * the front storage.  bits: 64 or 32. */
int vef_mf_create(const vef_mf_plan* plan, int32_t bits, vef_mf** out);

/* Element offset of each front in the storage (front_base may be NULL) and
 * the storage's total element count. */
int vef_mf_layout(const vef_mf* h, int64_t* front_base, int64_t* total);

/* Bytes of device memory the handle holds. */
int64_t vef_mf_device_bytes(const vef_mf* h);

/* Clears the fronts, puts ones on the padded pivots' diagonal, then
 * front_storage[dst[k]] = values[k] for k < count with dst[k] >= 0 (device
 * arrays; dst values must be distinct). */
int vef_mf_set(vef_mf* h, int64_t count, const int64_t* dst, const double* values, void* hip_stream);

/* Multifrontal factorisation of the values set last: by height, extend-add
 * of the children's update blocks (first children, then second), then per
 * bucket getrf (partial pivoting within the pivot block), getrs for
 * F11^-1 F12 and gemm for F22 -= F21 F11^-1 F12. */
int vef_mf_factor(vef_mf* h, void* hip_stream);

/* Timing of the factorisation by height of the elimination tree (no
 * counterpart in the reference: the library's own profiling).  enable != 0
 * records HIP events on the caller's stream around each height of every
 * later vef_mf_factor; vef_mf_profile_read waits for the last one and writes
 * ms[2 h] = the extend-add of height h and ms[2 h + 1] = its buckets (getrf,
 * the solves for F11^-1 F12, the Schur update), for h < vef_mf_heights. */
int vef_mf_profile(vef_mf* h, int32_t enable);
int32_t vef_mf_heights(const vef_mf* h);
int vef_mf_profile_read(vef_mf* h, float* ms);

/* Row interchanges within each front's pivot block (on != 0, the default:
 * getrf's partial pivoting, as MUMPS pivots) or none (on = 0: getrf without
 * interchanges, for a factorisation whose solves are refined by the caller;
 * the nested-dissection order puts each node's velocities before its
 * pressure, so the pressure pivots are the velocity block's Schur
 * complement).  Takes effect at the next vef_mf_factor.  No counterpart in
 * the reference's interface (MUMPS's pivoting is its ICNTL setting). */
int vef_mf_set_pivoting(vef_mf* h, int32_t on);

/* Fronts whose last factorisation met a zero pivot (synchronises). */
int vef_mf_zero_pivots(vef_mf* h, int32_t* count);

/* x = A^-1 b with the last factors (device arrays of n doubles in the plan's
 * unknown order; b is used as work space and overwritten). */
int vef_mf_solve(vef_mf* h, double* b, double* x, void* hip_stream);

/* y = A x for a CSR matrix of n rows (FP64; device arrays, rowptr[n + 1]
 * and col of 64-bit integers), summed in a fixed order (deterministic).
 * Every row holding more than 16384 entries (the arclength row) must be
 * listed in long_rows[n_long] (a device array): with work (a device array
 * of n_long * 256 doubles) each such row is split into 256 fixed chunks
 * summed by a workgroup each, then the chunks by one workgroup; with work
 * NULL one workgroup sums the whole row.  The other rows are summed by a
 * wavefront each.  For the residual checks and the refinement of the
 * factorisation's callers; no counterpart in the reference (PETSc's
 * MatMult). */
int vef_mf_csr_mv(int64_t n, const int64_t* rowptr, const int64_t* col, const double* values, const double* x,
                  double* y, int32_t n_long, const int64_t* long_rows, double* work, void* hip_stream);

const char* vef_mf_last_error(const vef_mf* h);
void vef_mf_destroy(vef_mf* h);

#ifdef __cplusplus
}
#endif

#endif
