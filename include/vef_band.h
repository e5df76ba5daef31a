/* vef_band.h — banded LU without pivoting on the GPU, for the Newton matrix
 * of the cylinder solver in a bandwidth-reducing node order.
 *
 * Replaces the direct factorisation and solve the reference's Newton
 * iteration hands to PETSc/MUMPS (solveLinearSystem's KSPSolve,
 * src/Petsc_mod.F90:807-871, with KSPPREONLY, PCLU and MATSOLVERMUMPS set in
 * src/Petsc_mod.F90:127-135) for the node unknowns; the caller handles the
 * dense arclength row and column as a border (Schur complement of one
 * scalar).  No pivoting: the Newton matrices of this formulation factor
 * stably in the node order (the caller checks the first solve's residual).
 *
 * Storage: n unknowns, half-bandwidth w; the library stores A(i, j) for
 * |i - j| <= K = w + nb in a column-major band, AB[(K + i - j) + j*ld] with
 * ld = 2K + 1 (the LAPACK band layout with kl = ku = K; the extra nb keeps
 * every block the blocked elimination touches inside the storage).  The
 * caller fills AB (zero outside the band) and calls vef_band_factor; the
 * factor overwrites AB with L (unit lower, below the diagonal) and U, and
 * keeps the inverses of the nb x nb diagonal blocks of L and U for the
 * solves.  All calls are ordered on the given HIP stream; one host thread
 * per handle.
 */
#ifndef VEF_BAND_H
#define VEF_BAND_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vef_band vef_band;

/* nb: block size, 64 (the only value this build supports). */
int vef_band_create(int64_t n, int32_t w, int32_t nb, vef_band** out);

/* Device pointer of the band storage, its leading dimension and K. */
int vef_band_storage(vef_band* h, double** ab, int64_t* ld, int32_t* K);

/* Zeroes the band, then AB[dst[i]] = values[src[i]] for i < count (device
 * arrays; dst indices are band positions K + i - j + j*ld). */
int vef_band_set(vef_band* h, int64_t count, const int64_t* src, const int64_t* dst, const double* values,
                 void* hip_stream);

/* In-place LU without pivoting of the stored band; inverses of the diagonal
 * blocks.  A zero pivot gives non-finite factors (the caller's residual
 * check sees them). */
int vef_band_factor(vef_band* h, void* hip_stream);

/* x = A^-1 b with the last factors (device arrays of n doubles; b is used as
 * work space and overwritten). */
int vef_band_solve(vef_band* h, double* b, double* x, void* hip_stream);

const char* vef_band_last_error(const vef_band* h);
void vef_band_destroy(vef_band* h);

#ifdef __cplusplus
}
#endif

#endif
