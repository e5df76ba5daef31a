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
 * every block the blocked elimination touches inside the storage; nb more
 * columns past column n - 1 are zero padding the caller leaves alone).  The
 * caller fills AB (zero outside the band) and calls vef_band_factor; the
 * factor overwrites AB with L (unit lower, below the diagonal) and U, and
 * keeps the inverses of the nb x nb diagonal blocks of L and U for the
 * solves.  All calls are ordered on the given HIP stream; one host thread
 * per handle.
 *
 * A handle made with vef_band_create_prec(..., 32, ...) stores the band, the
 * block inverses and the panel products in FP32 (the values handed to
 * vef_band_set are rounded to float; the factorisation runs in FP32 with
 * sgemm) while the solve's vectors and sums stay FP64: the single-precision
 * factorisation of an inexact Newton iteration whose residual and update are
 * FP64.
 */
#ifndef VEF_BAND_H
#define VEF_BAND_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vef_band vef_band;

/* nb: block size, 64 (the only value this build supports).  FP64 storage. */
int vef_band_create(int64_t n, int32_t w, int32_t nb, vef_band** out);

/* As vef_band_create with the stored precision: bits = 64 (double) or 32
 * (float; see above). */
int vef_band_create_prec(int64_t n, int32_t w, int32_t nb, int32_t bits, vef_band** out);

/* Device pointer of the band storage (FP64 handles only: pass ab = NULL for
 * an FP32 handle), its leading dimension and K. */
int vef_band_storage(vef_band* h, double** ab, int64_t* ld, int32_t* K);

/* Zeroes the band, then AB[dst[i]] = values[src[i]] for i < count (device
 * arrays; dst indices are band positions K + i - j + j*ld; values are FP64
 * and rounded to the handle's precision). */
int vef_band_set(vef_band* h, int64_t count, const int64_t* src, const int64_t* dst, const double* values,
                 void* hip_stream);

/* In-place LU without pivoting of the stored band; inverses of the diagonal
 * blocks.  A zero pivot gives non-finite factors (the caller's residual
 * check sees them).  From the second call on a handle the launch sequence
 * is replayed as one HIP graph (captured once; direct launches if the
 * capture fails). */
int vef_band_factor(vef_band* h, void* hip_stream);

/* 1 when vef_band_factor replays its captured graph, else 0. */
int vef_band_uses_graph(const vef_band* h);

/* x = A^-1 b with the last factors (device arrays of n doubles; b is used as
 * work space and overwritten). */
int vef_band_solve(vef_band* h, double* b, double* x, void* hip_stream);

const char* vef_band_last_error(const vef_band* h);
void vef_band_destroy(vef_band* h);

#ifdef __cplusplus
}
#endif

#endif
