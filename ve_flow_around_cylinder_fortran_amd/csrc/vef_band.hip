// vef_band.hip — banded LU without pivoting (see include/vef_band.h).
//
// Blocked right-looking elimination in blocks of NB = 64 columns.  Step k
// (columns k0 .. k0+kb-1, m2 = min(w, n - k0 - kb) rows/columns below/right):
//   1. band_diag_lu     LU of the kb x kb diagonal block and the inverses of
//                       its L and U, one workgroup, rows in registers;
//   2. rocblas_dgemm    T1 = A21 U11^-1 and T2 = L11^-1 A12 (the panels);
//   3. rocblas_dgemm    A22 -= T1 T2 (the m2 x m2 trailing block);
//   4. band_panel_store T1 and T2 into the band (L21, U12 for the solves).
// The inverses are kept for the solves.  Without pivoting the factors have no
// bitwise target (the caller checks the first solve's residual), so this
// file is compiled with FMA contraction.
// The band view A(i, j) = AB[K + i + j*(ld - 1)] makes every block a plain
// column-major matrix with leading dimension ld - 1, which is what the dgemm
// takes.  A block of the last step with kb < NB is padded with the identity
// so every loop has the compile-time length NB.
// The solves push the updates: forward, per block, y_k = L11^-1 b_k and
// b[below] -= L21 y_k; backward, per block from the last, x_k = U11^-1 y_k and
// y[above] -= U(above, k) x_k.  One launch per block and sweep, threads over
// the rows (coalesced down the band's columns).
//
// The factorisation's launch sequence depends only on the handle (its
// buffers are fixed), so from the second call on it is replayed as a HIP
// graph captured once on a private stream: the 5 launches per step otherwise
// cost more host time than their GPU time.  If the capture fails the
// launches stay direct.
//
// Every kernel is a template on the stored type T: double, or float for a
// single-precision factorisation (the band, the block inverses and the panels
// in FP32, sgemm for the products; the solves' vectors and sums stay FP64).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cstdio>
#include <utility>
#include <cstring>
#include <string>

#include "vef_band.h"

namespace {

constexpr int NB = 64;

// LU of the diagonal block and the inverses of its L and U.  Thread
// (i, g) = (tid / 4, tid % 4) keeps entries j = g, g+4, ... of row i of the
// block, of L^-1 and of U^-1 in registers.  At each step the finished row is
// published to LDS (double-buffered, so a step needs one barrier) and read
// back as a broadcast; the element of a thread's own row that a step needs
// comes from the lane of the same row that holds it (a shuffle: the four
// threads of a row are adjacent lanes of one wavefront).  The steps are
// template instances, so every register index is a compile-time constant.
constexpr int RC = NB / 4;  // register entries per thread and row

template <class T>
struct DiagRows {
  T d[RC], x[RC], y[RC];
  int i, g, row0;
  T (*P)[NB];
  T (*Q)[NB];
};

// Reciprocal estimate refined by one Newton step.
__device__ __forceinline__ double recip(double v) {
  double r = __builtin_amdgcn_rcp(v);
  return r * (2.0 - v * r);
}
__device__ __forceinline__ float recip(float v) {
  float r = __builtin_amdgcn_rcpf(v);
  return r * (2.0f - v * r);
}

template <int K, class T>
__device__ __forceinline__ void lu_step(DiagRows<T>& R) {
  constexpr int b = K & 1, kq = K >> 2, kr = K & 3;
  // D[i][K] is final after step K-1: fetched before the barrier so the
  // shuffle's latency overlaps the wait
  const T dik = __shfl(R.d[kq], R.row0 | kr);
  if (R.i == K) {
#pragma unroll
    for (int q = 0; q < RC; q++) R.P[b][R.g + 4 * q] = R.d[q];
  }
  __syncthreads();
  if (R.i > K) {
    const T l = dik * recip(R.P[b][K]);
#pragma unroll
    for (int q = 0; q < RC; q++) {
      const int j = R.g + 4 * q;
      if (j > K) R.d[q] -= l * R.P[b][j];
    }
    if (R.g == kr) R.d[kq] = l;
  }
}

// Step K of L^-1 (X[i] -= L[i][K] X[K] over j <= K) together with step
// KU = NB-1-K of U^-1 (Y[KU] /= U[KU][KU]; Y[i] -= U[i][KU] Y[KU] over j >= KU
// for the rows above).
template <int K, class T>
__device__ __forceinline__ void inv_step(DiagRows<T>& R) {
  constexpr int b = K & 1, KU = NB - 1 - K;
  const T uik = __shfl(R.d[KU >> 2], R.row0 | (KU & 3));  // U[i][KU]
  const T lik = __shfl(R.d[K >> 2], R.row0 | (K & 3));    // L[i][K] (D is final here)
  if (R.i == KU) {
    const T inv = recip(uik);
#pragma unroll
    for (int q = 0; q < RC; q++) {
      R.y[q] *= inv;
      R.Q[b][R.g + 4 * q] = R.y[q];
    }
  }
  if (R.i == K) {
#pragma unroll
    for (int q = 0; q < RC; q++) R.P[b][R.g + 4 * q] = R.x[q];
  }
  __syncthreads();
  if (R.i > K) {
#pragma unroll
    for (int q = 0; q < RC; q++) {
      const int j = R.g + 4 * q;
      if (j <= K) R.x[q] -= lik * R.P[b][j];
    }
  }
  if (R.i < KU) {
#pragma unroll
    for (int q = 0; q < RC; q++) {
      const int j = R.g + 4 * q;
      if (j >= KU) R.y[q] -= uik * R.Q[b][j];
    }
  }
}

template <class T, int... Ks>
__device__ __forceinline__ void lu_steps(DiagRows<T>& R, std::integer_sequence<int, Ks...>) {
  (lu_step<Ks>(R), ...);
}
template <class T, int... Ks>
__device__ __forceinline__ void inv_steps(DiagRows<T>& R, std::integer_sequence<int, Ks...>) {
  (inv_step<Ks>(R), ...);
}

template <class T>
__global__ void __launch_bounds__(256)
band_diag_lu(T* __restrict__ AB, int64_t ld, int32_t K, int64_t k0, int kb, T* __restrict__ invL,
             T* __restrict__ invU) {
  __shared__ T P[2][NB];  // published row: of L\U, then of L^-1
  __shared__ T Q[2][NB];  // published row of U^-1
  DiagRows<T> R;
  const int tid = threadIdx.x;
  R.i = tid >> 2;
  R.g = tid & 3;
  R.row0 = (tid & 63) & ~3;
  R.P = P;
  R.Q = Q;
  const int64_t lda = ld - 1;
  T* base = AB + K + k0 + k0 * lda;  // A(k0, k0)
#pragma unroll
  for (int q = 0; q < RC; q++) {
    const int j = R.g + 4 * q;
    R.d[q] = (R.i < kb && j < kb) ? base[R.i + j * lda] : T(R.i == j ? 1 : 0);  // identity padding
    R.x[q] = T(R.i == j ? 1 : 0);
    R.y[q] = R.x[q];
  }
  lu_steps(R, std::make_integer_sequence<int, NB - 1>{});
  inv_steps(R, std::make_integer_sequence<int, NB>{});
#pragma unroll
  for (int q = 0; q < RC; q++) {
    const int j = R.g + 4 * q;
    if (R.i < kb && j < kb) base[R.i + j * lda] = R.d[q];
    invL[R.i + j * NB] = R.x[q];  // column-major NB x NB
    invU[R.i + j * NB] = R.y[q];
  }
}

// Loads a column-major NB x NB block into LDS (blockDim.x = 256, all loads in
// flight first).
template <class T>
__device__ __forceinline__ void stage_block(const T* __restrict__ src, T (*M)[NB + 1]) {
  constexpr int Q = NB * NB / 256;
  T v[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) v[q] = src[threadIdx.x + 256 * q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int e = threadIdx.x + 256 * q;
    M[e % NB][e / NB] = v[q];
  }
}

// The panels, computed by dgemm into T1 (m2 x kb, ld m2) and T2 (kb x m2, ld
// NB), go back into the band: A21 = T1, A12 = T2.
template <class T>
__global__ void __launch_bounds__(256)
band_panel_store(T* __restrict__ AB, int64_t ld, int32_t K, int64_t k0, int kb, int64_t m2, const T* __restrict__ T1,
                 const T* __restrict__ T2) {
  const int64_t lda = ld - 1;
  const int64_t total = m2 * kb;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e % m2, c = e / m2;             // A21(r, c) = T1[r + c*m2]
    AB[K + (k0 + kb + r) + (k0 + c) * lda] = T1[e];
    const int64_t i = e % kb, jc = e / kb;             // A12(i, jc) = T2[i + jc*NB]
    AB[K + (k0 + i) + (k0 + kb + jc) * lda] = T2[i + jc * NB];
  }
}

// The block inverse (column-major NB x NB) and the block's right-hand side go
// to LDS with all loads in flight; v = M r (padding rows of r are zero), summed
// in FP64 by four adjacent threads per row (16 terms each, two chains) and a
// shuffle reduction, so the dependent chain is 8 deep instead of 64.
template <class T>
__device__ __forceinline__ void block_gemv(const T* __restrict__ M, const double* __restrict__ rhs, int kb,
                                           T (*sM)[NB + 1], double* sr, double* out) {
  const int t = threadIdx.x;
  const double rv = (t < kb) ? rhs[t] : 0.0;
  stage_block(M, sM);
  if (t < NB) sr[t] = rv;
  __syncthreads();
  const int row = t >> 2, part = t & 3;
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int i = 0; i < NB / 4; i += 2) {
    s0 += (double)sM[row][part + 4 * i] * sr[part + 4 * i];
    s1 += (double)sM[row][part + 4 * i + 4] * sr[part + 4 * i + 4];
  }
  double s = s0 + s1;
  s += __shfl_xor(s, 1);
  s += __shfl_xor(s, 2);
  if (part == 0) out[row] = s;
  __syncthreads();
}

// sum_j a[j] y[j] over NB terms in four interleaved chains.
template <class T>
__device__ __forceinline__ double row_dot(const T (&a)[NB], const double* y) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
  for (int j = 0; j < NB; j += 4) {
    s0 += (double)a[j] * y[j];
    s1 += (double)a[j + 1] * y[j + 1];
    s2 += (double)a[j + 2] * y[j + 2];
    s3 += (double)a[j + 3] * y[j + 3];
  }
  return (s0 + s1) + (s2 + s3);
}

// Forward sweep, block k: y_k = L11^-1 b_k; b[k0+kb+r] -= L21(r, :) y_k.
template <class T>
__global__ void __launch_bounds__(256)
band_fwd(const T* __restrict__ AB, int64_t ld, int32_t K, int64_t k0, int kb, int64_t m2,
         const T* __restrict__ invLk, double* __restrict__ b, double* __restrict__ y) {
  __shared__ T sM[NB][NB + 1];
  __shared__ double sr[NB], yk[NB];
  const int t = threadIdx.x;
  const int64_t lda = ld - 1;
  // The row's band entries are loaded before the block product needs the
  // wait, unconditionally (rows past m2 read row k0 + kb; columns past n are
  // the storage's zero padding) so that all NB loads are in flight together.
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + t;
  const int64_t rr = r < m2 ? r : 0;
  const T* a = AB + K + (k0 + kb + rr) + k0 * lda;
  T av[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) av[j] = a[j * lda];
  block_gemv(invLk, b + k0, kb, sM, sr, yk);
  if (blockIdx.x == 0 && t < kb) y[k0 + t] = yk[t];
  if (r < m2) b[k0 + kb + r] -= row_dot(av, yk);  // yk[j] = 0 and av[j] = 0 past kb
}

// Backward sweep, block k: x_k = U11^-1 y_k; y[k0-n2+r] -= U(k0-n2+r, k0+j) x_k.
template <class T>
__global__ void __launch_bounds__(256)
band_bwd(const T* __restrict__ AB, int64_t ld, int32_t K, int64_t k0, int kb, int64_t n2,
         const T* __restrict__ invUk, double* __restrict__ y, double* __restrict__ x) {
  __shared__ T sM[NB][NB + 1];
  __shared__ double sr[NB], xk[NB];
  const int t = threadIdx.x;
  const int64_t lda = ld - 1;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + t;
  const int64_t row = k0 - n2 + (r < n2 ? r : 0);
  const T* a = AB + K + row + k0 * lda;  // A(row, k0); unconditional loads as in band_fwd
  T av[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) av[j] = a[j * lda];
  block_gemv(invUk, y + k0, kb, sM, sr, xk);
  if (blockIdx.x == 0 && t < kb) x[k0 + t] = xk[t];
  if (r < n2) y[row] -= row_dot(av, xk);  // xk[j] = 0 and av[j] = 0 past kb
}

template <class T>
__global__ void band_zero(T* __restrict__ a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = T(0);
}

// AB[dst[i]] = values[src[i]], rounded to T.
template <class T>
__global__ void band_scatter(T* __restrict__ AB, int64_t count, const int64_t* __restrict__ src,
                             const int64_t* __restrict__ dst, const double* __restrict__ v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    AB[dst[i]] = (T)v[src[i]];
}

rocblas_status gemm(rocblas_handle h, rocblas_int m, rocblas_int n, rocblas_int k, double alpha, const double* A,
                    rocblas_int lda, const double* B, rocblas_int ldb, double beta, double* C, rocblas_int ldc) {
  return rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb, &beta, C,
                       ldc);
}
rocblas_status gemm(rocblas_handle h, rocblas_int m, rocblas_int n, rocblas_int k, float alpha, const float* A,
                    rocblas_int lda, const float* B, rocblas_int ldb, float beta, float* C, rocblas_int ldc) {
  return rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb, &beta, C,
                       ldc);
}

}  // namespace

struct vef_band {
  int64_t n = 0;
  int32_t w = 0, K = 0;
  int64_t ld = 0;
  int64_t nblk = 0;
  int32_t bits = 64;     // stored type: 64 double, 32 float
  void* AB = nullptr;    // band, block inverses and panel products in the stored type
  void* invL = nullptr;
  void* invU = nullptr;
  void* T1 = nullptr;    // panel products of one step: w x NB and NB x w
  void* T2 = nullptr;
  double* y = nullptr;
  rocblas_handle rb = nullptr;
  hipStream_t cs = nullptr;       // capture stream of the factorisation graph
  hipGraphExec_t gexec = nullptr; // the factorisation as one graph (second call on)
  int64_t factor_calls = 0;
  bool graph_off = false;         // capture failed: launch directly
  std::string err;
};

namespace {
int fail(vef_band* h, const std::string& m) {
  if (h) h->err = m;
  return 1;
}
#define BAND_HIP(h, call)                                                       \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) return fail(h, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
}  // namespace

namespace {

template <class T>
int factor_impl(vef_band* h, hipStream_t st) {
  const int64_t lda = h->ld - 1;
  T* AB = (T*)h->AB;
  T* invL = (T*)h->invL;
  T* invU = (T*)h->invU;
  T* T1 = (T*)h->T1;
  T* T2 = (T*)h->T2;
  auto A = [&](int64_t i, int64_t j) { return AB + h->K + i + j * lda; };
  for (int64_t b = 0; b < h->nblk; b++) {
    const int64_t k0 = b * NB;
    const int kb = (int)std::min<int64_t>(NB, h->n - k0);
    const int64_t m2 = std::min<int64_t>(h->w, h->n - k0 - kb);
    hipLaunchKernelGGL(band_diag_lu<T>, dim3(1), dim3(256), 0, st, AB, h->ld, h->K, k0, kb, invL + b * NB * NB,
                       invU + b * NB * NB);
    BAND_HIP(h, hipGetLastError());
    if (m2 > 0) {
      const rocblas_int M2 = (rocblas_int)m2, KB = kb, LDA = (rocblas_int)lda;
      // T1 = A21 U11^-1 (m2 x kb), T2 = L11^-1 A12 (kb x m2), A22 -= T1 T2
      if (gemm(h->rb, M2, KB, KB, T(1), A(k0 + kb, k0), LDA, invU + b * NB * NB, NB, T(0), T1, M2) !=
              rocblas_status_success ||
          gemm(h->rb, KB, M2, KB, T(1), invL + b * NB * NB, NB, A(k0, k0 + kb), LDA, T(0), T2, NB) !=
              rocblas_status_success ||
          gemm(h->rb, M2, M2, KB, T(-1), T1, M2, T2, NB, T(1), A(k0 + kb, k0 + kb), LDA) != rocblas_status_success)
        return fail(h, "rocblas gemm failed");
      hipLaunchKernelGGL(band_panel_store<T>, dim3((unsigned)std::min<int64_t>(1024, (m2 * kb + 255) / 256)),
                         dim3(256), 0, st, AB, h->ld, h->K, k0, kb, m2, (const T*)T1, (const T*)T2);
      BAND_HIP(h, hipGetLastError());
    }
  }
  return 0;
}

template <class T>
int solve_impl(vef_band* h, double* b, double* x, hipStream_t st) {
  const T* AB = (const T*)h->AB;
  const T* invL = (const T*)h->invL;
  const T* invU = (const T*)h->invU;
  for (int64_t k = 0; k < h->nblk; k++) {
    const int64_t k0 = k * NB;
    const int kb = (int)std::min<int64_t>(NB, h->n - k0);
    const int64_t m2 = std::min<int64_t>(h->w, h->n - k0 - kb);
    const unsigned g = (unsigned)std::max<int64_t>(1, (m2 + 255) / 256);
    hipLaunchKernelGGL(band_fwd<T>, dim3(g), dim3(256), 0, st, AB, h->ld, h->K, k0, kb, m2, invL + k * NB * NB, b,
                       h->y);
    BAND_HIP(h, hipGetLastError());
  }
  for (int64_t k = h->nblk - 1; k >= 0; k--) {
    const int64_t k0 = k * NB;
    const int kb = (int)std::min<int64_t>(NB, h->n - k0);
    const int64_t n2 = std::min<int64_t>(h->w, k0);
    const unsigned g = (unsigned)std::max<int64_t>(1, (n2 + 255) / 256);
    hipLaunchKernelGGL(band_bwd<T>, dim3(g), dim3(256), 0, st, AB, h->ld, h->K, k0, kb, n2, invU + k * NB * NB, h->y,
                       x);
    BAND_HIP(h, hipGetLastError());
  }
  return 0;
}

}  // namespace

extern "C" {

int vef_band_create(int64_t n, int32_t w, int32_t nb, vef_band** out) {
  return vef_band_create_prec(n, w, nb, 64, out);
}

int vef_band_create_prec(int64_t n, int32_t w, int32_t nb, int32_t bits, vef_band** out) {
  if (!out) return 1;
  *out = nullptr;
  if (n < 1 || w < 0 || nb != NB || (bits != 64 && bits != 32)) return 1;
  vef_band* h = new vef_band();
  h->n = n;
  h->w = w;
  h->K = w + NB;
  h->ld = 2 * (int64_t)h->K + 1;
  h->nblk = (n + NB - 1) / NB;
  h->bits = bits;
  const size_t es = bits / 8;
  // NB columns of zero padding past column n - 1 (never written: every
  // factorisation step stays inside the n x n matrix) let the solves load a
  // whole block row without bounds tests at the last block.
  const size_t band_bytes = es * (n + NB) * h->ld;
  int rc = 0;
  if (hipMalloc(&h->AB, band_bytes) != hipSuccess ||
      hipMalloc(&h->invL, es * h->nblk * NB * NB) != hipSuccess ||
      hipMalloc(&h->invU, es * h->nblk * NB * NB) != hipSuccess ||
      hipMalloc((void**)&h->y, sizeof(double) * n) != hipSuccess ||
      hipMalloc(&h->T1, es * std::max<int64_t>(w, 1) * NB) != hipSuccess ||
      hipMalloc(&h->T2, es * std::max<int64_t>(w, 1) * NB) != hipSuccess)
    rc = 1;
  if (!rc && hipMemset(h->AB, 0, band_bytes) != hipSuccess) rc = 1;
  if (!rc && rocblas_create_handle(&h->rb) != rocblas_status_success) rc = 1;
  if (!rc && rocblas_set_atomics_mode(h->rb, rocblas_atomics_not_allowed) != rocblas_status_success) rc = 1;
  if (rc) {
    vef_band_destroy(h);
    return 1;
  }
  *out = h;
  return 0;
}

int vef_band_storage(vef_band* h, double** ab, int64_t* ld, int32_t* K) {
  if (!h || !ld || !K) return 1;
  if (ab) {
    if (h->bits != 64) return fail(h, "vef_band_storage: the band of an FP32 handle is not double (pass ab = NULL)");
    *ab = (double*)h->AB;
  }
  *ld = h->ld;
  *K = h->K;
  return 0;
}

int vef_band_set(vef_band* h, int64_t count, const int64_t* src, const int64_t* dst, const double* values,
                 void* stream) {
  if (!h) return 1;
  if (count < 0 || (count > 0 && (!src || !dst || !values))) return fail(h, "vef_band_set: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (h->bits == 64) {
    hipLaunchKernelGGL(band_zero<double>, dim3(2048), dim3(256), 0, st, (double*)h->AB, h->n * h->ld);
    BAND_HIP(h, hipGetLastError());
    if (count > 0) hipLaunchKernelGGL(band_scatter<double>, dim3(1024), dim3(256), 0, st, (double*)h->AB, count, src, dst,
                                      values);
  } else {
    hipLaunchKernelGGL(band_zero<float>, dim3(2048), dim3(256), 0, st, (float*)h->AB, h->n * h->ld);
    BAND_HIP(h, hipGetLastError());
    if (count > 0) hipLaunchKernelGGL(band_scatter<float>, dim3(1024), dim3(256), 0, st, (float*)h->AB, count, src, dst,
                                      values);
  }
  BAND_HIP(h, hipGetLastError());
  return 0;
}

int vef_band_factor(vef_band* h, void* stream) {
  if (!h) return 1;
  hipStream_t st = (hipStream_t)stream;
  if (h->gexec) {
    BAND_HIP(h, hipGraphLaunch(h->gexec, st));
    return 0;
  }
  if (h->factor_calls++ > 0 && !h->graph_off) {
    // Capture the sequence once (nothing runs while capturing), then launch it.
    if (!h->cs && hipStreamCreateWithFlags(&h->cs, hipStreamNonBlocking) != hipSuccess) h->cs = nullptr;
    if (h->cs && rocblas_set_stream(h->rb, h->cs) == rocblas_status_success &&
        hipStreamBeginCapture(h->cs, hipStreamCaptureModeRelaxed) == hipSuccess) {
      const int rc = h->bits == 64 ? factor_impl<double>(h, h->cs) : factor_impl<float>(h, h->cs);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(h->cs, &g);
      hipGraphExec_t x = nullptr;
      if (rc == 0 && e == hipSuccess && g && hipGraphInstantiate(&x, g, nullptr, nullptr, 0) == hipSuccess) h->gexec = x;
      if (g) (void)hipGraphDestroy(g);
    }
    (void)hipGetLastError();  // a failed capture leaves its error here; the direct path below reports its own
    h->err.clear();
    if (h->gexec) {
      BAND_HIP(h, hipGraphLaunch(h->gexec, st));
      return 0;
    }
    h->graph_off = true;
  }
  if (rocblas_set_stream(h->rb, st) != rocblas_status_success) return fail(h, "rocblas_set_stream failed");
  return h->bits == 64 ? factor_impl<double>(h, st) : factor_impl<float>(h, st);
}

int vef_band_uses_graph(const vef_band* h) { return h && h->gexec ? 1 : 0; }

int vef_band_solve(vef_band* h, double* b, double* x, void* stream) {
  if (!h || !b || !x) return 1;
  hipStream_t st = (hipStream_t)stream;
  return h->bits == 64 ? solve_impl<double>(h, b, x, st) : solve_impl<float>(h, b, x, st);
}

const char* vef_band_last_error(const vef_band* h) { return h ? h->err.c_str() : "handle is NULL"; }

void vef_band_destroy(vef_band* h) {
  if (!h) return;
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->cs) (void)hipStreamDestroy(h->cs);
  if (h->rb) rocblas_destroy_handle(h->rb);
  for (void* p : {h->AB, h->invL, h->invU, (void*)h->y, h->T1, h->T2})
    if (p) (void)hipFree(p);
  delete h;
}

}  // extern "C"
