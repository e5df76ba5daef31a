// Multifrontal LU of the Newton matrix's node block on the GPU (C ABI in
// include/vef_mf.h): the role of MUMPS in the reference's solveLinearSystem
// (src/Petsc_mod.F90:127-141, 807-871).
//
// The symbolic analysis (nested dissection, fronts, buckets, maps) is the
// caller's; this library checks the plan, lays the padded fronts out bucket
// by bucket, and runs the numeric phases as height-ordered sequences of
// kernels and strided-batched rocSOLVER/rocBLAS calls (the buckets of one
// height spread over eight internal streams):
//   set:    clear, identity on the padded pivots, scatter of the CSR values;
//   factor: per height, extend-add of the children's update blocks into the
//           parents (slot 0 children, then slot 1: a fixed order), then per
//           bucket getrf(F11), getrs(F11, F12) and gemm(F22 -= F21 X);
//   pivots: after the factorisation, each front's row interchanges become a
//           permutation of its pivot unknowns (mf_perm), which the solve's
//           gather applies;
//   solve:  forward by height (gather P b_s, the L and U solves of every
//           front of the height in one launch, F21 w in another, ordered
//           subtract into the update unknowns), backward by height (gather
//           x_u, w -= X x_u in one launch, x_s = w).
// Sums whose order matters (extend-add, the forward subtraction) run in a
// fixed order, and every solve kernel sums in a fixed order, so results are
// deterministic.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "vef_mf.h"

namespace {

struct EaTask {
  int64_t cbase, pbase, moff;
  int32_t cm, cs, pm, u;
};

// Per-height index lists, concatenated over heights (offsets per height).
struct Lists {
  std::vector<int64_t> off;  // [H + 1]
};

template <typename T>
__global__ void mf_fill(T* __restrict__ a, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

template <typename T>
__global__ void mf_set_pos(T* __restrict__ a, const int64_t* __restrict__ pos, int64_t n, T v) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[pos[i]] = v;
}

template <typename T>
__global__ void mf_scatter(T* __restrict__ buf, const int64_t* __restrict__ dst, const double* __restrict__ vals,
                           int64_t n) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) {
    int64_t d = dst[k];
    if (d >= 0) buf[d] = (T)vals[k];
  }
}

// One block per (task, update column j) — task on grid x, column on grid y:
// the child's update column j is added into the parent's column map[j] at
// rows map[i].
template <typename T>
__global__ void mf_extend_add(T* __restrict__ buf, const EaTask* __restrict__ tasks, const int32_t* __restrict__ map) {
  const EaTask t = tasks[blockIdx.x];
  const int j = blockIdx.y;
  if (j >= t.u) return;
  const int32_t* mp = map + t.moff;
  const T* src = buf + t.cbase + (int64_t)(t.cs + j) * t.cm + t.cs;
  T* dst = buf + t.pbase + (int64_t)mp[j] * t.pm;
  for (int i = threadIdx.x; i < t.u; i += blockDim.x) dst[mp[i]] += src[i];
}

template <typename T>
__global__ void mf_gather(T* __restrict__ w, const int64_t* __restrict__ dst, const int32_t* __restrict__ src,
                          const double* __restrict__ v, int64_t n) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) w[dst[k]] = (T)v[src[k]];
}

// r[tgt[t]] -= sum over p in [ptr[t], ptr[t+1]) of V[vidx[p]], in order.
template <typename T>
__global__ void mf_subtract(double* __restrict__ r, const int32_t* __restrict__ tgt, const int64_t* __restrict__ ptr,
                            const int64_t* __restrict__ vidx, const T* __restrict__ V, int64_t n) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) {
    double acc = 0.0;
    for (int64_t p = ptr[t]; p < ptr[t + 1]; p++) acc += (double)V[vidx[p]];
    r[tgt[t]] -= acc;
  }
}

template <typename T>
__global__ void mf_put(double* __restrict__ x, const int32_t* __restrict__ dst, const int64_t* __restrict__ src,
                       const T* __restrict__ w, int64_t n) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) x[dst[k]] = (double)w[src[k]];
}

// A front as the solve kernels see it: storage offset, padded pivot and
// update counts, offsets of its pieces of W (pivot rows) and V/XU (update
// rows), and its real pivot count (the padded rows' slots of W are kept 0).
struct MfFront {
  int64_t base, w, v;
  int32_t s, u, sr;
};

constexpr int kSolveThreads = 256;
constexpr int kDiag = 64;  // diagonal block of the blocked triangular solves

// Pivot rows of the gather in the order of the last factorisation's row
// interchanges: for front f (one workgroup each), apply getrf's swaps
// (ipiv, 1-based within the front) to its pivot list in LDS (one thread:
// they are sequential), then psrc[w + i] = the unknown now at row i.  Padded
// rows are never interchanged (their columns are zero above an identity
// diagonal), so they keep -1.
__global__ void __launch_bounds__(kSolveThreads)
mf_perm(const MfFront* __restrict__ fr, const int* __restrict__ ipiv, const int64_t* __restrict__ fipiv,
        const int32_t* __restrict__ prow, int32_t* __restrict__ psrc) {
  extern __shared__ int32_t sp[];
  const MfFront d = fr[blockIdx.x];
  if (d.s == 0) return;
  const int* piv = ipiv + fipiv[blockIdx.x];
  for (int i = threadIdx.x; i < d.s; i += blockDim.x) sp[i] = prow[d.w + i];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < d.s; i++) {
      const int k = piv[i] - 1;
      if (k != i && k > i && k < d.s) {
        const int32_t t = sp[i];
        sp[i] = sp[k];
        sp[k] = t;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < d.s; i += blockDim.x) psrc[d.w + i] = sp[i];
}

// W[dst[k]] = v[psrc[dst[k]]] (the permuted pivot rows; padded slots stay 0).
template <typename T>
__global__ void mf_gather_p(T* __restrict__ w, const int64_t* __restrict__ dst, const int32_t* __restrict__ psrc,
                            const double* __restrict__ v, int64_t n) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) {
    const int64_t d = dst[k];
    const int32_t src = psrc[d];
    w[d] = src >= 0 ? (T)v[src] : T(0);
  }
}

// acc - sum_k col[k * m] * v[k] for k < n (n <= kDiag), k in order; the loads
// are issued 16 at a time so that many are in flight per thread.
// The value lane k (uniform across the wavefront) holds, read from its
// register (v_readlane) rather than through LDS.
__device__ __forceinline__ float lane_value(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
__device__ __forceinline__ double lane_value(double v, int k) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
  return __hiloint2double(hi, lo);
}

template <typename T>
__device__ __forceinline__ T dot_sub(T acc, const T* __restrict__ col, int64_t m, const T* v, int n) {
  constexpr int kB = 16;
  for (int k0 = 0; k0 < n; k0 += kB) {
    T a[kB];
#pragma unroll
    for (int u = 0; u < kB; u++) a[u] = k0 + u < n ? col[(int64_t)(k0 + u) * m] : T(0);
#pragma unroll
    for (int u = 0; u < kB; u++)
      if (k0 + u < n) acc -= a[u] * v[k0 + u];
  }
  return acc;
}
// Two rows at once (their loads in flight together, kB per row per batch);
// each row's terms are subtracted in the order dot_sub takes, so the results
// are the same bits.
template <typename T, int kB>
__device__ __forceinline__ void dot_sub2(T& acc0, T& acc1, const T* __restrict__ c0, const T* __restrict__ c1,
                                         int64_t m, const T* v, int n) {
  for (int k0 = 0; k0 < n; k0 += kB) {
    T a[kB], b[kB];
#pragma unroll
    for (int u = 0; u < kB; u++) {
      a[u] = k0 + u < n ? c0[(int64_t)(k0 + u) * m] : T(0);
      b[u] = k0 + u < n ? c1[(int64_t)(k0 + u) * m] : T(0);
    }
#pragma unroll
    for (int u = 0; u < kB; u++)
      if (k0 + u < n) {
        const T vk = v[k0 + u];
        acc0 -= a[u] * vk;
        acc1 -= b[u] * vk;
      }
  }
}
// w[i] -= A(i, j0 .. j0 + jb) w[j0 ..] for the rows i in [r0, r1) of thread
// tid's stride, two rows per pass.
template <typename T, int NT>
__device__ __forceinline__ void update_rows(T* w, const T* __restrict__ Aj, int64_t m, int j0, int jb, int r0,
                                            int r1, int tid) {
  int i = r0 + tid;
  for (; i + NT < r1; i += 2 * NT) {
    T a0 = w[i], a1 = w[i + NT];
    // 32 loads per row in FP32 where the registers allow it (no spill at
    // 256 threads; 1024 threads leave 128 registers)
    dot_sub2<T, (sizeof(T) == 4 && NT <= 256) ? 32 : 16>(a0, a1, Aj + i, Aj + i + NT, m, w + j0, jb);
    w[i] = a0;
    w[i + NT] = a1;
  }
  if (i < r1) w[i] = dot_sub(w[i], Aj + i, m, w + j0, jb);
}

// w = U^-1 L^-1 w for one front per workgroup of NT threads (F11 = L U from
// getrf, the permutation already applied by the gather).  Blocked by kDiag
// columns: the diagonal block is staged in LDS and solved by the first
// wavefront with shuffles (lane l owns row l of the block); then every thread
// updates its rows outside the block with the block's columns (coalesced:
// consecutive rows of one column).  w lives in LDS when it fits (lds_w), else
// in W.
template <typename T, int NT>
__global__ void __launch_bounds__(NT)
mf_trsv(const T* __restrict__ buf, T* __restrict__ W, const MfFront* __restrict__ fr, int32_t f0, int32_t lds_w) {
  extern __shared__ unsigned char smem[];
  const MfFront d = fr[f0 + blockIdx.x];
  const int s = d.s;
  if (s == 0) return;
  const int tid = threadIdx.x;
  const int64_t m = (int64_t)d.s + d.u;
  const T* A = buf + d.base;
  T* sD = reinterpret_cast<T*>(smem);  // [kDiag][kDiag + 1], column k at k * (kDiag + 1)
  T* w = lds_w ? sD + kDiag * (kDiag + 1) : W + d.w;
  // padded slots start from 0 whatever an earlier solve left there (a
  // failed factorisation's NaN would otherwise reach every later solve
  // through the zero padding columns)
  for (int i = tid; i < s; i += NT) w[i] = i < d.sr ? W[d.w + i] : T(0);
  __syncthreads();
  auto stage = [&](int j0, int jb) {
    for (int t = tid; t < jb * kDiag; t += NT) {
      const int k = t / kDiag, l = t % kDiag;
      if (l < jb) sD[k * (kDiag + 1) + l] = A[(int64_t)(j0 + k) * m + j0 + l];
    }
  };
  // L (unit lower), blocks in order
  for (int j0 = 0; j0 < s; j0 += kDiag) {
    const int jb = min(kDiag, s - j0);
    stage(j0, jb);
    __syncthreads();
    if (tid < kDiag) {
      T val = tid < jb ? w[j0 + tid] : T(0);
      // the block's columns read 8 ahead of the chain (rows past jb clamped
      // into the staged block, their values unused)
      for (int k0 = 0; k0 < jb; k0 += 8) {
        T a[8];
#pragma unroll
        for (int u = 0; u < 8; u++) a[u] = sD[min(k0 + u, kDiag - 1) * (kDiag + 1) + tid];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int k = k0 + u;
          if (k < jb) {  // uniform
            const T wk = lane_value(val, k);
            const T nv = val - a[u] * wk;
            val = (tid > k && tid < jb) ? nv : val;
          }
        }
      }
      if (tid < jb) w[j0 + tid] = val;
    }
    __syncthreads();
    update_rows<T, NT>(w, A + (int64_t)j0 * m, m, j0, jb, j0 + jb, s, tid);
    __syncthreads();
  }
  // U (upper), blocks in reverse
  for (int j0 = ((s - 1) / kDiag) * kDiag; j0 >= 0; j0 -= kDiag) {
    const int jb = min(kDiag, s - j0);
    stage(j0, jb);
    __syncthreads();
    if (tid < kDiag) {
      T val = tid < jb ? w[j0 + tid] : T(0);
      for (int k0 = jb - 1; k0 >= 0; k0 -= 8) {
        T a[8], dg[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int kk = max(k0 - u, 0);
          a[u] = sD[kk * (kDiag + 1) + tid];
          dg[u] = sD[kk * (kDiag + 1) + kk];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int k = k0 - u;
          if (k >= 0) {  // uniform
            // lane k's value over its pivot, formed on every lane from the
            // same two operands (the same bits as lane k's own division)
            const T wk = lane_value(val, k) / dg[u];
            const T nv = val - a[u] * wk;
            val = tid == k ? wk : (tid < k ? nv : val);
          }
        }
      }
      if (tid < jb) w[j0 + tid] = val;
    }
    __syncthreads();
    update_rows<T, NT>(w, A + (int64_t)j0 * m, m, j0, jb, 0, j0, tid);
    __syncthreads();
  }
  if (lds_w) {
    for (int i = tid; i < s; i += NT) W[d.w + i] = w[i];
  }
}

// A chunk of up to kGemvRows rows of one front's product.
struct MfChunk {
  int32_t front, row0;
};
constexpr int kGemvRows = 64;

constexpr int kVecChunk = 4096;  // vector entries staged in LDS at a time

// Forward (BACK = false): V[v + i] = sum_j F21(i, j) W[w + j], i < u, j < s
// (F21 = rows s.., columns 0..s-1).  Backward (BACK = true):
// W[w + i] -= sum_j X(i, j) XU[v + j], i < s, j < u (X = F11^-1 F12 in the
// F12 block).  A workgroup takes kGemvRows rows (a wavefront's 64 lanes,
// consecutive rows of a column: coalesced) times G column groups: group g
// sums its contiguous share of each staged piece of the vector, in column
// order, and the G partial sums of a row are added by a fixed pairwise tree
// (deterministic).  G = 16 for long rows (the top fronts, whose few chunks
// would otherwise run long serial loops), 4 for short ones.
template <typename T, bool BACK, int G>
__global__ void __launch_bounds__(kGemvRows * G)
mf_gemv(const T* __restrict__ buf, T* __restrict__ W, T* __restrict__ V, const T* __restrict__ XU,
        const MfFront* __restrict__ fr, const MfChunk* __restrict__ ch) {
  __shared__ T sv[kVecChunk];
  __shared__ T part[G][kGemvRows];
  const MfChunk c = ch[blockIdx.x];
  const MfFront d = fr[c.front];
  const int64_t m = (int64_t)d.s + d.u;
  const int rows = BACK ? d.s : d.u, cols = BACK ? d.u : d.s;
  const T* A = buf + d.base + (BACK ? (int64_t)d.s * m : (int64_t)d.s);
  const T* x = BACK ? XU + d.v : W + d.w;
  const int r = threadIdx.x % kGemvRows, g = threadIdx.x / kGemvRows;
  const int i = c.row0 + r;
  T acc = T(0);
  for (int j0 = 0; j0 < cols; j0 += kVecChunk) {
    const int jb = min(kVecChunk, cols - j0);
    __syncthreads();
    for (int t = threadIdx.x; t < jb; t += kGemvRows * G) sv[t] = x[j0 + t];
    __syncthreads();
    const int k_lo = jb * g / G, k_hi = jb * (g + 1) / G;
    if (i < rows) {
      const T* col = A + (int64_t)j0 * m + i;
      constexpr int kB = 16;
      for (int k0 = k_lo; k0 < k_hi; k0 += kB) {
        T a[kB];
#pragma unroll
        for (int u = 0; u < kB; u++) a[u] = k0 + u < k_hi ? col[(int64_t)(k0 + u) * m] : T(0);
#pragma unroll
        for (int u = 0; u < kB; u++)
          if (k0 + u < k_hi) acc += a[u] * sv[k0 + u];
      }
    }
  }
  part[g][r] = acc;
  __syncthreads();
  if (g == 0 && i < rows) {
    T v[G];
#pragma unroll
    for (int q = 0; q < G; q++) v[q] = part[q][r];
#pragma unroll
    for (int width = G; width > 1; width /= 2) {
#pragma unroll
      for (int q = 0; q < width / 2; q++) v[q] = v[2 * q] + v[2 * q + 1];
    }
    if (BACK) W[d.w + i] -= v[0];
    else V[d.v + i] = v[0];
  }
}

inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// y = A x for a CSR matrix (FP64), one wavefront per row: lanes stride the
// row, then a fixed shuffle tree sums them (deterministic).  Rows longer than
// kLongRow (the dense arclength row) go to mf_csr_mv_long instead.
constexpr int kLongRow = 1 << 14;
__global__ void __launch_bounds__(256)
mf_csr_mv(int64_t n, const int64_t* __restrict__ rowptr, const int64_t* __restrict__ col,
          const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t row = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (row >= n) return;  // uniform per wavefront
  const int64_t p0 = rowptr[row], p1 = rowptr[row + 1];
  if (p1 - p0 > kLongRow) return;
  double acc = 0.0;
  for (int64_t p = p0 + lane; p < p1; p += 64) acc += val[p] * x[col[p]];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) y[row] = acc;
}
// The long rows (listed by the caller): one 1024-thread workgroup each,
// per-thread sums in a fixed stride, then a fixed tree in LDS.
__global__ void __launch_bounds__(1024)
mf_csr_mv_long(const int64_t* __restrict__ long_rows, const int64_t* __restrict__ rowptr,
               const int64_t* __restrict__ col, const double* __restrict__ val, const double* __restrict__ x,
               double* __restrict__ y) {
  __shared__ double red[1024];
  const int64_t row = long_rows[blockIdx.x];
  const int64_t p0 = rowptr[row], p1 = rowptr[row + 1];
  double acc = 0.0;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 1024) acc += val[p] * x[col[p]];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) y[row] = red[0];
}

// The long rows split over kLongChunks workgroups each (fixed chunk bounds):
// each workgroup sums its chunk by a fixed tree into work[row * kLongChunks +
// chunk], then one workgroup per row sums the chunks by a fixed tree.
constexpr int kLongChunks = 256;
__device__ inline double block_sum_256(double acc) {
  __shared__ double red[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}
__global__ void __launch_bounds__(256)
mf_csr_mv_long_part(const int64_t* __restrict__ long_rows, const int64_t* __restrict__ rowptr,
                    const int64_t* __restrict__ col, const double* __restrict__ val, const double* __restrict__ x,
                    double* __restrict__ work) {
  const int r = blockIdx.x / kLongChunks, c = blockIdx.x % kLongChunks;
  const int64_t row = long_rows[r];
  const int64_t p0 = rowptr[row], len = rowptr[row + 1] - p0;
  const int64_t q0 = p0 + len * c / kLongChunks, q1 = p0 + len * (c + 1) / kLongChunks;
  double acc = 0.0;
  for (int64_t p = q0 + threadIdx.x; p < q1; p += 256) acc += val[p] * x[col[p]];
  acc = block_sum_256(acc);
  if (threadIdx.x == 0) work[blockIdx.x] = acc;
}
__global__ void __launch_bounds__(256)
mf_csr_mv_long_sum(const int64_t* __restrict__ long_rows, const double* __restrict__ work, double* __restrict__ y) {
  const double acc = block_sum_256(work[(int64_t)blockIdx.x * kLongChunks + threadIdx.x]);
  if (threadIdx.x == 0) y[long_rows[blockIdx.x]] = acc;
}

// rocSOLVER / rocBLAS by precision
rocblas_status getrf_sb(rocblas_handle h, int n, double* A, int lda, rocblas_stride sA, int* ipiv, rocblas_stride sP,
                        int* info, int bc) {
  return rocsolver_dgetrf_strided_batched(h, n, n, A, lda, sA, ipiv, sP, info, bc);
}
rocblas_status getrf_sb(rocblas_handle h, int n, float* A, int lda, rocblas_stride sA, int* ipiv, rocblas_stride sP,
                        int* info, int bc) {
  return rocsolver_sgetrf_strided_batched(h, n, n, A, lda, sA, ipiv, sP, info, bc);
}
rocblas_status getrf_npvt_sb(rocblas_handle h, int n, double* A, int lda, rocblas_stride sA, int* info, int bc) {
  return rocsolver_dgetrf_npvt_strided_batched(h, n, n, A, lda, sA, info, bc);
}
rocblas_status getrf_npvt_sb(rocblas_handle h, int n, float* A, int lda, rocblas_stride sA, int* info, int bc) {
  return rocsolver_sgetrf_npvt_strided_batched(h, n, n, A, lda, sA, info, bc);
}
rocblas_status getrs_sb(rocblas_handle h, int n, int nrhs, double* A, int lda, rocblas_stride sA, const int* ipiv,
                        rocblas_stride sP, double* B, int ldb, rocblas_stride sB, int bc) {
  return rocsolver_dgetrs_strided_batched(h, rocblas_operation_none, n, nrhs, A, lda, sA, ipiv, sP, B, ldb, sB, bc);
}
rocblas_status getrs_sb(rocblas_handle h, int n, int nrhs, float* A, int lda, rocblas_stride sA, const int* ipiv,
                        rocblas_stride sP, float* B, int ldb, rocblas_stride sB, int bc) {
  return rocsolver_sgetrs_strided_batched(h, rocblas_operation_none, n, nrhs, A, lda, sA, ipiv, sP, B, ldb, sB, bc);
}
rocblas_status gemm_sb(rocblas_handle h, int m, int n, int k, const double* alpha, const double* A, int lda,
                       rocblas_stride sA, const double* B, int ldb, rocblas_stride sB, const double* beta, double* C,
                       int ldc, rocblas_stride sC, int bc) {
  return rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m, n, k, alpha, A, lda, sA,
                                       B, ldb, sB, beta, C, ldc, sC, bc);
}
rocblas_status gemm_sb(rocblas_handle h, int m, int n, int k, const float* alpha, const float* A, int lda,
                       rocblas_stride sA, const float* B, int ldb, rocblas_stride sB, const float* beta, float* C,
                       int ldc, rocblas_stride sC, int bc) {
  return rocblas_sgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m, n, k, alpha, A, lda, sA,
                                       B, ldb, sB, beta, C, ldc, sC, bc);
}

}  // namespace

struct vef_mf {
  int bits = 64;
  int pivoting = 1;           // 0: getrf without row interchanges (vef_mf_set_pivoting)
  int* ipiv_id = nullptr;     // the identity interchanges, for getrs and the gather without pivoting
  int64_t n = 0;
  int32_t nb = 0, nf = 0, H = 0;
  std::vector<int32_t> b_h, b_c, b_s, b_u;
  std::vector<int64_t> b_base, b_first, b_ipiv, b_w, b_v;  // per bucket: storage, first front, ipiv/W/V offsets
  std::vector<int64_t> f_base;
  int64_t total = 0, n_ipiv = 0, n_w = 0, n_v = 0;
  // device
  void* buf = nullptr;
  int* ipiv = nullptr;
  int* info = nullptr;
  void *W = nullptr, *V = nullptr, *XU = nullptr;
  int64_t* d_pad = nullptr;
  int64_t n_pad = 0;
  EaTask* d_ea = nullptr;
  int32_t* d_map = nullptr;
  std::vector<int64_t> ea_off;    // [2H + 1] tasks of (height, slot)
  std::vector<int32_t> ea_maxu;   // [2H]
  int64_t *d_fg_dst = nullptr, *d_bg_dst = nullptr, *d_bs_src = nullptr, *d_fs_ptr = nullptr, *d_fs_v = nullptr;
  int32_t *d_bg_src = nullptr, *d_bs_dst = nullptr, *d_fs_tgt = nullptr;
  std::vector<int64_t> fg_off, bg_off, bs_off, fs_off, fs_ptr_off;  // per height
  // solve kernels: front table, per-height front ranges and product chunks
  MfFront* d_fr = nullptr;
  int64_t* d_fipiv = nullptr;  // ipiv offset of each front
  int32_t* d_prow = nullptr;   // [n_w] pivot unknown of each W slot (-1: padding)
  int32_t* d_psrc = nullptr;   // [n_w] the same after the row interchanges (mf_perm)
  MfChunk *d_fch = nullptr, *d_bch = nullptr;
  std::vector<int32_t> h_f0, h_nf, h_smax;  // per height: first front, fronts, largest padded s
  std::vector<int32_t> h_umax;              // per height: largest padded u
  std::vector<int64_t> fch_off, bch_off;    // per height: forward / backward chunks
  int32_t s_max = 0;
  rocblas_handle rb = nullptr;
  // Buckets of one height are independent: they run on kSide internal
  // streams (default priority), each with its own rocBLAS handle, forked from
  // and joined back to the caller's stream by events around each height.
  static constexpr int kSide = 8;
  hipStream_t side[kSide] = {};
  rocblas_handle rbs[kSide] = {};
  hipEvent_t ev_main = nullptr;
  hipEvent_t ev_side[kSide] = {};
  std::vector<int> b_stream;  // side stream of each bucket (balanced by flops within its height)
  // vef_mf_profile: events on the caller's stream at each height's start,
  // after its extend-add and after its buckets (2H + 1 per factorisation)
  bool profiling = false;
  std::vector<hipEvent_t> prof_ev;
  int64_t bytes = 0;
  std::string err;
};

static int fail(vef_mf* h, const std::string& m) {
  if (h) h->err = m;
  return -1;
}

template <typename U>
static int upload(vef_mf* h, U** d, const std::vector<U>& v) {
  *d = nullptr;
  if (v.empty()) return 0;
  if (hipMalloc((void**)d, v.size() * sizeof(U)) != hipSuccess) return fail(h, "hipMalloc failed (index arrays)");
  h->bytes += v.size() * sizeof(U);
  if (hipMemcpy(*d, v.data(), v.size() * sizeof(U), hipMemcpyHostToDevice) != hipSuccess)
    return fail(h, "hipMemcpy failed (index arrays)");
  return 0;
}

static void release(vef_mf* h) {
  void* ptrs[] = {h->buf,      h->ipiv,     h->info,     h->W,        h->V,        h->XU,
                  h->d_pad,    h->d_ea,     h->d_map,    h->d_fg_dst, h->d_bg_dst, h->d_bs_src,
                  h->d_fs_ptr, h->d_fs_v,   h->d_bg_src, h->d_bs_dst, h->d_fs_tgt,
                  h->d_fr,     h->d_fipiv,  h->d_prow,   h->d_psrc,   h->d_fch,    h->d_bch,
                  h->ipiv_id};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (h->rb) rocblas_destroy_handle(h->rb);
  h->rb = nullptr;
  for (hipEvent_t e : h->prof_ev) (void)hipEventDestroy(e);
  h->prof_ev.clear();
  for (int k = 0; k < vef_mf::kSide; k++) {
    if (h->rbs[k]) rocblas_destroy_handle(h->rbs[k]);
    if (h->ev_side[k]) (void)hipEventDestroy(h->ev_side[k]);
    if (h->side[k]) (void)hipStreamDestroy(h->side[k]);
    h->rbs[k] = nullptr;
    h->ev_side[k] = nullptr;
    h->side[k] = nullptr;
  }
  if (h->ev_main) (void)hipEventDestroy(h->ev_main);
  h->ev_main = nullptr;
}

// The caller's stream st -> every side stream (at the start of a height).
static int fork_streams(vef_mf* h, hipStream_t st) {
  if (hipEventRecord(h->ev_main, st) != hipSuccess) return fail(h, "hipEventRecord failed");
  for (int k = 0; k < vef_mf::kSide; k++)
    if (hipStreamWaitEvent(h->side[k], h->ev_main, 0) != hipSuccess) return fail(h, "hipStreamWaitEvent failed");
  return 0;
}

// Every side stream -> the caller's stream (at the end of a height).
static int join_streams(vef_mf* h, hipStream_t st) {
  for (int k = 0; k < vef_mf::kSide; k++) {
    if (hipEventRecord(h->ev_side[k], h->side[k]) != hipSuccess) return fail(h, "hipEventRecord failed");
    if (hipStreamWaitEvent(st, h->ev_side[k], 0) != hipSuccess) return fail(h, "hipStreamWaitEvent failed");
  }
  return 0;
}

// The exported functions take C linkage from their declarations in vef_mf.h.

int vef_mf_create(const vef_mf_plan* p, int32_t bits, vef_mf** out) {
  if (!out) return -1;
  *out = nullptr;
  if (!p || (bits != 64 && bits != 32)) return -1;
  vef_mf* h = new vef_mf();
  h->bits = bits;
  h->n = p->n;
  h->nb = p->n_buckets;
  h->nf = p->n_fronts;
  const size_t es = bits == 64 ? sizeof(double) : sizeof(float);
  auto bad = [&](const std::string& m) {
    fail(h, m);
    *out = h;
    return -1;
  };
  if (h->n <= 0 || h->n > INT32_MAX || h->nb <= 0 || h->nf <= 0) return bad("plan: empty or too large");
  // buckets
  h->b_h.assign(p->bucket_height, p->bucket_height + h->nb);
  h->b_c.assign(p->bucket_count, p->bucket_count + h->nb);
  h->b_s.assign(p->bucket_s, p->bucket_s + h->nb);
  h->b_u.assign(p->bucket_u, p->bucket_u + h->nb);
  int64_t nfr = 0;
  for (int b = 0; b < h->nb; b++) {
    if (h->b_c[b] <= 0 || h->b_s[b] < 0 || h->b_u[b] < 0 || h->b_h[b] < 0) return bad("plan: bad bucket");
    if (b > 0 && h->b_h[b] < h->b_h[b - 1]) return bad("plan: buckets not ordered by height");
    if ((int64_t)h->b_s[b] + h->b_u[b] > 46340) return bad("plan: front too large");
    if (h->b_s[b] > 16384) return bad("plan: more than 16384 pivots in a front");
    nfr += h->b_c[b];
  }
  if (nfr != h->nf) return bad("plan: bucket counts do not add up to n_fronts");
  h->H = h->b_h[h->nb - 1] + 1;
  h->b_base.resize(h->nb + 1);
  h->b_first.resize(h->nb + 1);
  h->b_ipiv.resize(h->nb + 1);
  h->b_w.resize(h->nb + 1);
  h->b_v.resize(h->nb + 1);
  h->b_base[0] = h->b_first[0] = h->b_ipiv[0] = h->b_w[0] = h->b_v[0] = 0;
  for (int b = 0; b < h->nb; b++) {
    int64_t m = (int64_t)h->b_s[b] + h->b_u[b];
    h->b_base[b + 1] = h->b_base[b] + (int64_t)h->b_c[b] * m * m;
    h->b_first[b + 1] = h->b_first[b] + h->b_c[b];
    h->b_ipiv[b + 1] = h->b_ipiv[b] + (int64_t)h->b_c[b] * h->b_s[b];
    h->b_w[b + 1] = h->b_w[b] + (int64_t)h->b_c[b] * h->b_s[b];
    h->b_v[b + 1] = h->b_v[b] + (int64_t)h->b_c[b] * h->b_u[b];
  }
  h->total = h->b_base[h->nb];
  h->n_ipiv = h->b_ipiv[h->nb];
  h->n_w = h->b_w[h->nb];
  h->n_v = h->b_v[h->nb];
  // fronts: bucket, base, padded sizes, W/V offsets
  std::vector<int32_t> fb(h->nf);
  std::vector<int64_t> fw(h->nf), fv(h->nf);
  h->f_base.resize(h->nf);
  for (int b = 0; b < h->nb; b++) {
    int64_t m = (int64_t)h->b_s[b] + h->b_u[b];
    for (int k = 0; k < h->b_c[b]; k++) {
      int64_t f = h->b_first[b] + k;
      fb[f] = b;
      h->f_base[f] = h->b_base[b] + k * m * m;
      fw[f] = h->b_w[b] + (int64_t)k * h->b_s[b];
      fv[f] = h->b_v[b] + (int64_t)k * h->b_u[b];
    }
  }
  auto spad = [&](int64_t f) { return (int64_t)h->b_s[fb[f]]; };
  auto mpad = [&](int64_t f) { return (int64_t)h->b_s[fb[f]] + h->b_u[fb[f]]; };
  std::vector<int> nchild(h->nf, 0), slotmask(h->nf, 0);
  for (int64_t f = 0; f < h->nf; f++) {
    int b = fb[f];
    int32_t s = p->front_s[f], u = p->front_u[f];
    if (s < 0 || u < 0 || s > h->b_s[b] || u > h->b_u[b]) return bad("plan: front larger than its bucket");
    if (p->rows_off[f + 1] - p->rows_off[f] != (int64_t)s + u) return bad("plan: row list length");
    for (int64_t k = p->rows_off[f]; k < p->rows_off[f + 1]; k++)
      if (p->rows[k] < 0 || p->rows[k] >= h->n) return bad("plan: row out of range");
    int32_t par = p->parent[f];
    if (par >= 0) {
      if (par >= h->nf || h->b_h[fb[par]] <= h->b_h[b]) return bad("plan: parent not above its child");
      if (p->slot[f] != 0 && p->slot[f] != 1) return bad("plan: child slot must be 0 or 1");
      if (slotmask[par] & (1 << p->slot[f])) return bad("plan: two children in one slot");
      slotmask[par] |= 1 << p->slot[f];
      if (p->ea_off[f + 1] - p->ea_off[f] != u) return bad("plan: extend-add map length");
      for (int64_t k = p->ea_off[f]; k < p->ea_off[f + 1]; k++)
        if (p->ea_map[k] < 0 || p->ea_map[k] >= mpad(par)) return bad("plan: extend-add row out of range");
    } else if (u != 0) {
      return bad("plan: a root front has update rows");
    }
  }
  // every unknown is a pivot row of exactly one front
  {
    std::vector<char> seen(h->n, 0);
    for (int64_t f = 0; f < h->nf; f++)
      for (int64_t k = p->rows_off[f]; k < p->rows_off[f] + p->front_s[f]; k++) {
        if (seen[p->rows[k]]) return bad("plan: an unknown is a pivot of two fronts");
        seen[p->rows[k]] = 1;
      }
    for (int64_t i = 0; i < h->n; i++)
      if (!seen[i]) return bad("plan: an unknown is a pivot of no front");
  }
  // device storage
  auto dmalloc = [&](void** d, int64_t bytes) -> bool {
    *d = nullptr;
    if (bytes <= 0) return true;
    if (hipMalloc(d, (size_t)bytes) != hipSuccess) return false;
    h->bytes += bytes;
    return true;
  };
  if (!dmalloc(&h->buf, h->total * (int64_t)es) || !dmalloc((void**)&h->ipiv, h->n_ipiv * 4) ||
      !dmalloc((void**)&h->info, (int64_t)h->nf * 4) || !dmalloc(&h->W, h->n_w * (int64_t)es) ||
      !dmalloc(&h->V, h->n_v * (int64_t)es) || !dmalloc(&h->XU, h->n_v * (int64_t)es))
    return bad("hipMalloc failed (front storage)");
  // padded entries of the per-front vectors stay zero (only real rows are gathered)
  if (h->n_w > 0 && hipMemset(h->W, 0, h->n_w * es) != hipSuccess) return bad("hipMemset failed");
  if (h->n_v > 0 && hipMemset(h->XU, 0, h->n_v * es) != hipSuccess) return bad("hipMemset failed");
  if (h->info && hipMemset(h->info, 0, (size_t)h->nf * 4) != hipSuccess) return bad("hipMemset failed");
  // padded pivot diagonal
  std::vector<int64_t> pad;
  for (int64_t f = 0; f < h->nf; f++)
    for (int64_t i = p->front_s[f]; i < spad(f); i++) pad.push_back(h->f_base[f] + i + i * mpad(f));
  h->n_pad = (int64_t)pad.size();
  if (upload(h, &h->d_pad, pad)) return bad(h->err);
  // extend-add tasks per (height, slot), and the maps
  std::vector<EaTask> tasks;
  h->ea_off.assign(2 * h->H + 1, 0);
  h->ea_maxu.assign(2 * h->H, 0);
  std::vector<std::vector<EaTask>> per(2 * h->H);
  for (int64_t f = 0; f < h->nf; f++) {
    int32_t par = p->parent[f];
    if (par < 0 || p->front_u[f] == 0) continue;
    int k = 2 * h->b_h[fb[par]] + p->slot[f];
    EaTask t;
    t.cbase = h->f_base[f];
    t.pbase = h->f_base[par];
    t.moff = p->ea_off[f];
    t.cm = (int32_t)mpad(f);
    t.cs = (int32_t)spad(f);
    t.pm = (int32_t)mpad(par);
    t.u = p->front_u[f];
    per[k].push_back(t);
    h->ea_maxu[k] = std::max(h->ea_maxu[k], t.u);
  }
  for (int k = 0; k < 2 * h->H; k++) {
    tasks.insert(tasks.end(), per[k].begin(), per[k].end());
    h->ea_off[k + 1] = (int64_t)tasks.size();
  }
  if (upload(h, &h->d_ea, tasks)) return bad(h->err);
  std::vector<int32_t> map(p->ea_map, p->ea_map + p->ea_off[h->nf]);
  if (upload(h, &h->d_map, map)) return bad(h->err);
  // solve lists per height
  std::vector<int64_t> fg_dst, bg_dst, bs_src, fs_ptr, fs_v;
  std::vector<int32_t> bg_src, bs_dst, fs_tgt;
  h->fg_off.assign(h->H + 1, 0);
  h->bg_off.assign(h->H + 1, 0);
  h->bs_off.assign(h->H + 1, 0);
  h->fs_off.assign(h->H + 1, 0);
  h->fs_ptr_off.assign(h->H + 1, 0);
  for (int hh = 0; hh < h->H; hh++) {
    std::vector<std::pair<int32_t, int64_t>> contrib;  // (target unknown, V index), fronts in order
    for (int b = 0; b < h->nb; b++) {
      if (h->b_h[b] != hh) continue;
      for (int64_t f = h->b_first[b]; f < h->b_first[b + 1]; f++) {
        const int32_t* rw = p->rows + p->rows_off[f];
        int32_t s = p->front_s[f], u = p->front_u[f];
        for (int32_t i = 0; i < s; i++) {
          fg_dst.push_back(fw[f] + i);
          bs_dst.push_back(rw[i]);
          bs_src.push_back(fw[f] + i);
        }
        if (h->b_s[b] == 0) continue;  // no pivots in the bucket: no update contributions
        for (int32_t i = 0; i < u; i++) {
          bg_dst.push_back(fv[f] + i);
          bg_src.push_back(rw[s + i]);
          contrib.push_back({rw[s + i], fv[f] + i});
        }
      }
    }
    std::stable_sort(contrib.begin(), contrib.end(),
                     [](const std::pair<int32_t, int64_t>& a, const std::pair<int32_t, int64_t>& c) {
                       return a.first < c.first;
                     });
    int64_t start = (int64_t)fs_v.size();
    for (size_t k = 0; k < contrib.size(); k++) {
      if (k == 0 || contrib[k].first != contrib[k - 1].first) {
        fs_tgt.push_back(contrib[k].first);
        fs_ptr.push_back(start + (int64_t)k);
      }
      fs_v.push_back(contrib[k].second);
    }
    fs_ptr.push_back((int64_t)fs_v.size());  // closes this height's last target
    h->fg_off[hh + 1] = (int64_t)fg_dst.size();
    h->bg_off[hh + 1] = (int64_t)bg_dst.size();
    h->bs_off[hh + 1] = (int64_t)bs_dst.size();
    h->fs_off[hh + 1] = (int64_t)fs_tgt.size();
    h->fs_ptr_off[hh + 1] = (int64_t)fs_ptr.size();
  }
  if (upload(h, &h->d_fg_dst, fg_dst) || upload(h, &h->d_bg_dst, bg_dst) ||
      upload(h, &h->d_bg_src, bg_src) || upload(h, &h->d_bs_dst, bs_dst) || upload(h, &h->d_bs_src, bs_src) ||
      upload(h, &h->d_fs_tgt, fs_tgt) || upload(h, &h->d_fs_ptr, fs_ptr) || upload(h, &h->d_fs_v, fs_v))
    return bad(h->err);
  // solve kernels' tables: fronts (in plan order, so a height's fronts are
  // consecutive), their ipiv offsets and pivot unknowns, and the row chunks
  // of the two products per height
  {
    std::vector<MfFront> fr(h->nf);
    std::vector<int64_t> fipiv(h->nf);
    std::vector<int32_t> prow(std::max<int64_t>(h->n_w, 1), -1);
    std::vector<MfChunk> fch, bch;
    h->h_f0.assign(h->H, 0);
    h->h_nf.assign(h->H, 0);
    h->h_smax.assign(h->H, 0);
    h->h_umax.assign(h->H, 0);
    h->fch_off.assign(h->H + 1, 0);
    h->bch_off.assign(h->H + 1, 0);
    for (int64_t f = 0; f < h->nf; f++) {
      const int b = fb[f];
      const int64_t k = f - h->b_first[b];
      fr[f] = MfFront{h->f_base[f], fw[f], fv[f], h->b_s[b], h->b_u[b], p->front_s[f]};
      fipiv[f] = h->b_ipiv[b] + k * h->b_s[b];
      const int32_t* rw = p->rows + p->rows_off[f];
      for (int32_t i = 0; i < p->front_s[f]; i++) prow[fw[f] + i] = rw[i];
      h->s_max = std::max(h->s_max, h->b_s[b]);
    }
    for (int hh = 0, b = 0; hh < h->H; hh++) {
      h->h_f0[hh] = (int32_t)h->b_first[b];
      for (; b < h->nb && h->b_h[b] == hh; b++) {
        h->h_smax[hh] = std::max(h->h_smax[hh], h->b_s[b]);
        if (h->b_s[b] == 0 || h->b_u[b] == 0) continue;
        h->h_umax[hh] = std::max(h->h_umax[hh], h->b_u[b]);
        for (int64_t f = h->b_first[b]; f < h->b_first[b + 1]; f++) {
          for (int32_t r0 = 0; r0 < h->b_u[b]; r0 += kGemvRows) fch.push_back(MfChunk{(int32_t)f, r0});
          for (int32_t r0 = 0; r0 < h->b_s[b]; r0 += kGemvRows) bch.push_back(MfChunk{(int32_t)f, r0});
        }
      }
      h->h_nf[hh] = (int32_t)(h->b_first[b] - h->h_f0[hh]);
      h->fch_off[hh + 1] = (int64_t)fch.size();
      h->bch_off[hh + 1] = (int64_t)bch.size();
    }
    if (upload(h, &h->d_fr, fr) || upload(h, &h->d_fipiv, fipiv) || upload(h, &h->d_prow, prow) ||
        upload(h, &h->d_psrc, prow) || upload(h, &h->d_fch, fch) || upload(h, &h->d_bch, bch))
      return bad(h->err);
    std::vector<int32_t> id(std::max<int64_t>(h->n_ipiv, 1));
    for (int b = 0; b < h->nb; b++)
      for (int64_t k = 0; k < h->b_c[b]; k++)
        for (int32_t i = 0; i < h->b_s[b]; i++) id[h->b_ipiv[b] + k * h->b_s[b] + i] = i + 1;
    if (upload(h, &h->ipiv_id, id)) return bad(h->err);
  }
  if (rocblas_create_handle(&h->rb) != rocblas_status_success) return bad("rocblas_create_handle failed");
  // no atomic reductions inside rocBLAS/rocSOLVER: repeated factorisations and
  // solves of the same values give the same bits
  if (rocblas_set_atomics_mode(h->rb, rocblas_atomics_not_allowed) != rocblas_status_success)
    return bad("rocblas_set_atomics_mode failed");
  if (hipEventCreateWithFlags(&h->ev_main, hipEventDisableTiming) != hipSuccess) return bad("hipEventCreate failed");
  for (int k = 0; k < vef_mf::kSide; k++) {
    if (hipStreamCreateWithFlags(&h->side[k], hipStreamNonBlocking) != hipSuccess) return bad("hipStreamCreate failed");
    if (hipEventCreateWithFlags(&h->ev_side[k], hipEventDisableTiming) != hipSuccess)
      return bad("hipEventCreate failed");
    if (rocblas_create_handle(&h->rbs[k]) != rocblas_status_success ||
        rocblas_set_stream(h->rbs[k], h->side[k]) != rocblas_status_success ||
        rocblas_set_atomics_mode(h->rbs[k], rocblas_atomics_not_allowed) != rocblas_status_success)
      return bad("rocblas handle of a side stream failed");
  }
  // buckets of each height onto the side streams, largest first to the least loaded
  h->b_stream.assign(h->nb, 0);
  for (int b0 = 0; b0 < h->nb;) {
    int b1 = b0;
    while (b1 < h->nb && h->b_h[b1] == h->b_h[b0]) b1++;
    std::vector<std::pair<double, int>> cost;
    for (int b = b0; b < b1; b++) {
      double s = h->b_s[b], u = h->b_u[b];
      cost.push_back({h->b_c[b] * (2.0 / 3.0 * s * s * s + 2.0 * s * s * u + 2.0 * s * u * u + 1e5), b});
    }
    std::sort(cost.begin(), cost.end(), [](const std::pair<double, int>& a, const std::pair<double, int>& c) {
      return a.first > c.first || (a.first == c.first && a.second < c.second);
    });
    double load[vef_mf::kSide] = {};
    for (auto& cb : cost) {
      int k = 0;
      for (int j = 1; j < vef_mf::kSide; j++)
        if (load[j] < load[k]) k = j;
      h->b_stream[cb.second] = k;
      load[k] += cb.first;
    }
    b0 = b1;
  }
  *out = h;
  return 0;
}

int vef_mf_layout(const vef_mf* h, int64_t* front_base, int64_t* total) {
  if (!h) return -1;
  if (front_base) std::memcpy(front_base, h->f_base.data(), sizeof(int64_t) * h->f_base.size());
  if (total) *total = h->total;
  return 0;
}

int64_t vef_mf_device_bytes(const vef_mf* h) { return h ? h->bytes : 0; }

template <typename T>
static int set_t(vef_mf* h, int64_t count, const int64_t* dst, const double* values, hipStream_t st) {
  T* buf = (T*)h->buf;
  mf_fill<T><<<4096, 256, 0, st>>>(buf, h->total, T(0));
  if (h->n_pad) mf_set_pos<T><<<blocks_for(h->n_pad, 256), 256, 0, st>>>(buf, h->d_pad, h->n_pad, T(1));
  if (count) mf_scatter<T><<<blocks_for(count, 256), 256, 0, st>>>(buf, dst, values, count);
  return hipGetLastError() == hipSuccess ? 0 : fail(h, "kernel launch failed (set)");
}

int vef_mf_set(vef_mf* h, int64_t count, const int64_t* dst, const double* values, void* stream) {
  if (!h || !h->buf) return -1;
  if (count < 0 || (count > 0 && (!dst || !values))) return fail(h, "vef_mf_set: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  return h->bits == 64 ? set_t<double>(h, count, dst, values, st) : set_t<float>(h, count, dst, values, st);
}

template <typename T>
static int factor_t(vef_mf* h, hipStream_t st) {
  T* buf = (T*)h->buf;
  const T one = T(1), mone = T(-1);
  int b = 0;
  // the solves' per-front vectors start from zero with every factorisation
  // (their padded slots must be 0; nothing of an earlier, possibly failed,
  // solve is carried over)
  if (h->n_w) mf_fill<T><<<blocks_for(h->n_w, 256), 256, 0, st>>>((T*)h->W, h->n_w, T(0));
  if (h->n_v) {
    mf_fill<T><<<blocks_for(h->n_v, 256), 256, 0, st>>>((T*)h->V, h->n_v, T(0));
    mf_fill<T><<<blocks_for(h->n_v, 256), 256, 0, st>>>((T*)h->XU, h->n_v, T(0));
  }
  const bool prof = h->profiling && (int)h->prof_ev.size() == 2 * h->H + 1;
  if (prof && hipEventRecord(h->prof_ev[0], st) != hipSuccess) return fail(h, "hipEventRecord failed (profile)");
  for (int hh = 0; hh < h->H; hh++) {
    for (int sl = 0; sl < 2; sl++) {
      int k = 2 * hh + sl;
      int64_t nt = h->ea_off[k + 1] - h->ea_off[k];
      if (nt > 0) {
        dim3 grid((unsigned)nt, (unsigned)h->ea_maxu[k]);
        mf_extend_add<T><<<grid, 128, 0, st>>>(buf, h->d_ea + h->ea_off[k], h->d_map);
      }
    }
    if (hipGetLastError() != hipSuccess) return fail(h, "kernel launch failed (extend-add)");
    if (prof && hipEventRecord(h->prof_ev[2 * hh + 1], st) != hipSuccess)
      return fail(h, "hipEventRecord failed (profile)");
    if (fork_streams(h, st)) return -1;
    for (; b < h->nb && h->b_h[b] == hh; b++) {
      int s = h->b_s[b], u = h->b_u[b], c = h->b_c[b];
      if (s == 0) continue;
      rocblas_handle rb = h->rbs[h->b_stream[b]];
      int m = s + u;
      rocblas_stride sA = (rocblas_stride)m * m;
      T* A = buf + h->b_base[b];
      int* piv = (h->pivoting ? h->ipiv : h->ipiv_id) + h->b_ipiv[b];
      if (h->pivoting) {
        if (getrf_sb(rb, s, A, m, sA, piv, s, h->info + h->b_first[b], c) != rocblas_status_success)
          return fail(h, "getrf failed");
      } else if (getrf_npvt_sb(rb, s, A, m, sA, h->info + h->b_first[b], c) != rocblas_status_success) {
        return fail(h, "getrf_npvt failed");
      }
      if (u == 0) continue;
      if (getrs_sb(rb, s, u, A, m, sA, piv, s, A + (int64_t)s * m, m, sA, c) !=
          rocblas_status_success)
        return fail(h, "getrs failed");
      if (gemm_sb(rb, u, u, s, &mone, A + s, m, sA, A + (int64_t)s * m, m, sA, &one, A + (int64_t)s * m + s, m, sA, c) !=
          rocblas_status_success)
        return fail(h, "gemm failed");
    }
    if (join_streams(h, st)) return -1;
    if (prof && hipEventRecord(h->prof_ev[2 * hh + 2], st) != hipSuccess)
      return fail(h, "hipEventRecord failed (profile)");
  }
  // the row interchanges as permutations of the solve's gather
  mf_perm<<<(unsigned)h->nf, kSolveThreads, (size_t)std::max(h->s_max, 1) * sizeof(int32_t), st>>>(
      h->d_fr, h->pivoting ? h->ipiv : h->ipiv_id, h->d_fipiv, h->d_prow, h->d_psrc);
  if (hipGetLastError() != hipSuccess) return fail(h, "kernel launch failed (pivot permutation)");
  return 0;
}

int vef_mf_profile(vef_mf* h, int32_t enable) {
  if (!h) return -1;
  if (enable && h->prof_ev.empty()) {
    h->prof_ev.assign(2 * h->H + 1, nullptr);
    for (auto& e : h->prof_ev)
      if (hipEventCreate(&e) != hipSuccess) return fail(h, "hipEventCreate failed (profile)");
  }
  h->profiling = enable != 0;
  return 0;
}

int32_t vef_mf_heights(const vef_mf* h) { return h ? h->H : -1; }

int vef_mf_set_pivoting(vef_mf* h, int32_t on) {
  if (!h) return -1;
  h->pivoting = on != 0;
  return 0;
}

int vef_mf_profile_read(vef_mf* h, float* ms) {
  if (!h || !ms) return -1;
  if (h->prof_ev.size() != (size_t)(2 * h->H + 1)) return fail(h, "vef_mf_profile_read: profiling was never enabled");
  if (hipEventSynchronize(h->prof_ev.back()) != hipSuccess) return fail(h, "hipEventSynchronize failed (profile)");
  for (int k = 0; k < 2 * h->H; k++)
    if (hipEventElapsedTime(&ms[k], h->prof_ev[k], h->prof_ev[k + 1]) != hipSuccess)
      return fail(h, "hipEventElapsedTime failed (profile)");
  return 0;
}

int vef_mf_factor(vef_mf* h, void* stream) {
  if (!h || !h->buf) return -1;
  hipStream_t st = (hipStream_t)stream;
  return h->bits == 64 ? factor_t<double>(h, st) : factor_t<float>(h, st);
}

int vef_mf_zero_pivots(vef_mf* h, int32_t* count) {
  if (!h || !count) return -1;
  std::vector<int32_t> info(h->nf);
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(info.data(), h->info, sizeof(int32_t) * h->nf, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(h, "hipMemcpy failed (info)");
  int32_t z = 0;
  for (int32_t v : info) z += v > 0;
  *count = z;
  return 0;
}

template <typename T>
static int solve_t(vef_mf* h, double* r, double* x, hipStream_t st) {
  T* buf = (T*)h->buf;
  T* W = (T*)h->W;
  T* V = (T*)h->V;
  T* XU = (T*)h->XU;
  const size_t diag_bytes = (size_t)kDiag * (kDiag + 1) * sizeof(T);
  // forward: w = U^-1 L^-1 P b_s per front, then the update rows' b -= F21 w
  for (int hh = 0; hh < h->H; hh++) {
    const int64_t n = h->fg_off[hh + 1] - h->fg_off[hh];
    if (n) mf_gather_p<T><<<blocks_for(n, 256), 256, 0, st>>>(W, h->d_fg_dst + h->fg_off[hh], h->d_psrc, r, n);
    if (h->h_nf[hh] > 0 && h->h_smax[hh] > 0) {
      const size_t wb = (size_t)h->h_smax[hh] * sizeof(T);
      const int lds_w = diag_bytes + wb <= 65536;
      const size_t shm = diag_bytes + (lds_w ? wb : 0);
      // large fronts: a whole CU's worth of threads per front for the loads
      // in flight; small ones: more fronts per CU
      if (h->h_smax[hh] > 4 * kSolveThreads)
        mf_trsv<T, 1024><<<(unsigned)h->h_nf[hh], 1024, shm, st>>>(buf, W, h->d_fr, h->h_f0[hh], lds_w);
      else
        mf_trsv<T, kSolveThreads><<<(unsigned)h->h_nf[hh], kSolveThreads, shm, st>>>(buf, W, h->d_fr, h->h_f0[hh],
                                                                                   lds_w);
    }
    const int64_t nc = h->fch_off[hh + 1] - h->fch_off[hh];
    if (nc && h->h_smax[hh] > 256)
      mf_gemv<T, false, 16><<<(unsigned)nc, kGemvRows * 16, 0, st>>>(buf, W, V, XU, h->d_fr, h->d_fch + h->fch_off[hh]);
    else if (nc)
      mf_gemv<T, false, 4><<<(unsigned)nc, kGemvRows * 4, 0, st>>>(buf, W, V, XU, h->d_fr, h->d_fch + h->fch_off[hh]);
    const int64_t nt = h->fs_off[hh + 1] - h->fs_off[hh];
    if (nt)
      mf_subtract<T><<<blocks_for(nt, 256), 256, 0, st>>>(r, h->d_fs_tgt + h->fs_off[hh], h->d_fs_ptr + h->fs_ptr_off[hh],
                                                          h->d_fs_v, V, nt);
    if (hipGetLastError() != hipSuccess) return fail(h, "kernel launch failed (forward solve)");
  }
  // backward: x_s = w - X x_u per front
  for (int hh = h->H - 1; hh >= 0; hh--) {
    const int64_t n = h->bg_off[hh + 1] - h->bg_off[hh];
    if (n)
      mf_gather<T><<<blocks_for(n, 256), 256, 0, st>>>(XU, h->d_bg_dst + h->bg_off[hh], h->d_bg_src + h->bg_off[hh], x,
                                                      n);
    const int64_t nc = h->bch_off[hh + 1] - h->bch_off[hh];
    if (nc && h->h_umax[hh] > 256)
      mf_gemv<T, true, 16><<<(unsigned)nc, kGemvRows * 16, 0, st>>>(buf, W, V, XU, h->d_fr, h->d_bch + h->bch_off[hh]);
    else if (nc)
      mf_gemv<T, true, 4><<<(unsigned)nc, kGemvRows * 4, 0, st>>>(buf, W, V, XU, h->d_fr, h->d_bch + h->bch_off[hh]);
    const int64_t np = h->bs_off[hh + 1] - h->bs_off[hh];
    if (np)
      mf_put<T><<<blocks_for(np, 256), 256, 0, st>>>(x, h->d_bs_dst + h->bs_off[hh], h->d_bs_src + h->bs_off[hh], W,
                                                    np);
    if (hipGetLastError() != hipSuccess) return fail(h, "kernel launch failed (backward solve)");
  }
  return 0;
}

int vef_mf_csr_mv(int64_t n, const int64_t* rowptr, const int64_t* col, const double* values, const double* x,
                  double* y, int32_t n_long, const int64_t* long_rows, double* work, void* stream) {
  if (n < 0 || n_long < 0 || (n > 0 && (!rowptr || !col || !values || !x || !y)) || (n_long > 0 && !long_rows))
    return -1;
  if ((int64_t)n_long * kLongChunks > (int64_t)INT32_MAX) return -1;
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  mf_csr_mv<<<blocks_for(n, 4), 256, 0, st>>>(n, rowptr, col, values, x, y);
  if (n_long > 0 && work) {
    mf_csr_mv_long_part<<<(unsigned)(n_long * kLongChunks), 256, 0, st>>>(long_rows, rowptr, col, values, x, work);
    mf_csr_mv_long_sum<<<(unsigned)n_long, 256, 0, st>>>(long_rows, work, y);
  } else if (n_long > 0) {
    mf_csr_mv_long<<<(unsigned)n_long, 1024, 0, st>>>(long_rows, rowptr, col, values, x, y);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int vef_mf_solve(vef_mf* h, double* b, double* x, void* stream) {
  if (!h || !h->buf || !b || !x) return -1;
  hipStream_t st = (hipStream_t)stream;
  return h->bits == 64 ? solve_t<double>(h, b, x, st) : solve_t<float>(h, b, x, st);
}

const char* vef_mf_last_error(const vef_mf* h) { return h ? h->err.c_str() : "null handle"; }

void vef_mf_destroy(vef_mf* h) {
  if (!h) return;
  release(h);
  delete h;
}

