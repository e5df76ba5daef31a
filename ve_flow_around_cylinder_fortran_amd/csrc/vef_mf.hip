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
//   solve:  forward by height (gather b_s, getrs, F21 w, ordered subtract
//           into the update unknowns), backward by height (gather x_u,
//           w -= X x_u, x_s = w).
// Sums whose order matters (extend-add, the forward subtraction) run in a
// fixed order, so results are deterministic.
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

inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// rocSOLVER / rocBLAS by precision
rocblas_status getrf_sb(rocblas_handle h, int n, double* A, int lda, rocblas_stride sA, int* ipiv, rocblas_stride sP,
                        int* info, int bc) {
  return rocsolver_dgetrf_strided_batched(h, n, n, A, lda, sA, ipiv, sP, info, bc);
}
rocblas_status getrf_sb(rocblas_handle h, int n, float* A, int lda, rocblas_stride sA, int* ipiv, rocblas_stride sP,
                        int* info, int bc) {
  return rocsolver_sgetrf_strided_batched(h, n, n, A, lda, sA, ipiv, sP, info, bc);
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
rocblas_status gemv_sb(rocblas_handle h, int m, int n, const double* alpha, const double* A, int lda,
                       rocblas_stride sA, const double* x, rocblas_stride sx, const double* beta, double* y,
                       rocblas_stride sy, int bc) {
  return rocblas_dgemv_strided_batched(h, rocblas_operation_none, m, n, alpha, A, lda, sA, x, 1, sx, beta, y, 1, sy,
                                       bc);
}
rocblas_status gemv_sb(rocblas_handle h, int m, int n, const float* alpha, const float* A, int lda, rocblas_stride sA,
                       const float* x, rocblas_stride sx, const float* beta, float* y, rocblas_stride sy, int bc) {
  return rocblas_sgemv_strided_batched(h, rocblas_operation_none, m, n, alpha, A, lda, sA, x, 1, sx, beta, y, 1, sy,
                                       bc);
}

}  // namespace

struct vef_mf {
  int bits = 64;
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
  int32_t *d_fg_src = nullptr, *d_bg_src = nullptr, *d_bs_dst = nullptr, *d_fs_tgt = nullptr;
  std::vector<int64_t> fg_off, bg_off, bs_off, fs_off, fs_ptr_off;  // per height
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
                  h->d_fs_ptr, h->d_fs_v,   h->d_fg_src, h->d_bg_src, h->d_bs_dst, h->d_fs_tgt};
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
  std::vector<int32_t> fg_src, bg_src, bs_dst, fs_tgt;
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
          fg_src.push_back(rw[i]);
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
  if (upload(h, &h->d_fg_dst, fg_dst) || upload(h, &h->d_fg_src, fg_src) || upload(h, &h->d_bg_dst, bg_dst) ||
      upload(h, &h->d_bg_src, bg_src) || upload(h, &h->d_bs_dst, bs_dst) || upload(h, &h->d_bs_src, bs_src) ||
      upload(h, &h->d_fs_tgt, fs_tgt) || upload(h, &h->d_fs_ptr, fs_ptr) || upload(h, &h->d_fs_v, fs_v))
    return bad(h->err);
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
      if (getrf_sb(rb, s, A, m, sA, h->ipiv + h->b_ipiv[b], s, h->info + h->b_first[b], c) != rocblas_status_success)
        return fail(h, "getrf failed");
      if (u == 0) continue;
      if (getrs_sb(rb, s, u, A, m, sA, h->ipiv + h->b_ipiv[b], s, A + (int64_t)s * m, m, sA, c) !=
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
  const T one = T(1), zero = T(0), mone = T(-1);
  // forward
  int b = 0;
  for (int hh = 0; hh < h->H; hh++) {
    int64_t n = h->fg_off[hh + 1] - h->fg_off[hh];
    if (n) mf_gather<T><<<blocks_for(n, 256), 256, 0, st>>>(W, h->d_fg_dst + h->fg_off[hh], h->d_fg_src + h->fg_off[hh],
                                                           r, n);
    if (fork_streams(h, st)) return -1;
    for (; b < h->nb && h->b_h[b] == hh; b++) {
      int s = h->b_s[b], u = h->b_u[b], c = h->b_c[b];
      if (s == 0) continue;
      rocblas_handle rb = h->rbs[h->b_stream[b]];
      int m = s + u;
      rocblas_stride sA = (rocblas_stride)m * m;
      T* A = buf + h->b_base[b];
      if (getrs_sb(rb, s, 1, A, m, sA, h->ipiv + h->b_ipiv[b], s, W + h->b_w[b], s, s, c) != rocblas_status_success)
        return fail(h, "getrs failed (solve)");
      if (u == 0) continue;
      if (gemv_sb(rb, u, s, &one, A + s, m, sA, W + h->b_w[b], s, &zero, V + h->b_v[b], u, c) !=
          rocblas_status_success)
        return fail(h, "gemv failed (solve)");
    }
    if (join_streams(h, st)) return -1;
    int64_t nt = h->fs_off[hh + 1] - h->fs_off[hh];
    if (nt)
      mf_subtract<T><<<blocks_for(nt, 256), 256, 0, st>>>(r, h->d_fs_tgt + h->fs_off[hh], h->d_fs_ptr + h->fs_ptr_off[hh],
                                                          h->d_fs_v, V, nt);
    if (hipGetLastError() != hipSuccess) return fail(h, "kernel launch failed (forward solve)");
  }
  // backward
  b = h->nb - 1;
  for (int hh = h->H - 1; hh >= 0; hh--) {
    int64_t n = h->bg_off[hh + 1] - h->bg_off[hh];
    if (n)
      mf_gather<T><<<blocks_for(n, 256), 256, 0, st>>>(XU, h->d_bg_dst + h->bg_off[hh], h->d_bg_src + h->bg_off[hh], x,
                                                      n);
    if (fork_streams(h, st)) return -1;
    for (; b >= 0 && h->b_h[b] == hh; b--) {
      int s = h->b_s[b], u = h->b_u[b], c = h->b_c[b];
      if (s == 0 || u == 0) continue;
      rocblas_handle rb = h->rbs[h->b_stream[b]];
      int m = s + u;
      rocblas_stride sA = (rocblas_stride)m * m;
      T* A = buf + h->b_base[b];
      if (gemv_sb(rb, s, u, &mone, A + (int64_t)s * m, m, sA, XU + h->b_v[b], u, &one, W + h->b_w[b], s, c) !=
          rocblas_status_success)
        return fail(h, "gemv failed (backward solve)");
    }
    if (join_streams(h, st)) return -1;
    int64_t np = h->bs_off[hh + 1] - h->bs_off[hh];
    if (np)
      mf_put<T><<<blocks_for(np, 256), 256, 0, st>>>(x, h->d_bs_dst + h->bs_off[hh], h->d_bs_src + h->bs_off[hh], W,
                                                    np);
    if (hipGetLastError() != hipSuccess) return fail(h, "kernel launch failed (backward solve)");
  }
  return 0;
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

