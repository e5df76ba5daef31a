// vef_mlp.hip — per-node MLP (n_in -> 128 -> 128 -> 128 -> n_out, GELU) on the
// MI355X matrix cores: forward pass, back-propagation and Adam, behind the C
// ABI of include/vef_mlp.h.
//
// Matrix products use the 16x16x16 BF16 MFMA with FP32 accumulation.  For
// v_mfma_f32_16x16x16bf16_1k, lane l (0..63) supplies A[l%16][4*(l/16)+j] and
// B[4*(l/16)+j][l%16], j = 0..3, and receives D[4*(l/16)+j][l%16].
//
// Layout: a wavefront owns a tile of 16 rows (nodes).  Weights live in FP32
// (the Adam master copy) and, after every update, in BF16 "fragment" order —
// for each 16x16 block the 64 lanes' 4-element B operands contiguously — so a
// lane fetches its operand with one 8-byte load; the transposed weights are
// packed the same way for the backward products.  Activations of the tile sit
// in LDS in BF16.  Training stores each layer's input (BF16) and pre-activation
// (FP32) and the back-propagated errors (BF16) per node; the weight gradients
// are then products over all nodes, split into fixed row chunks whose partial
// sums are added in chunk order, so the result is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "vef_mlp.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 32;     // padded input width (feature 31 = 1 carries the first bias)
constexpr int kW = 128;     // hidden width
constexpr int kOut = 16;    // padded output width
constexpr int kLds = 136;   // LDS row stride (bf16 elements) of a tile's activations
constexpr int kWaves = 4;   // tiles per workgroup
constexpr int kMaxChunks = 64;  // row chunks of the weight-gradient products

// Layer l (1..4): K[l] inputs, N[l] outputs.
__host__ __device__ constexpr int layer_k(int l) { return l == 1 ? kIn : kW; }
__host__ __device__ constexpr int layer_n(int l) { return l == 4 ? kOut : kW; }

// Parameter offsets: W1[32][128] | W2[128][128] | b2 | W3[128][128] | b3 | W4[128][16] | b4
constexpr int64_t kOffW1 = 0;
constexpr int64_t kOffW2 = kOffW1 + kIn * kW;
constexpr int64_t kOffB2 = kOffW2 + kW * kW;
constexpr int64_t kOffW3 = kOffB2 + kW;
constexpr int64_t kOffB3 = kOffW3 + kW * kW;
constexpr int64_t kOffW4 = kOffB3 + kW;
constexpr int64_t kOffB4 = kOffW4 + kW * kOut;
constexpr int64_t kParams = kOffB4 + kOut;

__host__ __device__ constexpr int64_t w_off(int l) {
  return l == 1 ? kOffW1 : l == 2 ? kOffW2 : l == 3 ? kOffW3 : kOffW4;
}
__host__ __device__ constexpr int64_t b_off(int l) { return l == 2 ? kOffB2 : l == 3 ? kOffB3 : kOffB4; }

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2/pi)
__device__ __forceinline__ float gelu(float z) {
  const float u = kGeluC * (z + 0.044715f * z * z * z);
  return 0.5f * z * (1.0f + tanhf(u));
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float u = kGeluC * (z + 0.044715f * z * z * z);
  const float t = tanhf(u);
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * kGeluC * (1.0f + 3.0f * 0.044715f * z * z);
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// A operand of a 16-row tile held in LDS (row stride kLds).
__device__ __forceinline__ bf16x4 lds_a(const uint16_t* s, int lane, int kb) {
  const uint16_t* p = s + (lane & 15) * kLds + kb * 16 + 4 * (lane >> 4);
  bf16x4 v;
  v[0] = (short)p[0];
  v[1] = (short)p[1];
  v[2] = (short)p[2];
  v[3] = (short)p[3];
  return v;
}
// B operand from a packed fragment array [KB][NB][64][4].
__device__ __forceinline__ bf16x4 frag_b(const uint16_t* f, int NB, int kb, int nb, int lane) {
  return *reinterpret_cast<const bf16x4*>(f + ((int64_t)(kb * NB + nb) * 64 + lane) * 4);
}

// ----------------------------------------------------------------------------
// Packing: FP32 W[K][N] -> BF16 fragments of W (forward) and of W^T (backward).
// ----------------------------------------------------------------------------
__global__ void pack_kernel(const float* __restrict__ W, int K, int N, uint16_t* __restrict__ Wf,
                            uint16_t* __restrict__ WTf) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)K * N;
  if (t >= total) return;
  // t enumerates fragment elements: ((kb*NB + nb)*64 + lane)*4 + j
  const int j = (int)(t & 3), lane = (int)((t >> 2) & 63);
  const int64_t blk = t >> 8;
  {
    const int NB = N / 16;
    const int kb = (int)(blk / NB), nb = (int)(blk % NB);
    const int k = kb * 16 + 4 * (lane >> 4) + j, nn = nb * 16 + (lane & 15);
    Wf[t] = to_bf16(W[(int64_t)k * N + nn]);
  }
  if (WTf) {
    // W^T[N][K]: element (k', n') = W[n'][k']
    const int NB = K / 16;
    const int kb = (int)(blk / NB), nb = (int)(blk % NB);
    const int kk = kb * 16 + 4 * (lane >> 4) + j, nn = nb * 16 + (lane & 15);
    WTf[t] = to_bf16(W[(int64_t)nn * N + kk]);
  }
}

struct Net {
  const uint16_t* Wf[5];   // forward fragments, index 1..4
  const uint16_t* WTf[5];  // transposed fragments, index 2..4
  const float* P;          // FP32 parameters (biases)
};

struct Store {
  uint16_t* A[4];  // inputs of layers 1..4: A[0] [n][32], A[1..3] [n][128] (BF16)
  float* Z[4];     // pre-activations of layers 1..3: Z[1..3] [n][128] (FP32)
  float* Y;        // outputs [n][16]
};

// ----------------------------------------------------------------------------
// Forward pass.  STORE keeps what back-propagation needs.
// ----------------------------------------------------------------------------
template <bool STORE>
__global__ void __launch_bounds__(64 * kWaves)
mlp_forward_kernel(int64_t n, int n_in, int n_out, const float* __restrict__ X, Net net, Store st,
                   float* __restrict__ Yout) {
  __shared__ uint16_t sbuf[kWaves][2][16 * kLds];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t row0 = ((int64_t)blockIdx.x * kWaves + wave) * 16;
  uint16_t* cur = sbuf[wave][0];
  uint16_t* nxt = sbuf[wave][1];
  // layer-1 input: features, zero padding, 1 at feature 31
  for (int e = lane; e < 16 * kIn; e += 64) {
    const int r = e / kIn, c = e % kIn;
    const int64_t row = row0 + r;
    float v = 0.0f;
    if (c == kIn - 1) v = 1.0f;
    else if (c < n_in && row < n) v = X[row * n_in + c];
    const uint16_t h = to_bf16(v);
    cur[r * kLds + c] = h;
    if (STORE) st.A[0][row * kIn + c] = h;
  }
  __syncthreads();
  for (int l = 1; l <= 4; l++) {
    const int K = layer_k(l), N = layer_n(l);
    const int KB = K / 16, NB = N / 16;
    for (int nb = 0; nb < NB; nb++) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
      for (int kb = 0; kb < KB; kb++) acc = mfma(lds_a(cur, lane, kb), frag_b(net.Wf[l], NB, kb, nb, lane), acc);
      const int col = nb * 16 + (lane & 15);
      const float bias = (l == 1) ? 0.0f : net.P[b_off(l) + col];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int r = 4 * (lane >> 4) + j;
        const int64_t row = row0 + r;
        const float z = acc[j] + bias;
        if (l < 4) {
          const uint16_t h = to_bf16(gelu(z));
          nxt[r * kLds + col] = h;
          if (STORE) {
            st.Z[l][row * kW + col] = z;
            st.A[l][row * kW + col] = h;
          }
        } else {
          if (STORE) st.Y[row * kOut + col] = z;
          if (Yout && row < n && col < n_out) Yout[row * n_out + col] = z;
        }
      }
    }
    __syncthreads();
    uint16_t* t = cur;
    cur = nxt;
    nxt = t;
  }
}

// ----------------------------------------------------------------------------
// Back-propagation of the errors (dZ of every layer) and the tile's loss.
// dY = 2 (Y - T) / (n * n_out) on the n x n_out outputs, 0 on the padding.
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(64 * kWaves)
mlp_backward_kernel(int64_t n, int n_out, const float* __restrict__ T, Net net, Store st, uint16_t* dZ1,
                    uint16_t* dZ2, uint16_t* dZ3, uint16_t* dZ4, float* __restrict__ loss_part) {
  __shared__ uint16_t sbuf[kWaves][2][16 * kLds];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t tile = (int64_t)blockIdx.x * kWaves + wave;
  const int64_t row0 = tile * 16;
  uint16_t* cur = sbuf[wave][0];
  uint16_t* nxt = sbuf[wave][1];
  uint16_t* dZ[5] = {nullptr, dZ1, dZ2, dZ3, dZ4};
  const float scale = 2.0f / ((float)n * (float)n_out);
  float lsum = 0.0f;
  for (int e = lane; e < 16 * kOut; e += 64) {
    const int r = e / kOut, c = e % kOut;
    const int64_t row = row0 + r;
    float d = 0.0f;
    if (row < n && c < n_out) {
      const float diff = st.Y[row * kOut + c] - T[row * n_out + c];
      lsum += diff * diff;
      d = scale * diff;
    }
    const uint16_t h = to_bf16(d);
    cur[r * kLds + c] = h;
    dZ4[row * kOut + c] = h;
  }
  // the tile's squared error, summed over the wavefront in a fixed order
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off, 64);
  if (lane == 0) loss_part[tile] = lsum;
  __syncthreads();
  for (int l = 4; l >= 2; l--) {
    // dA_{l-1} = dZ_l W_l^T: K = N(l), N = K(l) = 128
    const int K = layer_n(l), N = layer_k(l);
    const int KB = K / 16, NB = N / 16;
    for (int nb = 0; nb < NB; nb++) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
      for (int kb = 0; kb < KB; kb++) acc = mfma(lds_a(cur, lane, kb), frag_b(net.WTf[l], NB, kb, nb, lane), acc);
      const int col = nb * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int r = 4 * (lane >> 4) + j;
        const int64_t row = row0 + r;
        const float dz = acc[j] * gelu_grad(st.Z[l - 1][row * kW + col]);
        const uint16_t h = to_bf16(dz);
        nxt[r * kLds + col] = h;
        dZ[l - 1][row * kW + col] = h;
      }
    }
    __syncthreads();
    uint16_t* t = cur;
    cur = nxt;
    nxt = t;
  }
}

// ----------------------------------------------------------------------------
// Weight gradients: dW_l[K][N] = sum over rows of A_{l-1}[row][k] dZ_l[row][n]
// and, for l >= 2, the bias gradient sum over rows of dZ_l.  Block (b, c):
// one wavefront for the 16x16 output block b over the c-th of C row chunks
// (rows ascending); the partial sums go to part[c] and mlp_reduce_kernel adds
// the chunks in order, so the result does not depend on scheduling.
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(64)
mlp_weight_grad_kernel(int l, int64_t tiles, int chunks, const uint16_t* __restrict__ A,
                       const uint16_t* __restrict__ dZ, float* __restrict__ part) {
  const int K = layer_k(l), N = layer_n(l);
  const int NB = N / 16;
  const int mb = blockIdx.x / NB, nb = blockIdx.x % NB;
  const int c = blockIdx.y;
  const int lane = threadIdx.x;
  const int64_t t0 = tiles * c / chunks, t1 = tiles * (c + 1) / chunks;
  float* out = part + (int64_t)c * kParams;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  float bsum = 0.0f;
  for (int64_t t = t0; t < t1; t++) {
    const int64_t r0 = t * 16;
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int64_t row = r0 + 4 * (lane >> 4) + j;
      a[j] = (short)A[row * K + mb * 16 + (lane & 15)];
      const uint16_t d = dZ[row * N + nb * 16 + (lane & 15)];
      b[j] = (short)d;
      bsum += from_bf16(d);
    }
    acc = mfma(a, b, acc);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int k = mb * 16 + 4 * (lane >> 4) + j, nn = nb * 16 + (lane & 15);
    out[w_off(l) + (int64_t)k * N + nn] = acc[j];
  }
  if (l >= 2 && mb == 0) {
    // lanes l, l+16, l+32, l+48 hold partial sums of column l%16
    bsum += __shfl_down(bsum, 32, 64);
    bsum += __shfl_down(bsum, 16, 64);
    if (lane < 16) out[b_off(l) + nb * 16 + lane] = bsum;
  }
}

__global__ void mlp_reduce_kernel(int chunks, const float* __restrict__ part, float* __restrict__ grad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  float s = part[i];
  for (int c = 1; c < chunks; c++) s += part[(int64_t)c * kParams + i];
  grad[i] = s;
}

__global__ void loss_sum_kernel(int64_t tiles, const float* __restrict__ part, float* __restrict__ loss, float inv) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float s = 0.0f;
  for (int64_t t = 0; t < tiles; t++) s += part[t];
  loss[0] = s * inv;
}

__global__ void adam_kernel(int64_t P, float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, float lr, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1.0e-8f);
}

}  // namespace

struct vef_mlp {
  int n_in = 0, n_out = 0;
  float* d_P = nullptr;  // parameters (FP32)
  float* d_m = nullptr;
  float* d_v = nullptr;
  float* d_g = nullptr;
  float* d_gpart = nullptr;  // [kMaxChunks][params] partial weight gradients
  uint16_t* d_Wf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  uint16_t* d_WTf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int64_t cap = 0;  // rows (multiple of 64) the scratch below holds
  uint16_t* d_A[4] = {nullptr, nullptr, nullptr, nullptr};
  float* d_Z[4] = {nullptr, nullptr, nullptr, nullptr};
  float* d_Y = nullptr;
  uint16_t* d_dZ[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  float* d_lossp = nullptr;
  float* d_loss = nullptr;
  int64_t t = 0;  // Adam steps taken
  bool packed = false;
  std::string err;
};

namespace {
std::string g_mlp_create_err;

int fail(vef_mlp* m, const std::string& msg) {
  if (m) m->err = msg;
  else g_mlp_create_err = msg;
  return 1;
}
#define MLP_HIP(m, call)                                                                        \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) return fail((m), std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int pack_all(vef_mlp* m, hipStream_t st) {
  for (int l = 1; l <= 4; l++) {
    const int K = layer_k(l), N = layer_n(l);
    const int64_t total = (int64_t)K * N;
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, m->d_P + w_off(l), K, N,
                       m->d_Wf[l], l >= 2 ? m->d_WTf[l] : nullptr);
    MLP_HIP(m, hipGetLastError());
  }
  m->packed = true;
  return 0;
}

int ensure_scratch(vef_mlp* m, int64_t n) {
  const int64_t need = ((n + 63) / 64) * 64;
  if (need <= m->cap) return 0;
  for (int i = 0; i < 4; i++) {
    if (m->d_A[i]) (void)hipFree(m->d_A[i]);
    if (m->d_Z[i]) (void)hipFree(m->d_Z[i]);
    m->d_A[i] = nullptr;
    m->d_Z[i] = nullptr;
  }
  for (int i = 0; i < 5; i++) {
    if (m->d_dZ[i]) (void)hipFree(m->d_dZ[i]);
    m->d_dZ[i] = nullptr;
  }
  if (m->d_Y) (void)hipFree(m->d_Y);
  if (m->d_lossp) (void)hipFree(m->d_lossp);
  MLP_HIP(m, hipMalloc((void**)&m->d_A[0], sizeof(uint16_t) * need * kIn));
  for (int i = 1; i < 4; i++) {
    MLP_HIP(m, hipMalloc((void**)&m->d_A[i], sizeof(uint16_t) * need * kW));
    MLP_HIP(m, hipMalloc((void**)&m->d_Z[i], sizeof(float) * need * kW));
    MLP_HIP(m, hipMalloc((void**)&m->d_dZ[i], sizeof(uint16_t) * need * kW));
  }
  MLP_HIP(m, hipMalloc((void**)&m->d_dZ[4], sizeof(uint16_t) * need * kOut));
  MLP_HIP(m, hipMalloc((void**)&m->d_Y, sizeof(float) * need * kOut));
  MLP_HIP(m, hipMalloc((void**)&m->d_lossp, sizeof(float) * (need / 16)));
  m->cap = need;
  return 0;
}

Net net_of(const vef_mlp* m) {
  Net net;
  for (int l = 0; l < 5; l++) {
    net.Wf[l] = m->d_Wf[l];
    net.WTf[l] = m->d_WTf[l];
  }
  net.P = m->d_P;
  return net;
}

Store store_of(const vef_mlp* m) {
  Store s;
  for (int i = 0; i < 4; i++) {
    s.A[i] = m->d_A[i];
    s.Z[i] = m->d_Z[i];
  }
  s.Y = m->d_Y;
  return s;
}

// Forward with storage, back-propagation and weight gradients into d_g; loss into d_loss.
int gradient(vef_mlp* m, int64_t n, const float* X, const float* T, hipStream_t st) {
  if (int rc = ensure_scratch(m, n)) return rc;
  if (!m->packed)
    if (int rc = pack_all(m, st)) return rc;
  const int64_t rows = ((n + 63) / 64) * 64;
  const unsigned blocks = (unsigned)(rows / 64);
  hipLaunchKernelGGL(mlp_forward_kernel<true>, dim3(blocks), dim3(64 * kWaves), 0, st, n, m->n_in, m->n_out, X,
                     net_of(m), store_of(m), (float*)nullptr);
  MLP_HIP(m, hipGetLastError());
  hipLaunchKernelGGL(mlp_backward_kernel, dim3(blocks), dim3(64 * kWaves), 0, st, n, m->n_out, T, net_of(m),
                     store_of(m), m->d_dZ[1], m->d_dZ[2], m->d_dZ[3], m->d_dZ[4], m->d_lossp);
  MLP_HIP(m, hipGetLastError());
  const int64_t tiles = rows / 16;
  const int chunks = (int)std::max<int64_t>(1, std::min<int64_t>(kMaxChunks, tiles / 8));
  for (int l = 1; l <= 4; l++) {
    const int blocks_out = (layer_k(l) / 16) * (layer_n(l) / 16);
    hipLaunchKernelGGL(mlp_weight_grad_kernel, dim3(blocks_out, chunks), dim3(64), 0, st, l, tiles, chunks,
                       m->d_A[l - 1], m->d_dZ[l], m->d_gpart);
    MLP_HIP(m, hipGetLastError());
  }
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((unsigned)((kParams + 255) / 256)), dim3(256), 0, st, chunks,
                     m->d_gpart, m->d_g);
  MLP_HIP(m, hipGetLastError());
  hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(64), 0, st, rows / 16, m->d_lossp, m->d_loss,
                     1.0f / ((float)n * (float)m->n_out));
  MLP_HIP(m, hipGetLastError());
  return 0;
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double normal(uint64_t& s) {
  const double u1 = ((splitmix(s) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
  const double u2 = (splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
}  // namespace

extern "C" {

int vef_mlp_create(int32_t n_in, int32_t n_out, uint64_t seed, vef_mlp** out) {
  if (!out) return fail(nullptr, "out is NULL");
  *out = nullptr;
  if (n_in < 1 || n_in > VEF_MLP_IN_MAX) return fail(nullptr, "n_in must be 1..31");
  if (n_out < 1 || n_out > VEF_MLP_OUT_MAX) return fail(nullptr, "n_out must be 1..16");
  vef_mlp* m = new vef_mlp();
  m->n_in = n_in;
  m->n_out = n_out;
  std::vector<float> P(kParams, 0.0f);
  uint64_t s = seed;
  auto init = [&](int l, int fan_in, int rows, int cols_used) {
    const double sd = 1.0 / std::sqrt((double)fan_in);
    const int N = layer_n(l);
    for (int k = 0; k < rows; k++)
      for (int c = 0; c < cols_used; c++) P[w_off(l) + (int64_t)k * N + c] = (float)(sd * normal(s));
  };
  init(1, n_in, n_in, kW);  // rows n_in..31 stay 0 (row 31 is the first bias)
  init(2, kW, kW, kW);
  init(3, kW, kW, kW);
  init(4, kW, kW, n_out);
  auto cleanup = [&]() { vef_mlp_destroy(m); };
#define CREATE_HIP(call)                                     \
  do {                                                       \
    hipError_t e_ = (call);                                  \
    if (e_ != hipSuccess) {                                  \
      g_mlp_create_err = std::string(#call) + ": " + hipGetErrorString(e_); \
      cleanup();                                             \
      return 1;                                              \
    }                                                        \
  } while (0)
  CREATE_HIP(hipMalloc((void**)&m->d_P, sizeof(float) * kParams));
  CREATE_HIP(hipMalloc((void**)&m->d_m, sizeof(float) * kParams));
  CREATE_HIP(hipMalloc((void**)&m->d_v, sizeof(float) * kParams));
  CREATE_HIP(hipMalloc((void**)&m->d_g, sizeof(float) * kParams));
  CREATE_HIP(hipMalloc((void**)&m->d_gpart, sizeof(float) * kParams * kMaxChunks));
  // entries of unused parameters (padding rows and columns) must read as 0 in every chunk
  CREATE_HIP(hipMemset(m->d_gpart, 0, sizeof(float) * kParams * kMaxChunks));
  CREATE_HIP(hipMalloc((void**)&m->d_loss, sizeof(float)));
  for (int l = 1; l <= 4; l++) {
    const size_t e = (size_t)layer_k(l) * layer_n(l);
    CREATE_HIP(hipMalloc((void**)&m->d_Wf[l], sizeof(uint16_t) * e));
    if (l >= 2) CREATE_HIP(hipMalloc((void**)&m->d_WTf[l], sizeof(uint16_t) * e));
  }
  CREATE_HIP(hipMemcpy(m->d_P, P.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  CREATE_HIP(hipMemset(m->d_m, 0, sizeof(float) * kParams));
  CREATE_HIP(hipMemset(m->d_v, 0, sizeof(float) * kParams));
#undef CREATE_HIP
  *out = m;
  return 0;
}

int64_t vef_mlp_num_params(const vef_mlp* m) { return m ? kParams : 0; }

int vef_mlp_get_params(const vef_mlp* m, float* host) {
  if (!m || !host) return fail(const_cast<vef_mlp*>(m), "bad argument");
  MLP_HIP(const_cast<vef_mlp*>(m), hipMemcpy(host, m->d_P, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  return 0;
}

int vef_mlp_set_params(vef_mlp* m, const float* host) {
  if (!m || !host) return fail(m, "bad argument");
  MLP_HIP(m, hipMemcpy(m->d_P, host, sizeof(float) * kParams, hipMemcpyHostToDevice));
  m->packed = false;
  return 0;
}

int vef_mlp_forward(vef_mlp* m, int64_t n, const float* X, float* Y, void* stream) {
  if (!m) return fail(nullptr, "handle is NULL");
  if (n < 0 || (n > 0 && (!X || !Y))) return fail(m, "bad arguments");
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (!m->packed)
    if (int rc = pack_all(m, st)) return rc;
  const unsigned blocks = (unsigned)((n + 63) / 64);
  Store none{};
  hipLaunchKernelGGL(mlp_forward_kernel<false>, dim3(blocks), dim3(64 * kWaves), 0, st, n, m->n_in, m->n_out, X,
                     net_of(m), none, Y);
  MLP_HIP(m, hipGetLastError());
  return 0;
}

int vef_mlp_gradient(vef_mlp* m, int64_t n, const float* X, const float* T, float* grad, float* loss, void* stream) {
  if (!m) return fail(nullptr, "handle is NULL");
  if (n <= 0 || !X || !T) return fail(m, "bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (int rc = gradient(m, n, X, T, st)) return rc;
  if (grad) MLP_HIP(m, hipMemcpyAsync(grad, m->d_g, sizeof(float) * kParams, hipMemcpyDeviceToDevice, st));
  if (loss) MLP_HIP(m, hipMemcpyAsync(loss, m->d_loss, sizeof(float), hipMemcpyDeviceToDevice, st));
  return 0;
}

int vef_mlp_adam_step(vef_mlp* m, int64_t n, const float* X, const float* T, float lr, float* loss, void* stream) {
  if (!m) return fail(nullptr, "handle is NULL");
  if (n <= 0 || !X || !T) return fail(m, "bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (int rc = gradient(m, n, X, T, st)) return rc;
  m->t += 1;
  const float c1 = 1.0f - std::pow(0.9f, (float)m->t);
  const float c2 = 1.0f - std::pow(0.999f, (float)m->t);
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((kParams + 255) / 256)), dim3(256), 0, st, kParams, m->d_P, m->d_m,
                     m->d_v, m->d_g, lr, c1, c2);
  MLP_HIP(m, hipGetLastError());
  if (int rc = pack_all(m, st)) return rc;
  if (loss) MLP_HIP(m, hipMemcpyAsync(loss, m->d_loss, sizeof(float), hipMemcpyDeviceToDevice, st));
  return 0;
}

const char* vef_mlp_last_error(const vef_mlp* m) { return m ? m->err.c_str() : g_mlp_create_err.c_str(); }

void vef_mlp_destroy(vef_mlp* m) {
  if (!m) return;
  void* ptrs[] = {m->d_P, m->d_m, m->d_v, m->d_g, m->d_gpart, m->d_Y, m->d_lossp, m->d_loss};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int i = 0; i < 5; i++) {
    if (m->d_Wf[i]) (void)hipFree(m->d_Wf[i]);
    if (m->d_WTf[i]) (void)hipFree(m->d_WTf[i]);
    if (m->d_dZ[i]) (void)hipFree(m->d_dZ[i]);
  }
  for (int i = 0; i < 4; i++) {
    if (m->d_A[i]) (void)hipFree(m->d_A[i]);
    if (m->d_Z[i]) (void)hipFree(m->d_Z[i]);
  }
  delete m;
}

}  // extern "C"
