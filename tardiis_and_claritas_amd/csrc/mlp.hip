// mlp.hip -- surrogate MLP ensemble on gfx950 BF16 MFMA (no reference
// counterpart).
//
// n_members independent networks in_dim -> hidden x n_hidden_layers -> out_dim
// with tanh-form GELU after every hidden layer, trained by Adam on the MSE loss.
// Master parameters, gradients and Adam moments are FP32; the matrices the
// hidden GEMMs read are BF16 copies (W [out][in] and W^T [in][out]) refreshed
// after every optimizer step.  Activations entering a GEMM are BF16, products
// accumulate in FP32 on v_mfma_f32_16x16x32_bf16.  The thin first layer
// (in_dim inputs) and last layer (out_dim outputs) run on the VALU in FP32 with
// the BF16 weights.
//
// Every hidden-size matrix product is C[m][n] = sum_k A[m][k] Bt[n][k] with both
// operands row-major and contiguous in k, so each lane loads its MFMA fragment
// (8 consecutive k of one row) with one 16-byte load.  The three products of a
// layer map onto that form by keeping activations and gradients in both
// layouts:
//   forward    Z_l  = Y_{l-1} W_l^T          A = Y_{l-1} [B][in],  Bt = W_l  [out][in]
//   input grad dY   = dZ_l W_l               A = dZ_l    [B][out], Bt = W_l^T [in][out]
//   weight grad dW_l = dZ_l^T Y_{l-1}        A = dZ_l^T  [out][B], Bt = Y_{l-1}^T [in][B]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "mlp.h"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;     // GEMM workgroup tile (rows and columns)
constexpr int kK = 32;        // MFMA depth
constexpr int kAlign = 64;    // device tensors start at multiples of 64 elements

// tanh-form GELU and its derivative
__device__ __forceinline__ float gelu(float z) {
  const float c = 0.7978845608028654f;  // sqrt(2/pi)
  const float t = tanhf(c * (z + 0.044715f * z * z * z));
  return 0.5f * z * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float c = 0.7978845608028654f;
  const float u = c * (z + 0.044715f * z * z * z);
  const float t = tanhf(u);
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * z * z);
}

enum : int { EPI_STORE = 0, EPI_FWD = 1, EPI_BWD = 2, EPI_INFER = 3 };

struct GemmArgs {
  const bf16* A;      // [M][lda] per member
  const bf16* Bt;     // [N][ldb] per member
  int M, N, K, lda, ldb;
  long long sA, sB;   // member strides (elements)
  // EPI_STORE: C[m][n] (FP32, ldc)
  float* C;
  int ldc;
  long long sC;
  // EPI_FWD / EPI_INFER: z = acc + bias[n]; EPI_FWD stores Z, Y and Yt,
  // EPI_INFER stores Y.  EPI_BWD: dz = acc * gelu'(Zin[m][n]) -> Y (= dZ) and Yt.
  const float* bias;
  long long sBias;
  float* Z;
  const float* Zin;
  long long sZ;       // Z/Zin: [M][N]
  bf16* Y;            // [M][N]
  bf16* Yt;           // [N][M]
  long long sY;
};

// 256 threads: wave w computes the 32x32 sub-tile at rows 32*(w/2), cols
// 32*(w%2) of the workgroup's 64x64 tile as 2x2 MFMA 16x16 tiles.  The host
// guarantees M % 64 == 0, N % 64 == 0, K % 32 == 0 and 8-element alignment.
template <int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
  const int member = blockIdx.z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m0 = blockIdx.y * kTile + 32 * (wave >> 1), n0 = blockIdx.x * kTile + 32 * (wave & 1);
  const int r = lane & 15, kb = (lane >> 4) * 8;
  const bf16* __restrict__ A = g.A + member * g.sA;
  const bf16* __restrict__ Bt = g.Bt + member * g.sB;
  const bf16* a_row0 = A + (size_t)(m0 + r) * g.lda + kb;
  const bf16* a_row1 = A + (size_t)(m0 + 16 + r) * g.lda + kb;
  const bf16* b_row0 = Bt + (size_t)(n0 + r) * g.ldb + kb;
  const bf16* b_row1 = Bt + (size_t)(n0 + 16 + r) * g.ldb + kb;
  f32x4 acc00 = {0.f, 0.f, 0.f, 0.f}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
  for (int k0 = 0; k0 < g.K; k0 += kK) {
    const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(a_row0 + k0);
    const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(a_row1 + k0);
    const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(b_row0 + k0);
    const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(b_row1 + k0);
    acc00 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc00, 0, 0, 0);
    acc01 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc01, 0, 0, 0);
    acc10 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc10, 0, 0, 0);
    acc11 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc11, 0, 0, 0);
  }
  // lane holds rows 4*(lane/16)+v of column lane%16 of each 16x16 tile
  const f32x4 acc[2][2] = {{acc00, acc01}, {acc10, acc11}};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + 16 * j + r;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + 16 * i + (lane >> 4) * 4 + v;
        const float a = acc[i][j][v];
        if constexpr (EPI == EPI_STORE) {
          g.C[member * g.sC + (size_t)m * g.ldc + n] = a;
        } else if constexpr (EPI == EPI_FWD || EPI == EPI_INFER) {
          const float z = a + g.bias[member * g.sBias + n];
          const bf16 y = (bf16)gelu(z);
          g.Y[member * g.sY + (size_t)m * g.N + n] = y;
          if constexpr (EPI == EPI_FWD) {
            g.Z[member * g.sZ + (size_t)m * g.N + n] = z;
            g.Yt[member * g.sY + (size_t)n * g.M + m] = y;
          }
        } else {
          const float dz = a * gelu_grad(g.Zin[member * g.sZ + (size_t)m * g.N + n]);
          const bf16 d = (bf16)dz;
          g.Y[member * g.sY + (size_t)m * g.N + n] = d;
          g.Yt[member * g.sY + (size_t)n * g.M + m] = d;
        }
      }
    }
  }
}

// ---- thin layers (VALU, FP32 with BF16 weights) ----

// first layer: z[b][o] = bias[o] + sum_i x[row(b)][i] W[o][i]; rows past n_valid
// (batch padding) read x = 0.  Stores the gathered FP32 input rows for the
// weight gradient.  grid (ceil(B*H/256), members)
struct FirstArgs {
  const float* x;        // [n_rows][din]
  const int* idx;        // [members][n_idx] row permutation or nullptr (identity)
  long long n_idx, base; // rows base .. base+B-1 of the permutation
  int n_valid, B, din, H;
  const bf16* W;         // member stride sP
  const float* bias;
  long long sP;
  float* xg;             // [members][B][din] gathered rows (FWD only)
  float* Z;              // [members][B][H]
  bf16* Y;               // [members][B][H]
  bf16* Yt;              // [members][H][B]
  int store_train;       // 1: Z, Yt and xg too
};
__global__ __launch_bounds__(256) void first_layer_kernel(FirstArgs a) {
  const int member = blockIdx.y;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)a.B * a.H) return;
  const int b = (int)(t / a.H), o = (int)(t % a.H);
  long long row = -1;
  if (b < a.n_valid) row = a.idx ? a.idx[member * a.n_idx + a.base + b] : a.base + b;
  const bf16* W = a.W + member * a.sP + (size_t)o * a.din;
  float z = a.bias[member * a.sP + o];
  for (int i = 0; i < a.din; ++i) {
    const float xi = (row >= 0) ? a.x[row * a.din + i] : 0.0f;
    z += xi * (float)W[i];
    if (a.store_train && o == 0) a.xg[((size_t)member * a.B + b) * a.din + i] = xi;
  }
  const bf16 y = (bf16)gelu(z);
  const size_t mo = (size_t)member * a.B * a.H;
  a.Y[mo + (size_t)b * a.H + o] = y;
  if (a.store_train) {
    a.Z[mo + (size_t)b * a.H + o] = z;
    a.Yt[mo + (size_t)o * a.B + b] = y;
  }
}

// last layer: out[b][j] = bias[j] + sum_o y[b][o] W[j][o].  grid (ceil(B/256), members)
struct LastArgs {
  const bf16* Y;     // [members][B][H]
  const bf16* W;     // [dout][H], member stride sP
  const float* bias;
  long long sP;
  int B, H, dout;
  float* out;        // [members][B][dout]
};
__global__ __launch_bounds__(256) void last_layer_kernel(LastArgs a) {
  const int member = blockIdx.y;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.B) return;
  const bf16* y = a.Y + ((size_t)member * a.B + b) * a.H;
  for (int j = 0; j < a.dout; ++j) {
    const bf16* W = a.W + member * a.sP + (size_t)j * a.H;
    float s = a.bias[member * a.sP + j];
    for (int o = 0; o < a.H; ++o) s += (float)y[o] * (float)W[o];
    a.out[((size_t)member * a.B + b) * a.dout + j] = s;
  }
}

// weighted MSE: dout = 2 w (out - target) / (n_valid dout);
// loss[member] += sum w d^2 / (n_valid dout); w = 1 without a weight array
struct LossArgs {
  const float* out;   // [members][B][dout]
  const float* y;     // [n_rows][dout]
  const float* w;     // [n_rows][dout] or null
  const int* idx;
  long long n_idx, base;
  int n_valid, B, dout;
  float* dout_buf;    // [members][B][dout]
  float* loss;        // [members]
};
__global__ __launch_bounds__(256) void loss_kernel(LossArgs a) {
  const int member = blockIdx.y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.B * a.dout) return;
  const int b = t / a.dout, j = t % a.dout;
  float d = 0.0f, wt = 0.0f;
  if (b < a.n_valid) {
    const long long row = a.idx ? a.idx[member * a.n_idx + a.base + b] : a.base + b;
    d = a.out[((size_t)member * a.B + b) * a.dout + j] - a.y[row * a.dout + j];
    wt = a.w ? a.w[row * a.dout + j] : 1.0f;
  }
  const float inv = 1.0f / (float)(a.n_valid * a.dout);
  a.dout_buf[((size_t)member * a.B + b) * a.dout + j] = 2.0f * wt * d * inv;
  if (d != 0.0f && wt != 0.0f) atomicAdd(a.loss + member, wt * d * d * inv);
}

// last-layer gradients: dW[j][o] = sum_b dout[b][j] y[b][o] (from Yt), db[j];
// grid (ceil((dout*H + dout)/256), members)
struct LastGradArgs {
  const float* dout_buf;
  const bf16* Yt;     // [members][H][B]
  int B, H, dout;
  float* gW;          // member stride sP
  float* gb;
  long long sP;
};
__global__ __launch_bounds__(256) void last_grad_kernel(LastGradArgs a) {
  const int member = blockIdx.y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const float* d = a.dout_buf + (size_t)member * a.B * a.dout;
  if (t < a.dout * a.H) {
    const int j = t / a.H, o = t % a.H;
    const bf16* yt = a.Yt + ((size_t)member * a.H + o) * a.B;
    float s = 0.0f;
    for (int b = 0; b < a.B; ++b) s += d[b * a.dout + j] * (float)yt[b];
    a.gW[member * a.sP + (size_t)j * a.H + o] = s;
  } else if (t < a.dout * a.H + a.dout) {
    const int j = t - a.dout * a.H;
    float s = 0.0f;
    for (int b = 0; b < a.B; ++b) s += d[b * a.dout + j];
    a.gb[member * a.sP + j] = s;
  }
}

// gradient into the last hidden layer: dz[b][o] = gelu'(z[b][o]) sum_j dout[b][j] W[j][o]
struct LastBackArgs {
  const float* dout_buf;
  const bf16* W;      // [dout][H], member stride sP
  long long sP;
  const float* Z;     // [members][B][H]
  int B, H, dout;
  bf16* DZ;           // [members][B][H]
  bf16* DZt;          // [members][H][B]
};
__global__ __launch_bounds__(256) void last_back_kernel(LastBackArgs a) {
  const int member = blockIdx.y;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)a.B * a.H) return;
  const int b = (int)(t / a.H), o = (int)(t % a.H);
  const float* d = a.dout_buf + ((size_t)member * a.B + b) * a.dout;
  float s = 0.0f;
  for (int j = 0; j < a.dout; ++j) s += d[j] * (float)a.W[member * a.sP + (size_t)j * a.H + o];
  const size_t mo = (size_t)member * a.B * a.H;
  const bf16 v = (bf16)(s * gelu_grad(a.Z[mo + (size_t)b * a.H + o]));
  a.DZ[mo + (size_t)b * a.H + o] = v;
  a.DZt[mo + (size_t)o * a.B + b] = v;
}

// bias gradient of a hidden layer: db[o] = sum_b dz[b][o] (from DZt rows)
__global__ __launch_bounds__(256) void bias_grad_kernel(const bf16* DZt, int B, int H, float* gb, long long sP) {
  const int member = blockIdx.y;
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= H) return;
  const bf16* r = DZt + ((size_t)member * H + o) * B;
  float s = 0.0f;
  for (int b = 0; b < B; ++b) s += (float)r[b];
  gb[member * sP + o] = s;
}

// first-layer weight gradient: dW[o][i] = sum_b dz[b][o] x[b][i]
__global__ __launch_bounds__(256) void first_grad_kernel(const bf16* DZt, const float* xg, int B, int H, int din,
                                                         float* gW, long long sP) {
  const int member = blockIdx.y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= H * din) return;
  const int o = t / din, i = t % din;
  const bf16* r = DZt + ((size_t)member * H + o) * B;
  const float* x = xg + (size_t)member * B * din;
  float s = 0.0f;
  for (int b = 0; b < B; ++b) s += (float)r[b] * x[b * din + i];
  gW[member * sP + (size_t)o * din + i] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters
__global__ __launch_bounds__(256) void adam_kernel(float* p, const float* g, float* m, float* v, long long n,
                                                   float lr, float b1, float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of one weight matrix W [out][in]: W and W^T
__global__ __launch_bounds__(256) void to_bf16_kernel(const float* p, bf16* w, bf16* wt, long long off, int out,
                                                      int in, long long sP) {
  const int member = blockIdx.y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= out * in) return;
  const int o = t / in, i = t % in;
  const float val = p[member * sP + off + t];
  w[member * sP + off + t] = (bf16)val;
  wt[member * sP + off + (size_t)i * out + o] = (bf16)val;
}

// ensemble mean and sample standard deviation over members
__global__ __launch_bounds__(256) void ensemble_kernel(const float* out, int members, int B, int dout, int n_valid,
                                                       float* mean, float* stdev) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_valid * dout) return;
  float s = 0.0f;
  for (int m = 0; m < members; ++m) s += out[(size_t)m * B * dout + t];
  const float mu = s / (float)members;
  float q = 0.0f;
  for (int m = 0; m < members; ++m) {
    const float d = out[(size_t)m * B * dout + t] - mu;
    q += d * d;
  }
  mean[t] = mu;
  if (stdev) stdev[t] = members > 1 ? sqrtf(q / (float)(members - 1)) : 0.0f;
}

__global__ void fill_f32_kernel(float* p, long long n, float v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---- fused ensemble inference ----
// One workgroup (4 waves) runs every member of the ensemble on 64 query rows
// with the activations held in LDS: the first layer on the VALU, each hidden
// matrix product on MFMA (wave w owns output columns 64w .. 64w+63 of all 64
// rows as 4 x 4 tiles of 16 x 16), the output layer from LDS, then the
// ensemble mean and spread.  Every value is rounded and summed in the order of
// the unfused path (first_layer_kernel, gemm_kernel<EPI_INFER>,
// last_layer_kernel, ensemble_kernel), so both give identical results; only
// the per-layer activations no longer travel through HBM.  Shapes: hidden
// <= 256, in_dim <= 32, members * out_dim <= 80 (the config-4 ensemble is
// 256 / 8 / 15).
constexpr int kFRows = 64;
constexpr int kFMaxH = 256, kFMaxIn = 32, kFMaxOutVals = 80, kFMaxLayers = 17;

struct FusedArgs {
  const float* x;  // [n][din]
  long long n;
  int din, H, dout, members, n_layers;
  const bf16* wbf;
  const float* params;
  long long sP;
  long long w_off[kFMaxLayers], b_off[kFMaxLayers];
  float* mean;   // [n][dout]
  float* stdev;  // [n][dout] or nullptr
};

__device__ __forceinline__ void fused_gemm(const bf16* __restrict__ A, bf16* __restrict__ Y, int S,
                                           const bf16* __restrict__ W, const float* __restrict__ bias, int H,
                                           int wave, int lane) {
  const int n0 = 64 * wave;
  if (n0 >= H) return;
  const int r = lane & 15, kb = (lane >> 4) * 8;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16* brow = W + (size_t)(n0 + r) * H + kb;
  bf16x8 bn[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) bn[j] = *reinterpret_cast<const bf16x8*>(brow + (size_t)16 * j * H);
  for (int k0 = 0; k0 < H; k0 += kK) {
    bf16x8 bv[4], av[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) bv[j] = bn[j];
    if (k0 + kK < H) {
#pragma unroll
      for (int j = 0; j < 4; ++j) bn[j] = *reinterpret_cast<const bf16x8*>(brow + (size_t)16 * j * H + k0 + kK);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) av[i] = *reinterpret_cast<const bf16x8*>(A + (16 * i + r) * S + kb + k0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + 16 * j + r;
      const float b = bias[n];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = 16 * i + (lane >> 4) * 4 + v;
        Y[m * S + n] = (bf16)gelu(acc[i][j][v] + b);
      }
    }
}

__global__ __launch_bounds__(256) void fused_infer_kernel(FusedArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int S = a.H + 8;  // padded row: rows 0..7 of a fragment fall in distinct LDS banks
  bf16* act0 = reinterpret_cast<bf16*>(smem);
  bf16* act1 = act0 + kFRows * S;
  float* xs = reinterpret_cast<float*>(act1 + kFRows * S);
  float* outv = xs + kFRows * a.din;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const long long row0 = (long long)blockIdx.x * kFRows;
  const int nv = (int)min((long long)kFRows, a.n - row0);
  for (int i = tid; i < kFRows * a.din; i += 256) {
    const int r = i / a.din, c = i - r * a.din;
    xs[i] = (r < nv) ? a.x[(row0 + r) * a.din + c] : 0.0f;  // padding rows read 0, as the batched path
  }
  __syncthreads();
  const int L = a.n_layers;
  for (int m = 0; m < a.members; ++m) {
    const bf16* Wm = a.wbf + m * a.sP;
    const float* Pm = a.params + m * a.sP;
    for (int o = tid; o < a.H; o += 256) {
      const bf16* W0 = Wm + a.w_off[0] + (size_t)o * a.din;
      float w[kFMaxIn];
#pragma unroll
      for (int i = 0; i < kFMaxIn; ++i) w[i] = (i < a.din) ? (float)W0[i] : 0.0f;
      const float b0 = Pm[a.b_off[0] + o];
      for (int r = 0; r < kFRows; ++r) {
        float z = b0;
#pragma unroll
        for (int i = 0; i < kFMaxIn; ++i)
          if (i < a.din) z += xs[r * a.din + i] * w[i];
        act0[r * S + o] = (bf16)gelu(z);
      }
    }
    __syncthreads();
    bf16* cur = act0;
    bf16* nxt = act1;
    for (int l = 1; l < L - 1; ++l) {
      fused_gemm(cur, nxt, S, Wm + a.w_off[l], Pm + a.b_off[l], a.H, wave, lane);
      __syncthreads();
      bf16* t = cur; cur = nxt; nxt = t;
    }
    for (int t = tid; t < kFRows * a.dout; t += 256) {
      const int r = t / a.dout, j = t - r * a.dout;
      const bf16* W = Wm + a.w_off[L - 1] + (size_t)j * a.H;
      float sum = Pm[a.b_off[L - 1] + j];
      const bf16* y = cur + r * S;
      for (int o = 0; o < a.H; ++o) sum += (float)y[o] * (float)W[o];
      outv[(m * kFRows + r) * a.dout + j] = sum;
    }
    __syncthreads();
  }
  for (int t = tid; t < nv * a.dout; t += 256) {
    const int r = t / a.dout, j = t - r * a.dout;
    float sum = 0.0f;
    for (int m = 0; m < a.members; ++m) sum += outv[(m * kFRows + r) * a.dout + j];
    const float mu = sum / (float)a.members;
    float q = 0.0f;
    for (int m = 0; m < a.members; ++m) {
      const float d = outv[(m * kFRows + r) * a.dout + j] - mu;
      q += d * d;
    }
    a.mean[(row0 + r) * a.dout + j] = mu;
    if (a.stdev) a.stdev[(row0 + r) * a.dout + j] = a.members > 1 ? sqrtf(q / (float)(a.members - 1)) : 0.0f;
  }
}

inline unsigned int blocks(long long n, int b = 256) { return (unsigned int)((n + b - 1) / b); }
inline long long round_up(long long v, long long a) { return (v + a - 1) / a * a; }

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t want) {
    if (want <= n) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc(&p, std::max<size_t>(want, 1) * sizeof(T));
    if (e == hipSuccess) n = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct Tensor {
  long long host_off, dev_off, size;  // in elements
  int out, in;                        // W: out x in; bias: out x 1 (in = 0)
};

}  // namespace

struct MlpState {
  clr_mlp_shape sh{};
  int n_layers = 0;                 // n_hidden_layers + 1
  std::vector<int> lin, lout;       // per layer
  std::vector<Tensor> W, b;         // per layer
  long long P = 0, sP = 0;          // host params per member, device stride
  DBuf<float> params, grads, am, av;
  DBuf<bf16> wbf, wtbf;
  int step = 0;
  // workspaces for a batch of B rows (B % 64 == 0)
  int B = 0;
  DBuf<float> xg, out, dout, loss;
  std::vector<DBuf<float>> Z;       // hidden layers: [M][B][H]
  std::vector<DBuf<bf16>> Y, Yt, DZ, DZt;
  DBuf<float> x_d, y_d, w_d;        // uploaded data
  DBuf<int> idx_d;
  DBuf<float> mean_d, std_d;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~MlpState() {
    for (auto* d : {&params, &grads, &am, &av, &xg, &out, &dout, &loss, &x_d, &y_d, &w_d, &mean_d, &std_d}) d->release();
    wbf.release();
    wtbf.release();
    idx_d.release();
    for (auto& v : Z) v.release();
    for (auto* vec : {&Y, &Yt, &DZ, &DZt})
      for (auto& v : *vec) v.release();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

void clr_mlp_free(MlpState* st) { delete st; }

namespace {

#define MLPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      err = std::string("HIP error in surrogate: ") + hipGetErrorString(e_);      \
      return CLR_E_HIP;                                                           \
    }                                                                             \
  } while (0)

int ensure_batch(MlpState* st, int B, std::string& err) {
  if (B <= st->B) return 0;
  const int M = st->sh.n_members, H = st->sh.hidden, L = st->n_layers;
  MLPCHK(st->xg.ensure((size_t)M * B * st->sh.in_dim));
  MLPCHK(st->out.ensure((size_t)M * B * st->sh.out_dim));
  MLPCHK(st->dout.ensure((size_t)M * B * st->sh.out_dim));
  st->Z.resize(L - 1);
  st->Y.resize(L - 1);
  st->Yt.resize(L - 1);
  st->DZ.resize(L - 1);
  st->DZt.resize(L - 1);
  const size_t act = (size_t)M * B * H;
  for (int l = 0; l < L - 1; ++l) {
    MLPCHK(st->Z[l].ensure(act));
    MLPCHK(st->Y[l].ensure(act));
    MLPCHK(st->Yt[l].ensure(act));
    MLPCHK(st->DZ[l].ensure(act));
    MLPCHK(st->DZt[l].ensure(act));
  }
  st->B = B;
  return 0;
}

void launch_gemm(int epi, const GemmArgs& g, int members, hipStream_t s) {
  const dim3 grid((unsigned)(g.N / kTile), (unsigned)(g.M / kTile), (unsigned)members);
  switch (epi) {
    case EPI_STORE: hipLaunchKernelGGL(gemm_kernel<EPI_STORE>, grid, dim3(256), 0, s, g); break;
    case EPI_FWD: hipLaunchKernelGGL(gemm_kernel<EPI_FWD>, grid, dim3(256), 0, s, g); break;
    case EPI_BWD: hipLaunchKernelGGL(gemm_kernel<EPI_BWD>, grid, dim3(256), 0, s, g); break;
    default: hipLaunchKernelGGL(gemm_kernel<EPI_INFER>, grid, dim3(256), 0, s, g); break;
  }
}

// forward pass of a padded batch of B rows.  train: keep Z, Yt and the gathered
// input for the backward pass.  Output in st->out [M][B][dout].
void forward(MlpState* st, const float* x, const int* idx, long long n_idx, long long base, int n_valid, int B,
             bool train, hipStream_t s) {
  const int M = st->sh.n_members, H = st->sh.hidden, L = st->n_layers;
  FirstArgs f{};
  f.x = x; f.idx = idx; f.n_idx = n_idx; f.base = base; f.n_valid = n_valid; f.B = B;
  f.din = st->sh.in_dim; f.H = H; f.W = st->wbf.p + st->W[0].dev_off; f.bias = st->params.p + st->b[0].dev_off;
  f.sP = st->sP; f.xg = st->xg.p; f.Z = st->Z[0].p; f.Y = st->Y[0].p; f.Yt = st->Yt[0].p;
  f.store_train = train ? 1 : 0;
  hipLaunchKernelGGL(first_layer_kernel, dim3(blocks((long long)B * H), M), dim3(256), 0, s, f);
  for (int l = 1; l < L - 1; ++l) {
    GemmArgs g{};
    g.A = st->Y[l - 1].p; g.Bt = st->wbf.p + st->W[l].dev_off;
    g.M = B; g.N = H; g.K = H; g.lda = H; g.ldb = H;
    g.sA = (long long)B * H; g.sB = st->sP;
    g.bias = st->params.p + st->b[l].dev_off; g.sBias = st->sP;
    g.Z = st->Z[l].p; g.sZ = (long long)B * H;
    g.Y = st->Y[l].p; g.Yt = st->Yt[l].p; g.sY = (long long)B * H;
    launch_gemm(train ? EPI_FWD : EPI_INFER, g, M, s);
  }
  LastArgs a{};
  a.Y = st->Y[L - 2].p; a.W = st->wbf.p + st->W[L - 1].dev_off; a.bias = st->params.p + st->b[L - 1].dev_off;
  a.sP = st->sP; a.B = B; a.H = H; a.dout = st->sh.out_dim; a.out = st->out.p;
  hipLaunchKernelGGL(last_layer_kernel, dim3(blocks(B), M), dim3(256), 0, s, a);
}

// loss and backward pass after forward(train=true); gradients into st->grads
void backward(MlpState* st, const float* y, const float* w, const int* idx, long long n_idx, long long base,
              int n_valid, int B, hipStream_t s) {
  const int M = st->sh.n_members, H = st->sh.hidden, L = st->n_layers, dout = st->sh.out_dim;
  LossArgs la{};
  la.out = st->out.p; la.y = y; la.w = w; la.idx = idx; la.n_idx = n_idx; la.base = base; la.n_valid = n_valid; la.B = B;
  la.dout = dout; la.dout_buf = st->dout.p; la.loss = st->loss.p;
  hipLaunchKernelGGL(loss_kernel, dim3(blocks((long long)B * dout), M), dim3(256), 0, s, la);
  LastGradArgs lg{};
  lg.dout_buf = st->dout.p; lg.Yt = st->Yt[L - 2].p; lg.B = B; lg.H = H; lg.dout = dout;
  lg.gW = st->grads.p + st->W[L - 1].dev_off; lg.gb = st->grads.p + st->b[L - 1].dev_off; lg.sP = st->sP;
  hipLaunchKernelGGL(last_grad_kernel, dim3(blocks((long long)dout * H + dout), M), dim3(256), 0, s, lg);
  LastBackArgs lb{};
  lb.dout_buf = st->dout.p; lb.W = st->wbf.p + st->W[L - 1].dev_off; lb.sP = st->sP; lb.Z = st->Z[L - 2].p;
  lb.B = B; lb.H = H; lb.dout = dout; lb.DZ = st->DZ[L - 2].p; lb.DZt = st->DZt[L - 2].p;
  hipLaunchKernelGGL(last_back_kernel, dim3(blocks((long long)B * H), M), dim3(256), 0, s, lb);
  for (int l = L - 2; l >= 1; --l) {
    // weight gradient dW_l = dZ_l^T Y_{l-1}
    GemmArgs gw{};
    gw.A = st->DZt[l].p; gw.Bt = st->Yt[l - 1].p;
    gw.M = H; gw.N = H; gw.K = B; gw.lda = B; gw.ldb = B;
    gw.sA = (long long)B * H; gw.sB = (long long)B * H;
    gw.C = st->grads.p + st->W[l].dev_off; gw.ldc = H; gw.sC = st->sP;
    launch_gemm(EPI_STORE, gw, M, s);
    hipLaunchKernelGGL(bias_grad_kernel, dim3(blocks(H), M), dim3(256), 0, s, st->DZt[l].p, B, H,
                       st->grads.p + st->b[l].dev_off, st->sP);
    // input gradient through W_l and GELU' of layer l-1
    GemmArgs gx{};
    gx.A = st->DZ[l].p; gx.Bt = st->wtbf.p + st->W[l].dev_off;
    gx.M = B; gx.N = H; gx.K = H; gx.lda = H; gx.ldb = H;
    gx.sA = (long long)B * H; gx.sB = st->sP;
    gx.Zin = st->Z[l - 1].p; gx.sZ = (long long)B * H;
    gx.Y = st->DZ[l - 1].p; gx.Yt = st->DZt[l - 1].p; gx.sY = (long long)B * H;
    launch_gemm(EPI_BWD, gx, M, s);
  }
  hipLaunchKernelGGL(bias_grad_kernel, dim3(blocks(H), M), dim3(256), 0, s, st->DZt[0].p, B, H,
                     st->grads.p + st->b[0].dev_off, st->sP);
  hipLaunchKernelGGL(first_grad_kernel, dim3(blocks((long long)H * st->sh.in_dim), M), dim3(256), 0, s,
                     st->DZt[0].p, st->xg.p, B, H, st->sh.in_dim, st->grads.p + st->W[0].dev_off, st->sP);
}

void refresh_bf16(MlpState* st, hipStream_t s) {
  const int M = st->sh.n_members;
  for (const Tensor& w : st->W)
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks((long long)w.out * w.in), M), dim3(256), 0, s, st->params.p,
                       st->wbf.p, st->wtbf.p, w.dev_off, w.out, w.in, st->sP);
}

bool fused_ok(const MlpState* st) {
  const clr_mlp_shape& h = st->sh;
  return h.hidden <= kFMaxH && h.in_dim <= kFMaxIn && h.n_members * h.out_dim <= kFMaxOutVals &&
         st->n_layers <= kFMaxLayers;
}

size_t fused_lds_bytes(const MlpState* st) {
  const clr_mlp_shape& h = st->sh;
  return (size_t)2 * kFRows * (h.hidden + 8) * sizeof(bf16) + (size_t)kFRows * h.in_dim * sizeof(float) +
         (size_t)h.n_members * kFRows * h.out_dim * sizeof(float);
}

// one fused pass over n device-resident query rows (fused_ok(st) holds)
hipError_t launch_fused(MlpState* st, const float* x, long long n, float* mean, float* stdev, hipStream_t s) {
  FusedArgs a{};
  a.x = x; a.n = n; a.din = st->sh.in_dim; a.H = st->sh.hidden; a.dout = st->sh.out_dim;
  a.members = st->sh.n_members; a.n_layers = st->n_layers; a.wbf = st->wbf.p; a.params = st->params.p;
  a.sP = st->sP; a.mean = mean; a.stdev = stdev;
  for (int l = 0; l < st->n_layers; ++l) {
    a.w_off[l] = st->W[l].dev_off;
    a.b_off[l] = st->b[l].dev_off;
  }
  const size_t lds = fused_lds_bytes(st);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fused_infer_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fused_infer_kernel, dim3((unsigned)((n + kFRows - 1) / kFRows)), dim3(256), lds, s, a);
  return hipGetLastError();
}

// host flat layout <-> device padded layout, tensor by tensor
hipError_t copy_params(MlpState* st, float* host, const float* dev_src, bool to_host, hipStream_t s) {
  for (int m = 0; m < st->sh.n_members; ++m) {
    for (int l = 0; l < st->n_layers; ++l) {
      for (const Tensor* t : {&st->W[l], &st->b[l]}) {
        float* h = host + (size_t)m * st->P + t->host_off;
        float* d = const_cast<float*>(dev_src) + (size_t)m * st->sP + t->dev_off;
        hipError_t e = to_host ? hipMemcpyAsync(h, d, sizeof(float) * t->size, hipMemcpyDeviceToHost, s)
                               : hipMemcpyAsync(d, h, sizeof(float) * t->size, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
      }
    }
  }
  return hipStreamSynchronize(s);
}

}  // namespace

int mlp_init(MlpState** out, const clr_mlp_shape* shape, uint64_t seed, hipStream_t s, std::string& err) {
  if (!shape || shape->n_members < 1 || shape->n_members > 64 || shape->in_dim < 1 || shape->in_dim > 256 ||
      shape->out_dim < 1 || shape->out_dim > 64 || shape->hidden < 64 || shape->hidden > 4096 ||
      shape->hidden % 64 != 0 || shape->n_hidden_layers < 1 || shape->n_hidden_layers > 16) {
    err = "surrogate shape: need 1..64 members, in_dim 1..256, out_dim 1..64, hidden a multiple of 64 "
          "in 64..4096, 1..16 hidden layers";
    return CLR_E_ARG;
  }
  MlpState* st = new MlpState();
  st->sh = *shape;
  const int L = shape->n_hidden_layers + 1;
  st->n_layers = L;
  for (int l = 0; l < L; ++l) {
    st->lin.push_back(l == 0 ? shape->in_dim : shape->hidden);
    st->lout.push_back(l == L - 1 ? shape->out_dim : shape->hidden);
  }
  long long ho = 0, dv = 0;
  for (int l = 0; l < L; ++l) {
    Tensor w{ho, dv, (long long)st->lout[l] * st->lin[l], st->lout[l], st->lin[l]};
    ho += w.size;
    dv = round_up(dv + w.size, kAlign);
    Tensor b{ho, dv, (long long)st->lout[l], st->lout[l], 0};
    ho += b.size;
    dv = round_up(dv + b.size, kAlign);
    st->W.push_back(w);
    st->b.push_back(b);
  }
  st->P = ho;
  st->sP = dv;
  const size_t total = (size_t)shape->n_members * st->sP;
  auto fail_hip = [&](hipError_t e) {
    err = std::string("HIP error in surrogate init: ") + hipGetErrorString(e);
    delete st;
    return CLR_E_HIP;
  };
  hipError_t e;
  if ((e = st->params.ensure(total)) != hipSuccess || (e = st->grads.ensure(total)) != hipSuccess ||
      (e = st->am.ensure(total)) != hipSuccess || (e = st->av.ensure(total)) != hipSuccess ||
      (e = st->wbf.ensure(total)) != hipSuccess || (e = st->wtbf.ensure(total)) != hipSuccess ||
      (e = st->loss.ensure((size_t)shape->n_members)) != hipSuccess)
    return fail_hip(e);
  if ((e = hipEventCreate(&st->e0)) != hipSuccess || (e = hipEventCreate(&st->e1)) != hipSuccess) return fail_hip(e);
  // Glorot-uniform weights, zero biases, one stream per member from `seed`
  std::vector<float> host((size_t)shape->n_members * st->P, 0.0f);
  for (int m = 0; m < shape->n_members; ++m) {
    std::mt19937_64 gen(seed * 0x9E3779B97F4A7C15ull + (uint64_t)m);
    for (int l = 0; l < L; ++l) {
      const double a = std::sqrt(6.0 / (double)(st->lin[l] + st->lout[l]));
      std::uniform_real_distribution<double> u(-a, a);
      float* w = host.data() + (size_t)m * st->P + st->W[l].host_off;
      for (long long i = 0; i < st->W[l].size; ++i) w[i] = (float)u(gen);
    }
  }
  for (auto* d : {&st->params, &st->grads, &st->am, &st->av})
    hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks((long long)total)), dim3(256), 0, s, d->p, (long long)total, 0.0f);
  if ((e = copy_params(st, host.data(), st->params.p, false, s)) != hipSuccess) return fail_hip(e);
  refresh_bf16(st, s);
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail_hip(e);
  *out = st;
  return 0;
}

size_t mlp_param_count(const MlpState* st) { return st ? (size_t)st->sh.n_members * st->P : 0; }

int mlp_get_params(MlpState* st, float* out, size_t n, hipStream_t s, std::string& err) {
  if (n < mlp_param_count(st)) { err = "output buffer smaller than clr_mlp_param_count()"; return CLR_E_ARG; }
  MLPCHK(copy_params(st, out, st->params.p, true, s));
  return 0;
}

int mlp_set_params(MlpState* st, const float* in, size_t n, hipStream_t s, std::string& err) {
  if (n != mlp_param_count(st)) { err = "parameter count differs from clr_mlp_param_count()"; return CLR_E_ARG; }
  MLPCHK(copy_params(st, const_cast<float*>(in), st->params.p, false, s));
  refresh_bf16(st, s);
  MLPCHK(hipStreamSynchronize(s));
  st->step = 0;
  const size_t total = (size_t)st->sh.n_members * st->sP;
  for (auto* d : {&st->am, &st->av})
    hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks((long long)total)), dim3(256), 0, s, d->p, (long long)total, 0.0f);
  MLPCHK(hipStreamSynchronize(s));
  return 0;
}

int mlp_grad(MlpState* st, const float* x, const float* y, const float* w, int32_t n, float* loss, float* grads,
             hipStream_t s, std::string& err) {
  if (n < 1 || n > (1 << 20)) { err = "gradient batch must hold 1..2^20 rows"; return CLR_E_ARG; }
  const int B = (int)round_up(n, kTile);
  if (int rc = ensure_batch(st, B, err)) return rc;
  const int M = st->sh.n_members;
  MLPCHK(st->x_d.ensure((size_t)n * st->sh.in_dim));
  MLPCHK(st->y_d.ensure((size_t)n * st->sh.out_dim));
  MLPCHK(hipMemcpyAsync(st->x_d.p, x, sizeof(float) * n * st->sh.in_dim, hipMemcpyHostToDevice, s));
  MLPCHK(hipMemcpyAsync(st->y_d.p, y, sizeof(float) * n * st->sh.out_dim, hipMemcpyHostToDevice, s));
  if (w) {
    MLPCHK(st->w_d.ensure((size_t)n * st->sh.out_dim));
    MLPCHK(hipMemcpyAsync(st->w_d.p, w, sizeof(float) * n * st->sh.out_dim, hipMemcpyHostToDevice, s));
  }
  const size_t total = (size_t)M * st->sP;
  hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks((long long)total)), dim3(256), 0, s, st->grads.p, (long long)total, 0.0f);
  hipLaunchKernelGGL(fill_f32_kernel, dim3(1), dim3(64), 0, s, st->loss.p, (long long)M, 0.0f);
  forward(st, st->x_d.p, nullptr, 0, 0, n, B, true, s);
  backward(st, st->y_d.p, w ? st->w_d.p : nullptr, nullptr, 0, 0, n, B, s);
  MLPCHK(hipGetLastError());
  if (loss) MLPCHK(hipMemcpyAsync(loss, st->loss.p, sizeof(float) * M, hipMemcpyDeviceToHost, s));
  if (grads) MLPCHK(copy_params(st, grads, st->grads.p, true, s));
  MLPCHK(hipStreamSynchronize(s));
  return 0;
}

int mlp_train(MlpState* st, const float* x, const float* y, const float* w, int64_t n, int32_t batch, int32_t epochs,
              float lr, uint64_t seed, float* loss, hipStream_t s, std::string& err) {
  if (n < 1 || n > ((int64_t)1 << 31) - 1 || batch < 1 || batch > (1 << 16) || epochs < 1 || !(lr > 0.0f)) {
    err = "train: need 1 <= n < 2^31, 1 <= batch <= 65536, epochs >= 1, lr > 0";
    return CLR_E_ARG;
  }
  if (batch > n) batch = (int32_t)n;
  const int B = (int)round_up(batch, kTile);
  if (int rc = ensure_batch(st, B, err)) return rc;
  const int M = st->sh.n_members;
  MLPCHK(st->x_d.ensure((size_t)n * st->sh.in_dim));
  MLPCHK(st->y_d.ensure((size_t)n * st->sh.out_dim));
  MLPCHK(st->idx_d.ensure((size_t)M * n));
  MLPCHK(hipMemcpyAsync(st->x_d.p, x, sizeof(float) * n * st->sh.in_dim, hipMemcpyHostToDevice, s));
  MLPCHK(hipMemcpyAsync(st->y_d.p, y, sizeof(float) * n * st->sh.out_dim, hipMemcpyHostToDevice, s));
  if (w) {
    MLPCHK(st->w_d.ensure((size_t)n * st->sh.out_dim));
    MLPCHK(hipMemcpyAsync(st->w_d.p, w, sizeof(float) * n * st->sh.out_dim, hipMemcpyHostToDevice, s));
  }
  const long long steps = n / batch;
  std::vector<int> perm((size_t)M * n);
  std::vector<float> epoch_loss(M, 0.0f);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    // each member sees its own shuffle of the rows
    for (int m = 0; m < M; ++m) {
      int* p = perm.data() + (size_t)m * n;
      std::iota(p, p + n, 0);
      std::mt19937_64 gen(seed + 0x632BE59BD9B4E019ull * (uint64_t)(ep + 1) + (uint64_t)m);
      std::shuffle(p, p + n, gen);
    }
    MLPCHK(hipMemcpyAsync(st->idx_d.p, perm.data(), sizeof(int) * perm.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(fill_f32_kernel, dim3(1), dim3(64), 0, s, st->loss.p, (long long)M, 0.0f);
    for (long long k = 0; k < steps; ++k) {
      forward(st, st->x_d.p, st->idx_d.p, n, k * batch, batch, B, true, s);
      backward(st, st->y_d.p, w ? st->w_d.p : nullptr, st->idx_d.p, n, k * batch, batch, B, s);
      st->step++;
      const float c1 = 1.0f - std::pow(b1, (float)st->step), c2 = 1.0f - std::pow(b2, (float)st->step);
      const long long total = (long long)M * st->sP;
      hipLaunchKernelGGL(adam_kernel, dim3(blocks(total)), dim3(256), 0, s, st->params.p, st->grads.p, st->am.p,
                         st->av.p, total, lr, b1, b2, eps, c1, c2);
      refresh_bf16(st, s);
    }
    MLPCHK(hipGetLastError());
    MLPCHK(hipMemcpyAsync(epoch_loss.data(), st->loss.p, sizeof(float) * M, hipMemcpyDeviceToHost, s));
    MLPCHK(hipStreamSynchronize(s));
  }
  if (loss)
    for (int m = 0; m < M; ++m) loss[m] = epoch_loss[m] / (float)std::max<long long>(steps, 1);
  return 0;
}

namespace {
constexpr int kInferChunk = 65536;  // rows per forward pass of inference
}

int mlp_infer(MlpState* st, const float* x, int64_t n, float* mean, float* stdev, hipStream_t s, std::string& err) {
  if (n < 1) { err = "inference needs n >= 1"; return CLR_E_ARG; }
  const int M = st->sh.n_members, dout = st->sh.out_dim;
  const bool fused = fused_ok(st);
  const int chunk = (int)std::min<int64_t>(kInferChunk, round_up(n, kTile));
  if (!fused)
    if (int rc = ensure_batch(st, chunk, err)) return rc;
  MLPCHK(st->x_d.ensure((size_t)chunk * st->sh.in_dim));
  MLPCHK(st->mean_d.ensure((size_t)chunk * dout));
  MLPCHK(st->std_d.ensure((size_t)chunk * dout));
  for (int64_t r0 = 0; r0 < n; r0 += chunk) {
    const int nv = (int)std::min<int64_t>(chunk, n - r0);
    const int B = (int)round_up(nv, kTile);
    MLPCHK(hipMemcpyAsync(st->x_d.p, x + (size_t)r0 * st->sh.in_dim, sizeof(float) * nv * st->sh.in_dim,
                          hipMemcpyHostToDevice, s));
    if (fused) {
      MLPCHK(launch_fused(st, st->x_d.p, nv, st->mean_d.p, st->std_d.p, s));
    } else {
      forward(st, st->x_d.p, nullptr, 0, 0, nv, B, false, s);
      hipLaunchKernelGGL(ensemble_kernel, dim3(blocks((long long)nv * dout)), dim3(256), 0, s, st->out.p, M, B, dout,
                         nv, st->mean_d.p, st->std_d.p);
    }
    MLPCHK(hipGetLastError());
    MLPCHK(hipMemcpyAsync(mean + (size_t)r0 * dout, st->mean_d.p, sizeof(float) * nv * dout, hipMemcpyDeviceToHost, s));
    if (stdev)
      MLPCHK(hipMemcpyAsync(stdev + (size_t)r0 * dout, st->std_d.p, sizeof(float) * nv * dout, hipMemcpyDeviceToHost,
                            s));
    MLPCHK(hipStreamSynchronize(s));
  }
  return 0;
}

int mlp_time_infer(MlpState* st, int64_t n, int repeats, double* ms, hipStream_t s, std::string& err) {
  if (n < 1 || n > kInferChunk || repeats < 1) {
    err = "timed inference: 1 <= n <= 65536 queries (one device-resident chunk), repeats >= 1";
    return CLR_E_ARG;
  }
  const int M = st->sh.n_members, dout = st->sh.out_dim;
  const int B = (int)round_up(n, kTile);
  const bool fused = fused_ok(st);
  if (!fused)
    if (int rc = ensure_batch(st, B, err)) return rc;
  MLPCHK(st->x_d.ensure((size_t)B * st->sh.in_dim));
  MLPCHK(st->mean_d.ensure((size_t)B * dout));
  MLPCHK(st->std_d.ensure((size_t)B * dout));
  hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks((long long)B * st->sh.in_dim)), dim3(256), 0, s, st->x_d.p,
                     (long long)B * st->sh.in_dim, 0.25f);
  auto pass = [&]() -> hipError_t {
    if (fused) return launch_fused(st, st->x_d.p, n, st->mean_d.p, st->std_d.p, s);
    forward(st, st->x_d.p, nullptr, 0, 0, (int)n, B, false, s);
    hipLaunchKernelGGL(ensemble_kernel, dim3(blocks((long long)n * dout)), dim3(256), 0, s, st->out.p, M, B, dout,
                       (int)n, st->mean_d.p, st->std_d.p);
    return hipGetLastError();
  };
  MLPCHK(pass());  // warm-up
  MLPCHK(hipEventRecord(st->e0, s));
  for (int r = 0; r < repeats; ++r) MLPCHK(pass());
  MLPCHK(hipEventRecord(st->e1, s));
  MLPCHK(hipEventSynchronize(st->e1));
  MLPCHK(hipGetLastError());
  float t = 0.0f;
  MLPCHK(hipEventElapsedTime(&t, st->e0, st->e1));
  *ms = (double)t / repeats;
  return 0;
}

const clr_mlp_shape* mlp_get_shape(const MlpState* st) { return &st->sh; }
bool mlp_fused_ok(const MlpState* st) { return fused_ok(st); }

int mlp_infer_device(MlpState* st, const float* x, int64_t n, float* mean, float* stdev, hipStream_t s,
                     std::string& err) {
  if (!fused_ok(st)) {
    err = "device-resident inference needs the fused shape: hidden <= 256, in_dim <= 32, members * out_dim <= 80";
    return CLR_E_UNSUPPORTED;
  }
  if (n < 1) { err = "inference needs n >= 1"; return CLR_E_ARG; }
  MLPCHK(launch_fused(st, x, n, mean, stdev, s));
  return 0;
}
