// Per-contact warm-start network 16 -> 128 -> 128 -> 1 (SiLU) on gfx950 matrix cores.
//
// Layers 1 and 2 run v_mfma_f32_16x16x16_bf16: BF16 inputs and weights, FP32 accumulation.  Operand
// layouts of that instruction (lane l = 0..63, kg = l / 16):
//   A[i][k] (16 x 16): lane i + 16 * (k / 4), element k % 4
//   B[k][j] (16 x 16): lane j + 16 * (k / 4), element k % 4
//   C[i][j] (16 x 16): lane j + 16 * (i / 4), element i % 4
// Biases, the output layer and all gradients are FP32.  BF16 conversion rounds to nearest even.
//
// Inference (mlp_forward_kernel): each wave takes 16 contacts at a time; W1, W2 (padded rows), biases
// and the output weights stay in LDS for the whole launch, the 16 x 128 hidden tile goes through LDS
// between layers.  Training: mlp_train_forward_kernel stores the pre-activations, mlp_backward_rows_kernel
// forms the per-row gradients (dZ2, dZ1 through an MFMA product with W2^T), mlp_wgrad_kernel the weight
// gradients sum_r dZ[r][o] x[r][i] (MFMA over the batch), mlp_colsum_kernel the bias gradients.
#include <hip/hip_runtime.h>

#include "nmc_internal.h"

namespace nmc {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NF = 16;    // inputs
constexpr int NH = 128;   // hidden width
constexpr int W2S = 136;  // padded LDS row stride (bf16) of W2 and the hidden tile

// FP32 -> BF16, round to nearest even: gfx950's v_cvt_pk_bf16_f32 (one instruction instead of the integer
// rounding sequence; the same bits for every finite input).
__device__ __forceinline__ unsigned short to_bf16(float x) {
  return __builtin_bit_cast(unsigned short, static_cast<__bf16>(x));
}
__device__ __forceinline__ float from_bf16(unsigned short b) { return __uint_as_float(static_cast<unsigned>(b) << 16); }
__device__ __forceinline__ float silu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// LDS hand-off between the lanes of one wave: a wave's LDS operations complete in order, so only
// compiler reordering has to be fenced.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bf16x4 ld4(const unsigned short* p) {
  bf16x4 v;
  v[0] = static_cast<short>(p[0]);
  v[1] = static_cast<short>(p[1]);
  v[2] = static_cast<short>(p[2]);
  v[3] = static_cast<short>(p[3]);
  return v;
}

struct MlpWeights {
  const unsigned short* w1;  // [128][16] bf16
  const float* b1;           // [128]
  const unsigned short* w2;  // [128][128] bf16
  const float* b2;           // [128]
  const float* w3;           // [128]
  const float* b3;           // [1]
  const float* x_mean;       // [16] input normalization: x = (f - mean) * inv_std
  const float* x_inv_std;    // [16]
  float out_scale;           // lambda = y * out_scale
};

struct MlpLds {
  unsigned short* w1;  // 128 x 16
  unsigned short* w2;  // 128 x W2S
  float* b1;
  float* b2;
  float* w3;
  unsigned short* h;   // per wave: 16 x W2S
};

constexpr int MLP_BLOCK = 256;
constexpr size_t MLP_LDS = sizeof(unsigned short) * (NH * NF + NH * W2S + 4 * 16 * W2S) + sizeof(float) * 3 * NH;

__device__ void mlp_load_weights(const MlpWeights& W, MlpLds& S, unsigned char* smem) {
  unsigned char* p = smem;
  S.w1 = reinterpret_cast<unsigned short*>(p); p += sizeof(unsigned short) * NH * NF;
  S.w2 = reinterpret_cast<unsigned short*>(p); p += sizeof(unsigned short) * NH * W2S;
  S.h = reinterpret_cast<unsigned short*>(p); p += sizeof(unsigned short) * 4 * 16 * W2S;
  S.b1 = reinterpret_cast<float*>(p); p += sizeof(float) * NH;
  S.b2 = reinterpret_cast<float*>(p); p += sizeof(float) * NH;
  S.w3 = reinterpret_cast<float*>(p);
  // 16-byte copies: W1 is 256 chunks of 8 BF16, W2 128 rows x 16 chunks (rows padded to W2S in LDS)
  for (int i = threadIdx.x; i < NH * NF / 8; i += blockDim.x)
    reinterpret_cast<uint4*>(S.w1)[i] = reinterpret_cast<const uint4*>(W.w1)[i];
  for (int i = threadIdx.x; i < NH * NH / 8; i += blockDim.x) {
    int row = i >> 4, ch = i & 15;
    *reinterpret_cast<uint4*>(S.w2 + row * W2S + 8 * ch) = reinterpret_cast<const uint4*>(W.w2)[i];
  }
  for (int i = threadIdx.x; i < NH; i += blockDim.x) {
    S.b1[i] = W.b1[i];
    S.b2[i] = W.b2[i];
    S.w3[i] = W.w3[i];
  }
  __syncthreads();
}

// One wave: 16 rows starting at `row0` (rows >= n are zero-padded).  Returns in `y4` the outputs of rows
// 4*kg .. 4*kg+3 on lanes with l % 16 == 0.  z1 / z2 (optional, FP32 [n][128]) receive the
// pre-activations for training.
// The lane's four inputs of tile row0: row row0 + l % 16, features 4 (l / 16) .. +3.
__device__ __forceinline__ float4 mlp_load_input(const float* __restrict__ feat, int row0, int n) {
  const int l = threadIdx.x & 63;
  const int r = row0 + (l & 15);
  return r < n ? *reinterpret_cast<const float4*>(feat + NF * static_cast<long long>(r) + 4 * (l >> 4))
               : make_float4(0, 0, 0, 0);
}

__device__ void mlp_tile(const MlpWeights& W, const MlpLds& S, unsigned short* h, float4 f, int row0, int n,
                         float* __restrict__ z1, float* __restrict__ z2, float y4[4]) {
  const int l = threadIdx.x & 63, col = l & 15, kg = l >> 4;
  // ---- input tile (A operand): row col, features 4kg .. 4kg+3, normalized, BF16
  bf16x4 ax;
  {
    int r = row0 + col;
    float v[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int k = 4 * kg + e;
      float x = r < n ? (v[e] - W.x_mean[k]) * W.x_inv_std[k] : 0.0f;
      ax[e] = static_cast<short>(to_bf16(x));
    }
  }
  // ---- layer 1: 8 output blocks of 16
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    bf16x4 bw = ld4(S.w1 + (nb * 16 + col) * NF + 4 * kg);
    f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ax, bw, c, 0, 0, 0);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int row = 4 * kg + e, o = nb * 16 + col;
      float z = c[e] + S.b1[o];
      if (z1 && row0 + row < n) z1[NH * static_cast<long long>(row0 + row) + o] = z;
      h[row * W2S + o] = to_bf16(silu(z));
    }
  }
  wave_sync();
  // ---- layer 2: K = 128 in 8 steps, 8 output blocks
  f32x4 acc[8];
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) acc[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ks = 0; ks < 8; ++ks) {
    bf16x4 ah = ld4(h + col * W2S + ks * 16 + 4 * kg);
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
      bf16x4 bw = ld4(S.w2 + (nb * 16 + col) * W2S + ks * 16 + 4 * kg);
      acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bw, acc[nb], 0, 0, 0);
    }
  }
  // ---- layer 3 (FP32): y[row] = sum_o silu(z2[row][o]) w3[o] + b3
  float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    int o = nb * 16 + col;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int row = 4 * kg + e;
      float z = acc[nb][e] + S.b2[o];
      if (z2 && row0 + row < n) z2[NH * static_cast<long long>(row0 + row) + o] = z;
      part[e] += silu(z) * S.w3[o];
    }
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    float v = part[e];
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    y4[e] = v + W.b3[0];
  }
  wave_sync();
}

// lambda0[r] = MLP(feat[r]) * out_scale for r < *count (capped at capacity).
__global__ void __launch_bounds__(MLP_BLOCK) mlp_forward_kernel(MlpWeights W, const float* __restrict__ feat,
                                                                const int* __restrict__ count, int capacity,
                                                                float* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  MlpLds S;
  mlp_load_weights(W, S, smem);
  int n = *count;
  if (n > capacity) n = capacity;
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  unsigned short* h = S.h + wave * 16 * W2S;
  const int tiles = (n + 15) / 16;
  const int stride = gridDim.x * 4;
  int t = blockIdx.x * 4 + wave;
  float4 f = t < tiles ? mlp_load_input(feat, t * 16, n) : make_float4(0, 0, 0, 0);
  for (; t < tiles; t += stride) {
    const float4 cur = f;
    if (t + stride < tiles) f = mlp_load_input(feat, (t + stride) * 16, n);  // next tile in flight
    float y4[4];
    mlp_tile(W, S, h, cur, t * 16, n, nullptr, nullptr, y4);
    if ((l & 15) == 0) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        int r = t * 16 + 4 * (l >> 4) + e;
        if (r < n) out[r] = y4[e] * W.out_scale;
      }
    }
  }
}

// Training forward over n rows: y, and the pre-activations z1, z2 ([n][128] FP32).
__global__ void __launch_bounds__(MLP_BLOCK) mlp_train_forward_kernel(MlpWeights W, const float* __restrict__ feat,
                                                                      int n, float* __restrict__ y,
                                                                      float* __restrict__ z1, float* __restrict__ z2) {
  extern __shared__ __align__(16) unsigned char smem[];
  MlpLds S;
  mlp_load_weights(W, S, smem);
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  unsigned short* h = S.h + wave * 16 * W2S;
  const int tiles = (n + 15) / 16;
  for (int t = blockIdx.x * 4 + wave; t < tiles; t += gridDim.x * 4) {
    float y4[4];
    mlp_tile(W, S, h, mlp_load_input(feat, t * 16, n), t * 16, n, z1, z2, y4);
    if ((l & 15) == 0) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        int r = t * 16 + 4 * (l >> 4) + e;
        if (r < n) y[r] = y4[e];
      }
    }
  }
}

// Per-row gradients of L = 0.5 * mean((y - t)^2):
//   dy = (y - t) / n; dz2 = dy w3 silu'(z2); dz1 = (dz2 W2) silu'(z1)   (the product with BF16 MFMA)
// Outputs (row-major, [n][.]): xb = BF16 normalized inputs [16], h1 = BF16 silu(z1) [128],
// dz2b / dz1b = BF16 [128] (weight-gradient operands), dz2f / dz1f / g3 = FP32 [128] (bias gradients and
// dw3 terms dy * silu(z2)), dyf = FP32 [n].
__global__ void __launch_bounds__(MLP_BLOCK) mlp_backward_rows_kernel(
    MlpWeights W, const float* __restrict__ feat, const float* __restrict__ target, int n, const float* __restrict__ y,
    const float* __restrict__ z1, const float* __restrict__ z2, unsigned short* __restrict__ xb,
    unsigned short* __restrict__ h1, unsigned short* __restrict__ dz2b, unsigned short* __restrict__ dz1b,
    float* __restrict__ dz2f, float* __restrict__ dz1f, float* __restrict__ g3, float* __restrict__ dyf) {
  extern __shared__ __align__(16) unsigned char smem[];
  MlpLds S;
  mlp_load_weights(W, S, smem);
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, col = l & 15, kg = l >> 4;
  unsigned short* tile = S.h + wave * 16 * W2S;  // dz2 rows of the tile (BF16)
  const int tiles = (n + 15) / 16;
  const float inv_n = 1.0f / static_cast<float>(n);
  for (int t = blockIdx.x * 4 + wave; t < tiles; t += gridDim.x * 4) {
    const int row0 = t * 16;
    // dz2 for the tile, one row per 4 lanes: lane l covers row l / 4, outputs (l % 4) * 32 .. +31
    {
      const int row = l >> 2, r = row0 + row, ob = (l & 3) * 32;
      float dy = r < n ? (y[r] - target[r]) * inv_n : 0.0f;
      if (r < n && (l & 3) == 0) dyf[r] = dy;
      for (int o = ob; o < ob + 32; ++o) {
        float z = r < n ? z2[NH * static_cast<long long>(r) + o] : 0.0f;
        float g = dy * S.w3[o] * silu_grad(z);
        unsigned short gb = to_bf16(g);
        tile[row * W2S + o] = gb;
        if (r < n) {
          long long i = NH * static_cast<long long>(r) + o;
          dz2b[i] = gb;
          dz2f[i] = g;
          g3[i] = dy * silu(z);
        }
      }
    }
    wave_sync();
    // dh1[row][i] = sum_o dz2[row][o] W2[o][i]: A = dz2 tile (16 x 128), B[k=o][j=i] = W2[o][i]
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
      f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
        bf16x4 a = ld4(tile + col * W2S + ks * 16 + 4 * kg);
        bf16x4 b;
#pragma unroll
        for (int e = 0; e < 4; ++e) b[e] = static_cast<short>(S.w2[(ks * 16 + 4 * kg + e) * W2S + nb * 16 + col]);
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        int row = 4 * kg + e, r = row0 + row, i = nb * 16 + col;
        if (r < n) {
          long long idx = NH * static_cast<long long>(r) + i;
          float z = z1[idx];
          float g = c[e] * silu_grad(z);
          dz1f[idx] = g;
          dz1b[idx] = to_bf16(g);
          h1[idx] = to_bf16(silu(z));
        }
      }
    }
    // BF16 normalized inputs
    if (l < 16 && row0 + l < n) {
      int r = row0 + l;
      for (int k = 0; k < NF; ++k)
        xb[NF * static_cast<long long>(r) + k] = to_bf16((feat[NF * static_cast<long long>(r) + k] - W.x_mean[k]) * W.x_inv_std[k]);
    }
    wave_sync();
  }
}

// g[o][i] = sum_r a[r][o] b[r][i] over r < n (a: [n][M] BF16, b: [n][N] BF16, M, N multiples of 16);
// one wave per 16 x 16 output tile, K = the batch in steps of 16 rows.
__global__ void __launch_bounds__(64) mlp_wgrad_kernel(const unsigned short* __restrict__ a, int M,
                                                       const unsigned short* __restrict__ b, int N, int n,
                                                       float* __restrict__ g) {
  const int l = threadIdx.x, col = l & 15, kg = l >> 4;
  const int tm = blockIdx.x / (N / 16), tn = blockIdx.x % (N / 16);
  f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int r0 = 0; r0 < n; r0 += 16) {
    bf16x4 av, bv;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int r = r0 + 4 * kg + e;
      av[e] = r < n ? static_cast<short>(a[static_cast<long long>(M) * r + tm * 16 + col]) : static_cast<short>(0);
      bv[e] = r < n ? static_cast<short>(b[static_cast<long long>(N) * r + tn * 16 + col]) : static_cast<short>(0);
    }
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, c, 0, 0, 0);
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) g[(tm * 16 + 4 * kg + e) * N + tn * 16 + col] = c[e];
}

// s[c] = sum_r x[r][c] for r < n (x: [n][cols] FP32), one workgroup per column.
__global__ void __launch_bounds__(256) mlp_colsum_kernel(const float* __restrict__ x, int cols, int n,
                                                         float* __restrict__ s) {
  __shared__ float red[4];
  const int c = blockIdx.x;
  float v = 0.0f;
  for (int r = threadIdx.x; r < n; r += blockDim.x) v += x[static_cast<long long>(cols) * r + c];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) s[c] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace nmc

using namespace nmc;

static MlpWeights make_weights(const NmcMlp* w) {
  MlpWeights W;
  W.w1 = static_cast<const unsigned short*>(w->w1_bf16);
  W.b1 = w->b1;
  W.w2 = static_cast<const unsigned short*>(w->w2_bf16);
  W.b2 = w->b2;
  W.w3 = w->w3;
  W.b3 = w->b3;
  W.x_mean = w->x_mean;
  W.x_inv_std = w->x_inv_std;
  W.out_scale = w->out_scale;
  return W;
}

static bool weights_ok(const NmcMlp* w) {
  return w && w->w1_bf16 && w->b1 && w->w2_bf16 && w->b2 && w->w3 && w->b3 && w->x_mean && w->x_inv_std;
}

extern "C" int nmc_mlp_forward(const NmcMlp* w, const float* feat, const int* count, int capacity, float* lambda0,
                               nmc_stream_t stream) {
  if (!weights_ok(w) || !feat || !count || !lambda0) return nmc::fail("nmc_mlp_forward: null argument");
  if (capacity <= 0) return 0;
  int tiles = (capacity + 15) / 16;
  int blocks = (tiles + 3) / 4;
  if (blocks > 512) blocks = 512;  // two workgroups per CU (LDS-bound) resident at once; the rest is strided
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(MLP_BLOCK), MLP_LDS, reinterpret_cast<hipStream_t>(stream),
                     make_weights(w), feat, count, capacity, lambda0);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : nmc::hip_fail(e, "nmc_mlp_forward launch");
}

extern "C" int nmc_mlp_train_grad(const NmcMlp* w, const float* feat, const float* target, int n,
                                  const NmcMlpTrainBuffers* tb, NmcMlpGrads* g, nmc_stream_t stream) {
  if (!weights_ok(w) || !feat || !target || !tb || !g) return nmc::fail("nmc_mlp_train_grad: null argument");
  if (n <= 0) return nmc::fail("nmc_mlp_train_grad: empty batch");
  if (!tb->y || !tb->z1 || !tb->z2 || !tb->xb || !tb->h1 || !tb->dz2b || !tb->dz1b || !tb->dz2f || !tb->dz1f ||
      !tb->g3 || !tb->dy || !g->w1 || !g->b1 || !g->w2 || !g->b2 || !g->w3 || !g->b3)
    return nmc::fail("nmc_mlp_train_grad: null buffer");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  MlpWeights W = make_weights(w);
  int tiles = (n + 15) / 16;
  int blocks = (tiles + 3) / 4;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(mlp_train_forward_kernel, dim3(blocks), dim3(MLP_BLOCK), MLP_LDS, s, W, feat, n, tb->y, tb->z1,
                     tb->z2);
  hipLaunchKernelGGL(mlp_backward_rows_kernel, dim3(blocks), dim3(MLP_BLOCK), MLP_LDS, s, W, feat, target, n, tb->y,
                     tb->z1, tb->z2, static_cast<unsigned short*>(tb->xb), static_cast<unsigned short*>(tb->h1),
                     static_cast<unsigned short*>(tb->dz2b), static_cast<unsigned short*>(tb->dz1b), tb->dz2f, tb->dz1f,
                     tb->g3, tb->dy);
  hipLaunchKernelGGL(mlp_wgrad_kernel, dim3((NH / 16) * (NH / 16)), dim3(64), 0, s,
                     static_cast<const unsigned short*>(tb->dz2b), NH, static_cast<const unsigned short*>(tb->h1), NH, n,
                     g->w2);
  hipLaunchKernelGGL(mlp_wgrad_kernel, dim3((NH / 16) * (NF / 16)), dim3(64), 0, s,
                     static_cast<const unsigned short*>(tb->dz1b), NH, static_cast<const unsigned short*>(tb->xb), NF, n,
                     g->w1);
  hipLaunchKernelGGL(mlp_colsum_kernel, dim3(NH), dim3(256), 0, s, tb->dz2f, NH, n, g->b2);
  hipLaunchKernelGGL(mlp_colsum_kernel, dim3(NH), dim3(256), 0, s, tb->dz1f, NH, n, g->b1);
  hipLaunchKernelGGL(mlp_colsum_kernel, dim3(NH), dim3(256), 0, s, tb->g3, NH, n, g->w3);
  hipLaunchKernelGGL(mlp_colsum_kernel, dim3(1), dim3(256), 0, s, tb->dy, 1, n, g->b3);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : nmc::hip_fail(e, "nmc_mlp_train_grad launch");
}
