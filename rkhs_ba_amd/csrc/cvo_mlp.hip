// Learned pose proposer on gfx950 (no counterpart in the reference): a point
// set encoder shared by both clouds and a head proposing K rigid transforms
// with confidences, trained and run with hand-written BF16 MFMA kernels.
//
// Network (per pair; a cloud is npts points of 8 channels: xyz in metres and
// 5 features):
//   frame   per cloud: centroid c and principal axes V = [e1, e2, e1 x e2]
//           (covariance eigenvectors, descending eigenvalues; e1, e2 signed so
//           the third moment along them is >= 0)
//   input   x = [V^T (xyz - c) / 10, features] as BF16, padded to 16
//   encoder H1 = relu(x W1 + b1) (64), H2 = relu(H1 W2 + b2) (128),
//           H3 = relu(H2 W3 + b3) (256), g = max over points of H3
//   head    z = [g_src, g_tgt] (512), z1 = relu(z W4 + b4) (256),
//           o = z1 W5 + b5 (64) = K = 8 x (quaternion w,x,y,z; translation
//           residual / 10) then 8 confidence logits
//   proposal k: R_k = V_tgt rot(q_k / |q_k|) V_src^T (the quaternion lives in
//           the canonical frames), t_k = c_tgt - R_k c_src + 10 o_t,k
//           (the align-init convention: target ~ R source + t)
// Weights: FP32 master copies (Adam), BF16 copies feed the MFMA tiles;
// activations between layers are stored in BF16; every product accumulates
// in FP32.  Training loss per pair (relaxed winner-takes-all), q* the label's
// rotation in the canonical frames (V_tgt^T R* V_src):
//   L_k = |q^_k - s_k q*|^2 + |o_t,k - d*|^2, s_k = sign(q^_k . q*),
//   k* = argmin L_k, loss = sum_k w_k L_k + 0.1 CE(conf, k*),
//   w_k* = 0.95, others 0.05 / 7;  batch loss = mean over pairs.
// Backward through the max-pool only touches each cloud's argmax points: one
// row per (cloud, channel), so the encoder's backward GEMMs run over
// 2 * 256 * pairs rows instead of every point.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "cvo_internal.h"

namespace cvo {

constexpr int kMlpK = 8;          // proposals
constexpr int kMlpOut = 64;       // 8 x 7 + 8
constexpr int kMlpIn = 8;         // channels per point
constexpr int kMlpInPad = 16;     // padded to one MFMA k-step
constexpr int kH1 = 64, kH2 = 128, kH3 = 256, kZ = 512, kZ1 = 256;
constexpr float kMlpScale = 10.0f;

// parameter offsets in the flat FP32 vector
constexpr int oW1 = 0, oB1 = oW1 + kMlpIn * kH1;
constexpr int oW2 = oB1 + kH1, oB2 = oW2 + kH1 * kH2;
constexpr int oW3 = oB2 + kH2, oB3 = oW3 + kH2 * kH3;
constexpr int oW4 = oB3 + kH3, oB4 = oW4 + kZ * kZ1;
constexpr int oW5 = oB4 + kZ1, oB5 = oW5 + kZ1 * kMlpOut;
constexpr int kMlpParams = oB5 + kMlpOut;

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// ---------------------------------------------------------------------------
// BF16 MFMA GEMM: C[M x N] = sum_k A[m, k] B[k, n] with strided operands
// (transposes are stride swaps).  64 x 64 block tile, four waves of 32 x 32
// (2 x 2 tiles of v_mfma_f32_16x16x16_bf16), k-steps of 16 staged in LDS.
// Epilogue: optional bias (per n), ReLU, mask (zero where a BF16 matrix of the
// same shape is zero), FP32 store or accumulate, BF16 store, or a max-pool over
// row groups into packed (value, ~row) keys.
// ---------------------------------------------------------------------------

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

struct GemmArgs {
  const uint16_t* A;
  int64_t a_sm, a_sk;
  const uint16_t* B;
  int64_t b_sk, b_sn;
  int M, N, K;
  float* C;            // optional FP32 output
  int64_t ldc;
  int accumulate;      // C += result
  uint16_t* Cb;        // optional BF16 output
  int64_t ldcb;
  const float* bias;   // optional, per n
  int relu;
  const uint16_t* mask;  // optional, M x N BF16, zero -> output zero
  int64_t ldm;
  unsigned long long* pool;  // optional: pool[(m / group) * N + n] = max key
  int group;                 // rows per pool group (multiple of 64)
};

__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
  __shared__ uint16_t As[64][16 + 4];
  __shared__ uint16_t Bs[16][64 + 4];
  __shared__ unsigned long long s_pool[2][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wm = wave >> 1, wn = wave & 1;
  float4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (float4_t){0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    {  // A tile: 64 rows x 16 k, four consecutive k per thread
      const int r = tid >> 2, kq = (tid & 3) * 4;
      const int m = m0 + r;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = k0 + kq + i;
        As[r][kq + i] = (m < g.M && k < g.K) ? g.A[(int64_t)m * g.a_sm + (int64_t)k * g.a_sk] : (uint16_t)0;
      }
    }
    {  // B tile: 16 k x 64 n, four consecutive n per thread
      const int k = tid >> 4, nq = (tid & 15) * 4;
      const int kk = k0 + k;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int n = n0 + nq + i;
        Bs[k][nq + i] = (kk < g.K && n < g.N) ? g.B[(int64_t)kk * g.b_sk + (int64_t)n * g.b_sn] : (uint16_t)0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int r = wm * 32 + i * 16 + (lane & 15);
      const int kb = (lane >> 4) * 4;
      short4_t a;
      a[0] = (short)As[r][kb + 0];
      a[1] = (short)As[r][kb + 1];
      a[2] = (short)As[r][kb + 2];
      a[3] = (short)As[r][kb + 3];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int c = wn * 32 + j * 16 + (lane & 15);
        short4_t b;
        b[0] = (short)Bs[kb + 0][c];
        b[1] = (short)Bs[kb + 1][c];
        b[2] = (short)Bs[kb + 2][c];
        b[3] = (short)Bs[kb + 3][c];
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // epilogue: lane holds rows (lane >> 4) * 4 + v, column lane & 15 of each tile
  unsigned long long colmax[2] = {0ull, 0ull};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm * 32 + i * 16 + (lane >> 4) * 4 + v;
        if (m >= g.M || n >= g.N) continue;
        float x = acc[i][j][v];
        if (g.bias) x += g.bias[n];
        if (g.relu) x = x > 0.0f ? x : 0.0f;  // never -0 (pool keys use the bits)
        if (g.mask && g.mask[(int64_t)m * g.ldm + n] == 0) x = 0.0f;
        if (g.C) {
          float* cp = g.C + (int64_t)m * g.ldc + n;
          *cp = g.accumulate ? *cp + x : x;
        }
        if (g.Cb) g.Cb[(int64_t)m * g.ldcb + n] = f2bf(x);
        if (g.pool) {
          // x >= 0 (ReLU): float bits order like the values; ties -> lower row
          const unsigned long long key =
              ((unsigned long long)__float_as_uint(x) << 32) | (unsigned long long)(0xFFFFFFFFu - (uint32_t)(m % g.group));
          colmax[j] = colmax[j] > key ? colmax[j] : key;
        }
      }
    }
  if (g.pool) {
    // 64 rows of the block belong to one group (group % 64 == 0): reduce the
    // 4 lane groups, the 2 row tiles (already in colmax) and the 2 waves of
    // each column half, then one atomic per column
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      unsigned long long k = colmax[j];
      unsigned long long o = __shfl_xor(k, 16, 64);
      k = k > o ? k : o;
      o = __shfl_xor(k, 32, 64);
      k = k > o ? k : o;
      colmax[j] = k;
    }
    if (lane < 16) {
      s_pool[wm][wn * 32 + 0 * 16 + lane] = colmax[0];
      s_pool[wm][wn * 32 + 1 * 16 + lane] = colmax[1];
    }
    __syncthreads();
    if (tid < 64) {
      const int n = n0 + tid;
      const unsigned long long a = s_pool[0][tid], b = s_pool[1][tid];
      const unsigned long long k = a > b ? a : b;
      if (n < g.N && m0 < g.M) atomicMax(&g.pool[(int64_t)(m0 / g.group) * g.N + n], k);
    }
  }
}

// ---------------------------------------------------------------------------
// Fused point encoder for inference: the three per-point layers (16 -> 64 ->
// 128 -> 256, bias and ReLU, BF16 between layers as in the layer-by-layer
// path) and the max-pool over the cloud in one kernel.  One workgroup per
// cloud: the weights are staged once, transposed (n rows of k), so both MFMA
// operands are 16-byte LDS loads; the cloud's points go through in tiles of
// 128, each of the four waves owning 32 rows of a tile from the input to the
// pool, so the activations never leave LDS.  Layer 1 (k = 16) runs on
// v_mfma_f32_16x16x16_bf16, layers 2 and 3 on v_mfma_f32_16x16x32_bf16.  The
// pool keeps, per column, the largest (value bits, ~row) key of the cloud
// (ties to the lower row), folded in registers and reduced across lanes and
// waves in fixed order: one plain store per cloud and channel.
// ---------------------------------------------------------------------------

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

constexpr int kFT = 128;  // points per tile
struct alignas(16) EncLds {
  uint16_t w1t[kH1][kMlpInPad + 8];   // [n][k], rows padded to keep 16-byte alignment
  uint16_t w2t[kH2][kH1 + 8];
  uint16_t w3t[kH3][kH2 + 8];
  uint16_t x[kFT][kMlpInPad + 8];
  uint16_t h1[kFT][kH1 + 8];
  uint16_t h2[kFT][kH2 + 8];
  unsigned long long pool[4][kH3];
};

__device__ __forceinline__ bf16x8_t lds8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const uint4*>(p));
}
__device__ __forceinline__ short4_t lds4(const uint16_t* p) {
  return __builtin_bit_cast(short4_t, *reinterpret_cast<const uint2*>(p));
}

__global__ __launch_bounds__(256) void encoder_fused_kernel(const uint16_t* __restrict__ X, int npts,
                                                            const uint16_t* __restrict__ W1b,
                                                            const uint16_t* __restrict__ W2b,
                                                            const uint16_t* __restrict__ W3b,
                                                            const float* __restrict__ w,
                                                            unsigned long long* __restrict__ pool) {
  __shared__ EncLds L;  // 155 KB of the CU's 160 KB
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int cloud = blockIdx.x;
  const float* b1 = w + oB1;
  const float* b2 = w + oB2;
  const float* b3 = w + oB3;
  for (int i = tid; i < kMlpInPad * kH1; i += 256) L.w1t[i % kH1][i / kH1] = W1b[i];
  for (int i = tid; i < kH1 * kH2; i += 256) L.w2t[i % kH2][i / kH2] = W2b[i];
  for (int i = tid; i < kH2 * kH3; i += 256) L.w3t[i % kH3][i / kH3] = W3b[i];
  const int col = lane & 15, kq = lane >> 4;  // MFMA operand column / k group
  // running pool keys of this lane: column t * 16 + col, the wave's rows
  unsigned long long best[kH3 / 16];
#pragma unroll
  for (int t = 0; t < kH3 / 16; ++t) best[t] = 0ull;
  const uint16_t* Xc = X + (size_t)cloud * npts * kMlpInPad;
  for (int t0 = 0; t0 < npts; t0 += kFT) {
    __syncthreads();  // the weights are staged / the previous tile is done with x
    {
      const int r = tid >> 1, h = (tid & 1) * 8;
      if (t0 + r < npts)
        *reinterpret_cast<uint4*>(&L.x[r][h]) =
            *reinterpret_cast<const uint4*>(Xc + (size_t)(t0 + r) * kMlpInPad + h);
    }
    __syncthreads();
    const int rw = wave * 32;  // this wave's rows of the tile: rw .. rw + 31
    // npts is a multiple of 64: a wave's 32 rows are all in the cloud or
    // none; a wave past the end waits for the others at the next barrier
    if (t0 + rw >= npts) continue;
    // layer 1: 16 -> 64
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
      const short4_t a = lds4(&L.x[rw + rt * 16 + col][kq * 4]);
#pragma unroll
      for (int ct = 0; ct < kH1 / 16; ++ct) {
        const short4_t b = lds4(&L.w1t[ct * 16 + col][kq * 4]);
        float4_t acc = (float4_t){0.0f, 0.0f, 0.0f, 0.0f};
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        const int c = ct * 16 + col;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float x = acc[v] + b1[c];
          x = x > 0.0f ? x : 0.0f;
          L.h1[rw + rt * 16 + kq * 4 + v][c] = f2bf(x);
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): the wave's own h1 rows are written
    __builtin_amdgcn_wave_barrier();
    // layer 2: 64 -> 128
#pragma unroll 2
    for (int ct = 0; ct < kH2 / 16; ++ct) {
      float4_t acc[2] = {(float4_t){0.0f, 0.0f, 0.0f, 0.0f}, (float4_t){0.0f, 0.0f, 0.0f, 0.0f}};
#pragma unroll
      for (int ks = 0; ks < kH1 / 32; ++ks) {
        const bf16x8_t b = lds8(&L.w2t[ct * 16 + col][ks * 32 + kq * 8]);
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) {
          const bf16x8_t a = lds8(&L.h1[rw + rt * 16 + col][ks * 32 + kq * 8]);
          acc[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[rt], 0, 0, 0);
        }
      }
      const int c = ct * 16 + col;
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float x = acc[rt][v] + b2[c];
          x = x > 0.0f ? x : 0.0f;
          L.h2[rw + rt * 16 + kq * 4 + v][c] = f2bf(x);
        }
    }
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_wave_barrier();
    // layer 3: 128 -> 256, four column tiles at a time, folded into the pool
    for (int cc = 0; cc < kH3 / 16; cc += 4) {
      float4_t acc[2][4];
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[rt][j] = (float4_t){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < kH2 / 32; ++ks) {
        const bf16x8_t a0 = lds8(&L.h2[rw + col][ks * 32 + kq * 8]);
        const bf16x8_t a1 = lds8(&L.h2[rw + 16 + col][ks * 32 + kq * 8]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const bf16x8_t b = lds8(&L.w3t[(cc + j) * 16 + col][ks * 32 + kq * 8]);
          acc[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b, acc[0][j], 0, 0, 0);
          acc[1][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b, acc[1][j], 0, 0, 0);
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = (cc + j) * 16 + col;
        const float bias = b3[c];
        unsigned long long m = best[cc + j];
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            float x = acc[rt][j][v] + bias;
            x = x > 0.0f ? x : 0.0f;  // never -0: the key orders by the bits
            const uint32_t row = (uint32_t)(t0 + rw + rt * 16 + kq * 4 + v);
            const unsigned long long key =
                ((unsigned long long)__float_as_uint(x) << 32) | (unsigned long long)(0xFFFFFFFFu - row);
            m = m > key ? m : key;
          }
        best[cc + j] = m;
      }
    }
  }
  // lanes sharing a column (col, +16, +32, +48), then the four waves in order
#pragma unroll
  for (int t = 0; t < kH3 / 16; ++t) {
    unsigned long long k = best[t];
    unsigned long long o = __shfl_xor(k, 16, 64);
    k = k > o ? k : o;
    o = __shfl_xor(k, 32, 64);
    k = k > o ? k : o;
    if (lane < 16) L.pool[wave][t * 16 + lane] = k;
  }
  __syncthreads();
  {
    const int c = tid;  // 256 channels, one a thread
    unsigned long long k = L.pool[0][c];
    for (int wv = 1; wv < 4; ++wv) k = k > L.pool[wv][c] ? k : L.pool[wv][c];
    pool[(size_t)cloud * kH3 + c] = k;
  }
}

static void gemm(const GemmArgs& a, hipStream_t s) {
  dim3 grid((a.N + 63) / 64, (a.M + 63) / 64);
  hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, s, a);
}

static GemmArgs gemm_args(const uint16_t* A, int64_t a_sm, int64_t a_sk, const uint16_t* B, int64_t b_sk,
                          int64_t b_sn, int M, int N, int K) {
  GemmArgs g;
  std::memset(&g, 0, sizeof(g));
  g.A = A;
  g.a_sm = a_sm;
  g.a_sk = a_sk;
  g.B = B;
  g.b_sk = b_sk;
  g.b_sn = b_sn;
  g.M = M;
  g.N = N;
  g.K = K;
  g.group = 64;
  return g;
}

// ---------------------------------------------------------------------------
// Small kernels
// ---------------------------------------------------------------------------

// Eigen-decomposition of a symmetric 3x3 matrix (cyclic Jacobi, FP64):
// columns of V are the eigenvectors of ev.
__device__ inline void eig3_sym(double A[3][3], double ev[3], double V[3][3]) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) V[r][c] = (r == c) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 16; ++sweep) {
    const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    if (off < 1e-30 * (A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2] + 1e-300)) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (A[p][q] == 0.0) continue;
        const double th = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // A <- A J
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {  // A <- J^T A
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {  // V <- V J
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  for (int k = 0; k < 3; ++k) ev[k] = A[k][k];
}

// Per cloud: centroid, principal frame and the BF16 input rows
// [V^T (xyz - c) / 10, features, 0 x 8].  frame[12 * cloud] = c (3), V (3x3,
// row-major, columns e1, e2, e3).  Sums in FP32 with fixed-order block trees.
__global__ __launch_bounds__(256) void mlp_prep_kernel(const float* __restrict__ pts, int npts,
                                                       uint16_t* __restrict__ X, float* __restrict__ frame) {
  const int cl = blockIdx.x, tid = threadIdx.x;
  const float* P = pts + (int64_t)cl * npts * kMlpIn;
  __shared__ float s_sum[6][256];
  __shared__ float s_c[3];
  __shared__ float s_V[9];
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for (int i = tid; i < npts; i += 256) {
    sx += P[i * kMlpIn + 0];
    sy += P[i * kMlpIn + 1];
    sz += P[i * kMlpIn + 2];
  }
  s_sum[0][tid] = sx;
  s_sum[1][tid] = sy;
  s_sum[2][tid] = sz;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
      for (int d = 0; d < 3; ++d) s_sum[d][tid] += s_sum[d][tid + w];
    __syncthreads();
  }
  if (tid == 0)
    for (int d = 0; d < 3; ++d) s_c[d] = s_sum[d][0] / (float)npts;
  __syncthreads();
  const float cx = s_c[0], cy = s_c[1], cz = s_c[2];
  float a[6] = {0, 0, 0, 0, 0, 0};
  for (int i = tid; i < npts; i += 256) {
    const float x = P[i * kMlpIn + 0] - cx, y = P[i * kMlpIn + 1] - cy, z = P[i * kMlpIn + 2] - cz;
    a[0] += x * x; a[1] += x * y; a[2] += x * z; a[3] += y * y; a[4] += y * z; a[5] += z * z;
  }
  for (int d = 0; d < 6; ++d) s_sum[d][tid] = a[d];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
      for (int d = 0; d < 6; ++d) s_sum[d][tid] += s_sum[d][tid + w];
    __syncthreads();
  }
  if (tid == 0) {
    double A[3][3] = {{s_sum[0][0], s_sum[1][0], s_sum[2][0]},
                      {s_sum[1][0], s_sum[3][0], s_sum[4][0]},
                      {s_sum[2][0], s_sum[4][0], s_sum[5][0]}};
    double ev[3], V[3][3];
    eig3_sym(A, ev, V);
    int o[3] = {0, 1, 2};  // descending eigenvalues
    for (int x = 0; x < 2; ++x)
      for (int y = 0; y < 2 - x; ++y)
        if (ev[o[y]] < ev[o[y + 1]]) {
          const int t = o[y];
          o[y] = o[y + 1];
          o[y + 1] = t;
        }
    for (int r = 0; r < 3; ++r) {
      s_V[r * 3 + 0] = (float)V[r][o[0]];
      s_V[r * 3 + 1] = (float)V[r][o[1]];
    }
  }
  __syncthreads();
  // third moments along e1, e2 fix their signs
  float m3[2] = {0.0f, 0.0f};
  for (int i = tid; i < npts; i += 256) {
    const float x = P[i * kMlpIn + 0] - cx, y = P[i * kMlpIn + 1] - cy, z = P[i * kMlpIn + 2] - cz;
    for (int k = 0; k < 2; ++k) {
      const float u = x * s_V[0 * 3 + k] + y * s_V[1 * 3 + k] + z * s_V[2 * 3 + k];
      m3[k] += u * u * u;
    }
  }
  s_sum[0][tid] = m3[0];
  s_sum[1][tid] = m3[1];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
      for (int d = 0; d < 2; ++d) s_sum[d][tid] += s_sum[d][tid + w];
    __syncthreads();
  }
  if (tid == 0) {
    for (int k = 0; k < 2; ++k)
      if (s_sum[k][0] < 0.0f)
        for (int r = 0; r < 3; ++r) s_V[r * 3 + k] = -s_V[r * 3 + k];
    // e3 = e1 x e2
    s_V[0 * 3 + 2] = s_V[1 * 3 + 0] * s_V[2 * 3 + 1] - s_V[2 * 3 + 0] * s_V[1 * 3 + 1];
    s_V[1 * 3 + 2] = s_V[2 * 3 + 0] * s_V[0 * 3 + 1] - s_V[0 * 3 + 0] * s_V[2 * 3 + 1];
    s_V[2 * 3 + 2] = s_V[0 * 3 + 0] * s_V[1 * 3 + 1] - s_V[1 * 3 + 0] * s_V[0 * 3 + 1];
    float* fr = frame + 12 * cl;
    fr[0] = cx;
    fr[1] = cy;
    fr[2] = cz;
    for (int k = 0; k < 9; ++k) fr[3 + k] = s_V[k];
  }
  __syncthreads();
  uint16_t* Xc = X + (int64_t)cl * npts * kMlpInPad;
  for (int i = tid; i < npts; i += 256) {
    const float* p = P + i * kMlpIn;
    const float x = p[0] - cx, y = p[1] - cy, z = p[2] - cz;
    uint16_t* xo = Xc + i * kMlpInPad;
    for (int k = 0; k < 3; ++k)
      xo[k] = f2bf((x * s_V[0 * 3 + k] + y * s_V[1 * 3 + k] + z * s_V[2 * 3 + k]) / kMlpScale);
    for (int f = 3; f < kMlpIn; ++f) xo[f] = f2bf(p[f]);
    for (int f = kMlpIn; f < kMlpInPad; ++f) xo[f] = 0;
  }
}

__global__ void fill_u64_kernel(unsigned long long* p, int64_t n, unsigned long long v) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Pooled keys -> z (BF16, [src | tgt] per pair), pooled values (FP32) and the
// argmax point of every (cloud, channel).
__global__ void mlp_unpack_pool_kernel(const unsigned long long* __restrict__ pool, int n_pairs,
                                       uint16_t* __restrict__ z, float* __restrict__ gval,
                                       int32_t* __restrict__ garg) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;  // cloud * 256 + ch
  if (i >= (int64_t)n_pairs * 2 * kH3) return;
  const int cloud = (int)(i / kH3), ch = (int)(i % kH3);
  const unsigned long long k = pool[i];
  const float v = __uint_as_float((uint32_t)(k >> 32));
  gval[i] = v;
  garg[i] = (int32_t)(0xFFFFFFFFu - (uint32_t)(k & 0xFFFFFFFFull));
  const int pair = cloud >> 1, half = cloud & 1;
  z[(int64_t)pair * kZ + half * kH3 + ch] = f2bf(v);
}

__device__ inline void quat_to_rot_f(const float q[4], float R[9]) {
  const float w = q[0], x = q[1], y = q[2], zq = q[3];
  R[0] = 1 - 2 * (y * y + zq * zq); R[1] = 2 * (x * y - w * zq);     R[2] = 2 * (x * zq + w * y);
  R[3] = 2 * (x * y + w * zq);     R[4] = 1 - 2 * (x * x + zq * zq); R[5] = 2 * (y * zq - w * x);
  R[6] = 2 * (x * zq - w * y);     R[7] = 2 * (y * zq + w * x);     R[8] = 1 - 2 * (x * x + y * y);
}

// Unit quaternion (w >= 0) of a rotation (Shepperd's method).
__device__ inline void rot_to_quat_f(const float R[9], float q[4]) {
  const float tr = R[0] + R[4] + R[8];
  if (tr > 0.0f) {
    const float s = sqrtf(tr + 1.0f) * 2.0f;
    q[0] = 0.25f * s; q[1] = (R[7] - R[5]) / s; q[2] = (R[2] - R[6]) / s; q[3] = (R[3] - R[1]) / s;
  } else if (R[0] > R[4] && R[0] > R[8]) {
    const float s = sqrtf(1.0f + R[0] - R[4] - R[8]) * 2.0f;
    q[0] = (R[7] - R[5]) / s; q[1] = 0.25f * s; q[2] = (R[1] + R[3]) / s; q[3] = (R[2] + R[6]) / s;
  } else if (R[4] > R[8]) {
    const float s = sqrtf(1.0f + R[4] - R[0] - R[8]) * 2.0f;
    q[0] = (R[2] - R[6]) / s; q[1] = (R[1] + R[3]) / s; q[2] = 0.25f * s; q[3] = (R[5] + R[7]) / s;
  } else {
    const float s = sqrtf(1.0f + R[8] - R[0] - R[4]) * 2.0f;
    q[0] = (R[3] - R[1]) / s; q[1] = (R[2] + R[6]) / s; q[2] = (R[5] + R[7]) / s; q[3] = 0.25f * s;
  }
  const float n = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  const float sg = q[0] < 0.0f ? -1.0f : 1.0f;
  for (int i = 0; i < 4; ++i) q[i] = sg * q[i] / n;
}

// Loss and its gradient with respect to the head outputs; one thread per pair.
// targets: q* (w, x, y, z) and t* (metres) per pair.
__global__ void mlp_loss_kernel(const float* __restrict__ o, const float* __restrict__ targets,
                                const float* __restrict__ frame, int n_pairs, float* __restrict__ dout,
                                uint16_t* __restrict__ dout_b, float* __restrict__ loss_pair) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
  const float* op = o + (int64_t)p * kMlpOut;
  const float* tg = targets + 7 * p;
  float q0[4] = {tg[0], tg[1], tg[2], tg[3]};
  const float qn = sqrtf(q0[0] * q0[0] + q0[1] * q0[1] + q0[2] * q0[2] + q0[3] * q0[3]);
  for (int i = 0; i < 4; ++i) q0[i] /= qn;
  float Rs[9];
  quat_to_rot_f(q0, Rs);
  const float* cs = frame + 24 * p;       // source centroid, then V
  const float* ct = frame + 24 * p + 12;  // target centroid, then V
  const float* Vs = cs + 3;
  const float* Vt = ct + 3;
  // label rotation in the canonical frames: Vt^T Rs Vs
  float Rc[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      float acc = 0.0f;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) acc += Vt[i * 3 + r] * Rs[i * 3 + j] * Vs[j * 3 + c];
      Rc[r * 3 + c] = acc;
    }
  float qs[4];
  rot_to_quat_f(Rc, qs);
  float dstar[3];
  for (int r = 0; r < 3; ++r)
    dstar[r] = (tg[4 + r] - (ct[r] - (Rs[3 * r] * cs[0] + Rs[3 * r + 1] * cs[1] + Rs[3 * r + 2] * cs[2]))) / kMlpScale;
  float L[kMlpK], qh[kMlpK][4], sgn[kMlpK], nr[kMlpK];
  int best = 0;
  for (int k = 0; k < kMlpK; ++k) {
    const float* r = op + 7 * k;
    const float n = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]) + 1e-8f;
    nr[k] = n;
    float dot = 0.0f;
    for (int i = 0; i < 4; ++i) {
      qh[k][i] = r[i] / n;
      dot += qh[k][i] * qs[i];
    }
    sgn[k] = dot < 0.0f ? -1.0f : 1.0f;
    float lr = 0.0f, lt = 0.0f;
    for (int i = 0; i < 4; ++i) {
      const float d = qh[k][i] - sgn[k] * qs[i];
      lr += d * d;
    }
    for (int i = 0; i < 3; ++i) {
      const float d = r[4 + i] - dstar[i];
      lt += d * d;
    }
    L[k] = lr + lt;
    if (L[k] < L[best]) best = k;
  }
  // confidences
  const float* cl = op + 7 * kMlpK;
  float mx = cl[0];
  for (int k = 1; k < kMlpK; ++k) mx = fmaxf(mx, cl[k]);
  float se = 0.0f;
  for (int k = 0; k < kMlpK; ++k) se += expf(cl[k] - mx);
  const float lse = mx + logf(se);
  const float beta = 0.1f, eps = 0.05f;
  const float inv_b = 1.0f / (float)n_pairs;
  float loss = beta * (lse - cl[best]);
  float* dp = dout + (int64_t)p * kMlpOut;
  for (int k = 0; k < kMlpK; ++k) {
    const float w = (k == best) ? (1.0f - eps) : eps / (float)(kMlpK - 1);
    loss += w * L[k];
    const float* r = op + 7 * k;
    // d/dq^ = 2 w (q^ - s q*); d/draw = (I - q^ q^T) d/dq^ / |raw|
    float gq[4], dotg = 0.0f;
    for (int i = 0; i < 4; ++i) {
      gq[i] = 2.0f * w * (qh[k][i] - sgn[k] * qs[i]);
      dotg += gq[i] * qh[k][i];
    }
    for (int i = 0; i < 4; ++i) dp[7 * k + i] = inv_b * (gq[i] - dotg * qh[k][i]) / nr[k];
    for (int i = 0; i < 3; ++i) dp[7 * k + 4 + i] = inv_b * 2.0f * w * (r[4 + i] - dstar[i]);
    dp[7 * kMlpK + k] = inv_b * beta * (expf(cl[k] - lse) - (k == best ? 1.0f : 0.0f));
  }
  for (int i = 0; i < kMlpOut; ++i) dout_b[(int64_t)p * kMlpOut + i] = f2bf(dp[i]);
  loss_pair[p] = loss;
}

// Proposals from the head outputs: unit quaternion (w >= 0), translation in
// metres, softmax confidences.
__global__ void mlp_decode_kernel(const float* __restrict__ o, const float* __restrict__ frame, int n_pairs,
                                  float* __restrict__ quat_t, float* __restrict__ conf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
  const float* op = o + (int64_t)p * kMlpOut;
  const float* cs = frame + 24 * p;
  const float* ct = frame + 24 * p + 12;
  const float* Vs = cs + 3;
  const float* Vt = ct + 3;
  for (int k = 0; k < kMlpK; ++k) {
    const float* r = op + 7 * k;
    float qc[4];
    const float n = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]) + 1e-8f;
    for (int i = 0; i < 4; ++i) qc[i] = r[i] / n;
    float Rc[9], R[9];
    quat_to_rot_f(qc, Rc);
    // R = Vt Rc Vs^T
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        float acc = 0.0f;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) acc += Vt[a * 3 + i] * Rc[i * 3 + j] * Vs[b * 3 + j];
        R[a * 3 + b] = acc;
      }
    float q[4];
    rot_to_quat_f(R, q);
    float* out = quat_t + ((int64_t)p * kMlpK + k) * 7;
    for (int i = 0; i < 4; ++i) out[i] = q[i];
    for (int i = 0; i < 3; ++i)
      out[4 + i] = ct[i] - (R[3 * i] * cs[0] + R[3 * i + 1] * cs[1] + R[3 * i + 2] * cs[2]) + kMlpScale * r[4 + i];
  }
  const float* cl = op + 7 * kMlpK;
  float mx = cl[0];
  for (int k = 1; k < kMlpK; ++k) mx = fmaxf(mx, cl[k]);
  float se = 0.0f;
  for (int k = 0; k < kMlpK; ++k) se += expf(cl[k] - mx);
  for (int k = 0; k < kMlpK; ++k) conf[(int64_t)p * kMlpK + k] = expf(cl[k] - mx) / se;
}

// Encoder backward rows, one per (cloud, channel): the argmax point p and the
// pooled gradient v (zero where the pooled value is zero).  Gathers H1[p],
// x[p] and builds dH2 = v W3[:, ch] masked by H2[p] > 0.
__global__ __launch_bounds__(128) void mlp_rows_kernel(const float* __restrict__ dz, const float* __restrict__ gval,
                                                       const int32_t* __restrict__ garg, int npts,
                                                       const uint16_t* __restrict__ X, const uint16_t* __restrict__ H1,
                                                       const uint16_t* __restrict__ H2,
                                                       const uint16_t* __restrict__ W3b, uint16_t* __restrict__ Xg,
                                                       uint16_t* __restrict__ H1g, uint16_t* __restrict__ dH2,
                                                       float* __restrict__ rowv) {
  const int64_t row = blockIdx.x;  // cloud * 256 + ch
  const int cloud = (int)(row / kH3), ch = (int)(row % kH3);
  const int pair = cloud >> 1, half = cloud & 1;
  const float v = gval[row] > 0.0f ? dz[(int64_t)pair * kZ + half * kH3 + ch] : 0.0f;
  const int64_t pt = (int64_t)cloud * npts + garg[row];
  const int t = threadIdx.x;
  if (t == 0) rowv[row] = v;
  // dH2[row][i] = v * W3[i][ch] if H2[pt][i] > 0
  {
    const int i = t;  // 128 threads = 128 hidden units
    const float w = bf2f(W3b[(int64_t)i * kH3 + ch]);
    const bool on = H2[pt * kH2 + i] != 0;
    dH2[row * kH2 + i] = f2bf(on ? v * w : 0.0f);
  }
  if (t < kH1) H1g[row * kH1 + t] = H1[pt * kH1 + t];
  if (t < kMlpInPad) Xg[row * kMlpInPad + t] = X[pt * kMlpInPad + t];
}

// dW3[i][ch] = sum over clouds of v(cloud, ch) H2[argmax][i]; db3[ch] = sum v.
// One thread per (i, ch), clouds summed in index order.
__global__ void mlp_dw3_kernel(const float* __restrict__ rowv, const int32_t* __restrict__ garg, int n_clouds,
                               int npts, const uint16_t* __restrict__ H2, float* __restrict__ dW3,
                               float* __restrict__ db3) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= kH2 * kH3) return;
  const int i = idx / kH3, ch = idx % kH3;
  float s = 0.0f, sb = 0.0f;
  for (int c = 0; c < n_clouds; ++c) {
    const int64_t row = (int64_t)c * kH3 + ch;
    const float v = rowv[row];
    if (v == 0.0f) continue;
    const int64_t pt = (int64_t)c * npts + garg[row];
    s += v * bf2f(H2[pt * kH2 + i]);
    sb += v;
  }
  dW3[(int64_t)i * kH3 + ch] = s;
  if (i == 0) db3[ch] = sb;
}

// Column sums of an M x N matrix (FP32 or BF16), one block per column, fixed
// order.
__global__ __launch_bounds__(256) void mlp_colsum_kernel(const float* __restrict__ Af, const uint16_t* __restrict__ Ab,
                                                         int M, int N, float* __restrict__ out) {
  const int n = blockIdx.x, tid = threadIdx.x;
  __shared__ float s[256];
  float acc = 0.0f;
  for (int m = tid; m < M; m += 256) acc += Af ? Af[(int64_t)m * N + n] : bf2f(Ab[(int64_t)m * N + n]);
  s[tid] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) s[tid] += s[tid + w];
    __syncthreads();
  }
  if (tid == 0) out[n] = s[0];
}

// Adam on the flat parameter vector, then the BF16 weight copies.
__global__ void mlp_adam_kernel(float* __restrict__ w, const float* __restrict__ grad, float* __restrict__ m1,
                                float* __restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1,
                                float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gr = grad[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gr;
  const float b = b2 * m2[i] + (1.0f - b2) * gr * gr;
  m1[i] = a;
  m2[i] = b;
  w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// BF16 copies: W1 padded to 16 input rows, W2..W5 as they are.
__global__ void mlp_cast_kernel(const float* __restrict__ w, uint16_t* __restrict__ W1b, uint16_t* __restrict__ W2b,
                                uint16_t* __restrict__ W3b, uint16_t* __restrict__ W4b, uint16_t* __restrict__ W5b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kMlpInPad * kH1) {
    const int r = i / kH1, c = i % kH1;
    W1b[i] = r < kMlpIn ? f2bf(w[oW1 + r * kH1 + c]) : (uint16_t)0;
  }
  if (i < kH1 * kH2) W2b[i] = f2bf(w[oW2 + i]);
  if (i < kH2 * kH3) W3b[i] = f2bf(w[oW3 + i]);
  if (i < kZ * kZ1) W4b[i] = f2bf(w[oW4 + i]);
  if (i < kZ1 * kMlpOut) W5b[i] = f2bf(w[oW5 + i]);
}

}  // namespace cvo

// ---------------------------------------------------------------------------
// Model object and C ABI
// ---------------------------------------------------------------------------

using namespace cvo;

struct cvo_model {
  cvo_ctx* ctx = nullptr;
  int32_t k = kMlpK;
  int64_t step = 0;
  float* w = nullptr;      // FP32 master parameters
  float* grad = nullptr;
  float* m1 = nullptr;
  float* m2 = nullptr;
  uint16_t* W1b = nullptr;
  uint16_t* W2b = nullptr;
  uint16_t* W3b = nullptr;
  uint16_t* W4b = nullptr;
  uint16_t* W5b = nullptr;
  DevBuf ws[32];           // batch workspace (WS_COUNT used)
};

namespace {

enum {
  WS_PTS = 0, WS_X, WS_CEN, WS_H1, WS_H2, WS_POOL, WS_Z, WS_GVAL, WS_GARG, WS_Z1, WS_O, WS_TGT, WS_DO, WS_DOB,
  WS_DZ1, WS_DZ, WS_ROWV, WS_XG, WS_H1G, WS_DH2, WS_DH1, WS_LOSS, WS_QT, WS_CONF, WS_STAGE, WS_COUNT
};
static_assert(WS_COUNT <= 32, "the model's workspace array holds 32 buffers");

template <class T>
T* wsp(cvo_model* m, int i) {
  return reinterpret_cast<T*>(m->ws[i].p);
}

// Forward over n_pairs pairs of npts points; inputs already in WS_PTS
// ([pair][cloud][point][8] FP32).  Leaves z1 (BF16) and o (FP32).  fused:
// the point encoder in one kernel (inference); otherwise layer by layer with
// the activations h1, h2 kept for the backward pass (training).
cvo_status forward(cvo_model* m, int n_pairs, int npts, bool fused = false) {
  cvo_ctx* ctx = m->ctx;
  hipStream_t s = ctx->stream;
  const int clouds = 2 * n_pairs;
  const int64_t P = (int64_t)clouds * npts;
  CVO_TRY(ensure(ctx, m->ws[WS_X], sizeof(uint16_t) * P * kMlpInPad));
  CVO_TRY(ensure(ctx, m->ws[WS_CEN], sizeof(float) * 12 * clouds));
  if (!fused) {
    CVO_TRY(ensure(ctx, m->ws[WS_H1], sizeof(uint16_t) * P * kH1));
    CVO_TRY(ensure(ctx, m->ws[WS_H2], sizeof(uint16_t) * P * kH2));
  }
  CVO_TRY(ensure(ctx, m->ws[WS_POOL], sizeof(unsigned long long) * clouds * kH3));
  CVO_TRY(ensure(ctx, m->ws[WS_Z], sizeof(uint16_t) * n_pairs * kZ));
  CVO_TRY(ensure(ctx, m->ws[WS_GVAL], sizeof(float) * clouds * kH3));
  CVO_TRY(ensure(ctx, m->ws[WS_GARG], sizeof(int32_t) * clouds * kH3));
  CVO_TRY(ensure(ctx, m->ws[WS_Z1], sizeof(uint16_t) * n_pairs * kZ1));
  CVO_TRY(ensure(ctx, m->ws[WS_O], sizeof(float) * n_pairs * kMlpOut));
  hipLaunchKernelGGL(mlp_prep_kernel, dim3(clouds), dim3(256), 0, s, wsp<float>(m, WS_PTS), npts,
                     wsp<uint16_t>(m, WS_X), wsp<float>(m, WS_CEN));
  const float* w = m->w;
  const int64_t npool = (int64_t)clouds * kH3;
  GemmArgs g;
  if (fused) {
    hipLaunchKernelGGL(encoder_fused_kernel, dim3(clouds), dim3(256), 0, s, wsp<uint16_t>(m, WS_X), npts, m->W1b,
                       m->W2b, m->W3b, w, wsp<unsigned long long>(m, WS_POOL));
  } else {
  // layer 1: 16 (padded) -> 64
  g = gemm_args(wsp<uint16_t>(m, WS_X), kMlpInPad, 1, m->W1b, kH1, 1, (int)P, kH1, kMlpInPad);
  g.bias = w + oB1;
  g.relu = 1;
  g.Cb = wsp<uint16_t>(m, WS_H1);
  g.ldcb = kH1;
  gemm(g, s);
  // layer 2: 64 -> 128
  g = gemm_args(wsp<uint16_t>(m, WS_H1), kH1, 1, m->W2b, kH2, 1, (int)P, kH2, kH1);
  g.bias = w + oB2;
  g.relu = 1;
  g.Cb = wsp<uint16_t>(m, WS_H2);
  g.ldcb = kH2;
  gemm(g, s);
  // layer 3: 128 -> 256 with the max-pool epilogue
  hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((npool + 255) / 256)), dim3(256), 0, s,
                     wsp<unsigned long long>(m, WS_POOL), npool, 0ull);
  g = gemm_args(wsp<uint16_t>(m, WS_H2), kH2, 1, m->W3b, kH3, 1, (int)P, kH3, kH2);
  g.bias = w + oB3;
  g.relu = 1;
  g.pool = wsp<unsigned long long>(m, WS_POOL);
  g.group = npts;
  gemm(g, s);
  }
  hipLaunchKernelGGL(mlp_unpack_pool_kernel, dim3((unsigned)((npool + 255) / 256)), dim3(256), 0, s,
                     wsp<unsigned long long>(m, WS_POOL), n_pairs, wsp<uint16_t>(m, WS_Z), wsp<float>(m, WS_GVAL),
                     wsp<int32_t>(m, WS_GARG));
  // head
  g = gemm_args(wsp<uint16_t>(m, WS_Z), kZ, 1, m->W4b, kZ1, 1, n_pairs, kZ1, kZ);
  g.bias = w + oB4;
  g.relu = 1;
  g.Cb = wsp<uint16_t>(m, WS_Z1);
  g.ldcb = kZ1;
  gemm(g, s);
  g = gemm_args(wsp<uint16_t>(m, WS_Z1), kZ1, 1, m->W5b, kMlpOut, 1, n_pairs, kMlpOut, kZ1);
  g.bias = w + oB5;
  g.C = wsp<float>(m, WS_O);
  g.ldc = kMlpOut;
  gemm(g, s);
  return CVO_OK;
}

cvo_status check_batch(cvo_ctx* ctx, int n_pairs, int npts) {
  if (n_pairs <= 0 || n_pairs > 4096) return set_err(ctx, CVO_ERR_LIMIT, "n_pairs must be in [1, 4096]");
  if (npts < 64 || npts % 64 != 0 || npts > 65536)
    return set_err(ctx, CVO_ERR_LIMIT, "npts must be a multiple of 64 in [64, 65536]");
  if ((int64_t)n_pairs * 2 * npts > (int64_t)1 << 28) return set_err(ctx, CVO_ERR_LIMIT, "batch too large");
  return CVO_OK;
}

// The two batches into [pair][cloud][point][8] (each pair's two clouds
// interleaved).  Host input: one contiguous copy per batch into a staging
// buffer (a strided copy from pageable memory moves row by row), then a
// device-to-device interleave; device input: the interleave alone.
cvo_status upload_points(cvo_model* m, int n_pairs, int npts, const float* a, const float* b, bool on_device) {
  cvo_ctx* ctx = m->ctx;
  const size_t per = sizeof(float) * (size_t)npts * kMlpIn;
  CVO_TRY(ensure(ctx, m->ws[WS_PTS], per * 2 * n_pairs));
  float* d = wsp<float>(m, WS_PTS);
  const float* sa = a;
  const float* sb = b;
  if (!on_device) {
    CVO_TRY(ensure(ctx, m->ws[WS_STAGE], per * 2 * n_pairs));
    float* st = wsp<float>(m, WS_STAGE);
    CVO_HIP(ctx, hipMemcpyAsync(st, a, per * n_pairs, hipMemcpyHostToDevice, ctx->stream));
    CVO_HIP(ctx, hipMemcpyAsync(reinterpret_cast<char*>(st) + per * n_pairs, b, per * n_pairs,
                                hipMemcpyHostToDevice, ctx->stream));
    sa = st;
    sb = reinterpret_cast<const float*>(reinterpret_cast<const char*>(st) + per * n_pairs);
  }
  CVO_HIP(ctx, hipMemcpy2DAsync(d, 2 * per, sa, per, per, n_pairs, hipMemcpyDeviceToDevice, ctx->stream));
  CVO_HIP(ctx, hipMemcpy2DAsync(reinterpret_cast<char*>(d) + per, 2 * per, sb, per, per, n_pairs,
                                hipMemcpyDeviceToDevice, ctx->stream));
  return CVO_OK;
}

void refresh_bf16(cvo_model* m) {
  const int n = kZ * kZ1;  // largest matrix
  hipLaunchKernelGGL(mlp_cast_kernel, dim3((n + 255) / 256), dim3(256), 0, m->ctx->stream, m->w, m->W1b, m->W2b,
                     m->W3b, m->W4b, m->W5b);
}

uint64_t splitmix(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

extern "C" cvo_status cvo_model_create(cvo_ctx* ctx, int32_t k, uint64_t seed, cvo_model** out) {
  if (!ctx || !out) return CVO_ERR_INVALID_ARG;
  *out = nullptr;
  if (k != kMlpK) return set_err(ctx, CVO_ERR_UNSUPPORTED, "the proposer is built for k = 8");
  cvo_model* m = new cvo_model;
  m->ctx = ctx;
  hipError_t e = hipSuccess;
  const size_t pb = sizeof(float) * kMlpParams;
  e = hipMalloc(&m->w, pb);
  if (e == hipSuccess) e = hipMalloc(&m->grad, pb);
  if (e == hipSuccess) e = hipMalloc(&m->m1, pb);
  if (e == hipSuccess) e = hipMalloc(&m->m2, pb);
  if (e == hipSuccess) e = hipMalloc(&m->W1b, sizeof(uint16_t) * kMlpInPad * kH1);
  if (e == hipSuccess) e = hipMalloc(&m->W2b, sizeof(uint16_t) * kH1 * kH2);
  if (e == hipSuccess) e = hipMalloc(&m->W3b, sizeof(uint16_t) * kH2 * kH3);
  if (e == hipSuccess) e = hipMalloc(&m->W4b, sizeof(uint16_t) * kZ * kZ1);
  if (e == hipSuccess) e = hipMalloc(&m->W5b, sizeof(uint16_t) * kZ1 * kMlpOut);
  if (e != hipSuccess) {
    cvo_model_destroy(m);
    return hip_check(ctx, e, "model allocation");
  }
  // He-uniform initialisation from a seeded generator; biases zero
  std::vector<float> w(kMlpParams, 0.0f);
  uint64_t st = seed;
  auto uni = [&]() { return (float)((splitmix(st) >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f; };
  const int fan_in[5] = {kMlpIn, kH1, kH2, kZ, kZ1};
  const int offs[5] = {oW1, oW2, oW3, oW4, oW5};
  const int sizes[5] = {kMlpIn * kH1, kH1 * kH2, kH2 * kH3, kZ * kZ1, kZ1 * kMlpOut};
  for (int l = 0; l < 5; ++l) {
    const float lim = std::sqrt(6.0f / (float)fan_in[l]) * (l == 4 ? 0.1f : 1.0f);
    for (int i = 0; i < sizes[l]; ++i) w[offs[l] + i] = lim * uni();
  }
  cvo_status rs = cvo_model_set_weights(m, w.data(), kMlpParams);
  if (rs != CVO_OK) {
    cvo_model_destroy(m);
    return rs;
  }
  *out = m;
  return CVO_OK;
}

extern "C" void cvo_model_destroy(cvo_model* m) {
  if (!m) return;
  if (m->ctx && m->ctx->stream) (void)hipStreamSynchronize(m->ctx->stream);
  void* ps[] = {m->w, m->grad, m->m1, m->m2, m->W1b, m->W2b, m->W3b, m->W4b, m->W5b};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  for (auto& b : m->ws)
    if (b.p) (void)hipFree(b.p);
  delete m;
}

extern "C" int64_t cvo_model_num_params(void) { return kMlpParams; }

extern "C" cvo_status cvo_model_get_weights(const cvo_model* m, float* out, int64_t cap, int64_t* n_out) {
  if (!m || !out) return CVO_ERR_INVALID_ARG;
  if (n_out) *n_out = kMlpParams;
  if (cap < kMlpParams) return set_err(m->ctx, CVO_ERR_LIMIT, "weight buffer too small");
  CVO_HIP(m->ctx, hipMemcpyAsync(out, m->w, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m->ctx->stream));
  CVO_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return CVO_OK;
}

extern "C" cvo_status cvo_model_set_weights(cvo_model* m, const float* w, int64_t n) {
  if (!m || !w) return CVO_ERR_INVALID_ARG;
  if (n != kMlpParams) return set_err(m->ctx, CVO_ERR_INVALID_ARG, "weight count mismatch");
  hipStream_t s = m->ctx->stream;
  CVO_HIP(m->ctx, hipMemcpyAsync(m->w, w, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s));
  CVO_HIP(m->ctx, hipMemsetAsync(m->m1, 0, sizeof(float) * kMlpParams, s));
  CVO_HIP(m->ctx, hipMemsetAsync(m->m2, 0, sizeof(float) * kMlpParams, s));
  m->step = 0;
  refresh_bf16(m);
  CVO_HIP(m->ctx, hipStreamSynchronize(s));
  return CVO_OK;
}

// Gradient of the mean loss over the batch (FP32, parameter order), without an
// update; loss_out optional.
extern "C" cvo_status cvo_model_gradient(cvo_ctx* ctx, cvo_model* m, int32_t n_pairs, int32_t npts, const float* a,
                                         const float* b, const float* targets, float* grad_out, float* loss_out) {
  if (!ctx || !m || !a || !b || !targets) return CVO_ERR_INVALID_ARG;
  CVO_TRY(check_batch(ctx, n_pairs, npts));
  hipStream_t s = ctx->stream;
  CVO_TRY(upload_points(m, n_pairs, npts, a, b, false));
  CVO_TRY(forward(m, n_pairs, npts));
  const int clouds = 2 * n_pairs;
  const int R = clouds * kH3;
  CVO_TRY(ensure(ctx, m->ws[WS_TGT], sizeof(float) * 7 * n_pairs));
  CVO_TRY(ensure(ctx, m->ws[WS_DO], sizeof(float) * n_pairs * kMlpOut));
  CVO_TRY(ensure(ctx, m->ws[WS_DOB], sizeof(uint16_t) * n_pairs * kMlpOut));
  CVO_TRY(ensure(ctx, m->ws[WS_DZ1], sizeof(uint16_t) * n_pairs * kZ1));
  CVO_TRY(ensure(ctx, m->ws[WS_DZ], sizeof(float) * n_pairs * kZ));
  CVO_TRY(ensure(ctx, m->ws[WS_ROWV], sizeof(float) * R));
  CVO_TRY(ensure(ctx, m->ws[WS_XG], sizeof(uint16_t) * (size_t)R * kMlpInPad));
  CVO_TRY(ensure(ctx, m->ws[WS_H1G], sizeof(uint16_t) * (size_t)R * kH1));
  CVO_TRY(ensure(ctx, m->ws[WS_DH2], sizeof(uint16_t) * (size_t)R * kH2));
  CVO_TRY(ensure(ctx, m->ws[WS_DH1], sizeof(uint16_t) * (size_t)R * kH1));
  CVO_TRY(ensure(ctx, m->ws[WS_LOSS], sizeof(float) * n_pairs));
  CVO_HIP(ctx, hipMemcpyAsync(m->ws[WS_TGT].p, targets, sizeof(float) * 7 * n_pairs, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(mlp_loss_kernel, dim3((n_pairs + 63) / 64), dim3(64), 0, s, wsp<float>(m, WS_O),
                     wsp<float>(m, WS_TGT), wsp<float>(m, WS_CEN), n_pairs, wsp<float>(m, WS_DO),
                     wsp<uint16_t>(m, WS_DOB), wsp<float>(m, WS_LOSS));
  float* gr = m->grad;
  // head layer 2: dW5 = z1^T do, db5 = colsum(do), dz1 = (do W5^T) * [z1 > 0]
  GemmArgs g = gemm_args(wsp<uint16_t>(m, WS_Z1), 1, kZ1, wsp<uint16_t>(m, WS_DOB), kMlpOut, 1, kZ1, kMlpOut,
                         n_pairs);
  g.C = gr + oW5;
  g.ldc = kMlpOut;
  gemm(g, s);
  hipLaunchKernelGGL(mlp_colsum_kernel, dim3(kMlpOut), dim3(256), 0, s, wsp<float>(m, WS_DO), nullptr, n_pairs,
                     kMlpOut, gr + oB5);
  g = gemm_args(wsp<uint16_t>(m, WS_DOB), kMlpOut, 1, m->W5b, 1, kMlpOut, n_pairs, kZ1, kMlpOut);
  g.mask = wsp<uint16_t>(m, WS_Z1);
  g.ldm = kZ1;
  g.Cb = wsp<uint16_t>(m, WS_DZ1);
  g.ldcb = kZ1;
  gemm(g, s);
  // head layer 1: dW4 = z^T dz1, db4 = colsum(dz1), dz = dz1 W4^T
  g = gemm_args(wsp<uint16_t>(m, WS_Z), 1, kZ, wsp<uint16_t>(m, WS_DZ1), kZ1, 1, kZ, kZ1, n_pairs);
  g.C = gr + oW4;
  g.ldc = kZ1;
  gemm(g, s);
  hipLaunchKernelGGL(mlp_colsum_kernel, dim3(kZ1), dim3(256), 0, s, nullptr, wsp<uint16_t>(m, WS_DZ1), n_pairs,
                     kZ1, gr + oB4);
  g = gemm_args(wsp<uint16_t>(m, WS_DZ1), kZ1, 1, m->W4b, 1, kZ1, n_pairs, kZ, kZ1);
  g.C = wsp<float>(m, WS_DZ);
  g.ldc = kZ;
  gemm(g, s);
  // encoder, one row per (cloud, channel)
  hipLaunchKernelGGL(mlp_rows_kernel, dim3(R), dim3(128), 0, s, wsp<float>(m, WS_DZ), wsp<float>(m, WS_GVAL),
                     wsp<int32_t>(m, WS_GARG), npts, wsp<uint16_t>(m, WS_X), wsp<uint16_t>(m, WS_H1),
                     wsp<uint16_t>(m, WS_H2), m->W3b, wsp<uint16_t>(m, WS_XG), wsp<uint16_t>(m, WS_H1G),
                     wsp<uint16_t>(m, WS_DH2), wsp<float>(m, WS_ROWV));
  hipLaunchKernelGGL(mlp_dw3_kernel, dim3((kH2 * kH3 + 255) / 256), dim3(256), 0, s, wsp<float>(m, WS_ROWV),
                     wsp<int32_t>(m, WS_GARG), clouds, npts, wsp<uint16_t>(m, WS_H2), gr + oW3, gr + oB3);
  // dW2 = H1g^T dH2, db2 = colsum(dH2), dH1 = (dH2 W2^T) * [H1g > 0]
  g = gemm_args(wsp<uint16_t>(m, WS_H1G), 1, kH1, wsp<uint16_t>(m, WS_DH2), kH2, 1, kH1, kH2, R);
  g.C = gr + oW2;
  g.ldc = kH2;
  gemm(g, s);
  hipLaunchKernelGGL(mlp_colsum_kernel, dim3(kH2), dim3(256), 0, s, nullptr, wsp<uint16_t>(m, WS_DH2), R, kH2,
                     gr + oB2);
  g = gemm_args(wsp<uint16_t>(m, WS_DH2), kH2, 1, m->W2b, 1, kH2, R, kH1, kH2);
  g.mask = wsp<uint16_t>(m, WS_H1G);
  g.ldm = kH1;
  g.Cb = wsp<uint16_t>(m, WS_DH1);
  g.ldcb = kH1;
  gemm(g, s);
  // dW1 = Xg^T dH1 (first 8 input rows), db1 = colsum(dH1)
  g = gemm_args(wsp<uint16_t>(m, WS_XG), 1, kMlpInPad, wsp<uint16_t>(m, WS_DH1), kH1, 1, kMlpIn, kH1, R);
  g.C = gr + oW1;
  g.ldc = kH1;
  gemm(g, s);
  hipLaunchKernelGGL(mlp_colsum_kernel, dim3(kH1), dim3(256), 0, s, nullptr, wsp<uint16_t>(m, WS_DH1), R, kH1,
                     gr + oB1);
  CVO_HIP(ctx, hipGetLastError());
  if (grad_out) {
    CVO_HIP(ctx, hipMemcpyAsync(grad_out, gr, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, s));
  }
  if (loss_out) {
    std::vector<float> lp(n_pairs);
    CVO_HIP(ctx, hipMemcpyAsync(lp.data(), m->ws[WS_LOSS].p, sizeof(float) * n_pairs, hipMemcpyDeviceToHost, s));
    CVO_HIP(ctx, hipStreamSynchronize(s));
    double sum = 0.0;
    for (float v : lp) sum += v;
    *loss_out = (float)(sum / n_pairs);
  }
  CVO_HIP(ctx, hipStreamSynchronize(s));
  return CVO_OK;
}

extern "C" cvo_status cvo_model_train_step(cvo_ctx* ctx, cvo_model* m, int32_t n_pairs, int32_t npts,
                                           const float* a, const float* b, const float* targets, float lr,
                                           float* loss_out) {
  CVO_TRY(cvo_model_gradient(ctx, m, n_pairs, npts, a, b, targets, nullptr, loss_out));
  m->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, ctx->stream, m->w, m->grad,
                     m->m1, m->m2, kMlpParams, lr, b1, b2, 1e-8f, c1, c2);
  refresh_bf16(m);
  CVO_HIP(ctx, hipGetLastError());
  CVO_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CVO_OK;
}

static cvo_status propose_impl(cvo_ctx* ctx, const cvo_model* mc, int32_t n_pairs, int32_t npts, const float* a,
                               const float* b, bool on_device, float* quat_t, float* conf, float* raw_out) {
  if (!ctx || !mc || !a || !b || !quat_t || !conf) return CVO_ERR_INVALID_ARG;
  CVO_TRY(check_batch(ctx, n_pairs, npts));
  cvo_model* m = const_cast<cvo_model*>(mc);
  hipStream_t s = ctx->stream;
  CVO_TRY(upload_points(m, n_pairs, npts, a, b, on_device));
  if (ctx->profiling) {
    CVO_TRY(ensure_events(ctx, 2));
    CVO_HIP(ctx, hipEventRecord(ctx->ev_pool[0], s));
  }
  CVO_TRY(forward(m, n_pairs, npts, true));
  if (ctx->profiling) {
    CVO_HIP(ctx, hipEventRecord(ctx->ev_pool[1], s));
    CVO_HIP(ctx, hipEventSynchronize(ctx->ev_pool[1]));
    float ms = 0.0f;
    CVO_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_pool[0], ctx->ev_pool[1]));
    // algorithmic flops of the forward: 2 per multiply-add (input padded to 16)
    const double pts = 2.0 * n_pairs * npts;
    const double macs = pts * (kMlpInPad * kH1 + kH1 * kH2 + kH2 * kH3) + (double)n_pairs * (kZ * kZ1 + kZ1 * kMlpOut);
    ctx->stats.ms[CVO_KSTAT_MLP] += ms;
    ctx->stats.launches[CVO_KSTAT_MLP] += 1;
    ctx->stats.bytes[CVO_KSTAT_MLP] += 2.0 * macs;
  }
  CVO_TRY(ensure(ctx, m->ws[WS_QT], sizeof(float) * n_pairs * kMlpK * 7));
  CVO_TRY(ensure(ctx, m->ws[WS_CONF], sizeof(float) * n_pairs * kMlpK));
  hipLaunchKernelGGL(mlp_decode_kernel, dim3((n_pairs + 63) / 64), dim3(64), 0, s, wsp<float>(m, WS_O),
                     wsp<float>(m, WS_CEN), n_pairs, wsp<float>(m, WS_QT), wsp<float>(m, WS_CONF));
  CVO_HIP(ctx, hipGetLastError());
  CVO_HIP(ctx, hipMemcpyAsync(quat_t, m->ws[WS_QT].p, sizeof(float) * n_pairs * kMlpK * 7, hipMemcpyDeviceToHost, s));
  CVO_HIP(ctx, hipMemcpyAsync(conf, m->ws[WS_CONF].p, sizeof(float) * n_pairs * kMlpK, hipMemcpyDeviceToHost, s));
  if (raw_out)
    CVO_HIP(ctx, hipMemcpyAsync(raw_out, m->ws[WS_O].p, sizeof(float) * n_pairs * kMlpOut, hipMemcpyDeviceToHost, s));
  CVO_HIP(ctx, hipStreamSynchronize(s));
  return CVO_OK;
}

extern "C" cvo_status cvo_propose_poses(cvo_ctx* ctx, const cvo_model* m, int32_t n_pairs, int32_t npts,
                                        const float* a, const float* b, float* quat_t, float* conf,
                                        float* raw_out) {
  return propose_impl(ctx, m, n_pairs, npts, a, b, false, quat_t, conf, raw_out);
}

extern "C" cvo_status cvo_propose_poses_d(cvo_ctx* ctx, const cvo_model* m, int32_t n_pairs, int32_t npts,
                                          const float* a_dev, const float* b_dev, float* quat_t, float* conf,
                                          float* raw_out) {
  return propose_impl(ctx, m, n_pairs, npts, a_dev, b_dev, true, quat_t, conf, raw_out);
}

// BF16 MFMA GEMM on caller data (host arrays): C = A B with A (M x K) and
// B (K x N) row-major, rounded to BF16 on the way in.  Exposed so the tile
// layout can be checked against a host product.
extern "C" cvo_status cvo_gemm_bf16(cvo_ctx* ctx, int32_t M, int32_t N, int32_t K, const float* A, const float* B,
                                    float* C) {
  if (!ctx || !A || !B || !C || M <= 0 || N <= 0 || K <= 0) return CVO_ERR_INVALID_ARG;
  if ((int64_t)M * K > (1 << 26) || (int64_t)K * N > (1 << 26) || (int64_t)M * N > (1 << 26))
    return set_err(ctx, CVO_ERR_LIMIT, "gemm operands too large");
  hipStream_t s = ctx->stream;
  std::vector<uint16_t> a((size_t)M * K), b((size_t)K * N);
  for (size_t i = 0; i < a.size(); ++i) a[i] = f2bf(A[i]);
  for (size_t i = 0; i < b.size(); ++i) b[i] = f2bf(B[i]);
  CVO_TRY(ensure(ctx, ctx->b_mlp[0], a.size() * 2));
  CVO_TRY(ensure(ctx, ctx->b_mlp[1], b.size() * 2));
  CVO_TRY(ensure(ctx, ctx->b_mlp[2], sizeof(float) * (size_t)M * N));
  CVO_HIP(ctx, hipMemcpyAsync(ctx->b_mlp[0].p, a.data(), a.size() * 2, hipMemcpyHostToDevice, s));
  CVO_HIP(ctx, hipMemcpyAsync(ctx->b_mlp[1].p, b.data(), b.size() * 2, hipMemcpyHostToDevice, s));
  GemmArgs g = gemm_args(reinterpret_cast<uint16_t*>(ctx->b_mlp[0].p), K, 1, reinterpret_cast<uint16_t*>(ctx->b_mlp[1].p),
                         N, 1, M, N, K);
  g.C = reinterpret_cast<float*>(ctx->b_mlp[2].p);
  g.ldc = N;
  gemm(g, s);
  CVO_HIP(ctx, hipGetLastError());
  CVO_HIP(ctx, hipMemcpyAsync(C, ctx->b_mlp[2].p, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToHost, s));
  CVO_HIP(ctx, hipStreamSynchronize(s));
  return CVO_OK;
}
