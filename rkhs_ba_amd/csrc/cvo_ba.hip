// Multi-frame IRLS bundle adjustment on the device: CvoGPU::align over frames
// and edges (src/cvo/CvoGPU.cu:1614-1671) with CvoBatchIRLS::solve
// (src/cvo/IRLS.cpp:148-367) and BinaryStateGPU (src/cvo/IRLS_State_GPU.cu:
// 144-336, src/cvo/IRLS_State_GPU.cpp:10-146).
//
// One IRLS outer iteration is one replay of a captured graph:
//   ba_prepare        per-edge maps (world transforms of both frames), cap K,
//                     grid level
//   assoc / overflow  per-edge association between the frames' world points,
//                     with per-edge stored pairs and largest row
//   ba_edges          edge on/off; stop when no edge has enough pairs
//   (ba_pair -> ba_edge_reduce -> ba_solve) x (iterations_per_solve + 1)
//                     Levenberg-Marquardt over the fixed pairs
//   ba_finish         per-edge lengthscale, pose change, convergence
//
// Residuals and Jacobians are the reference's PairwiseAnalyticalDiffFunctor
// (src/cvo/IRLS_Cost_CPU.hpp:305-357): r = a |T1 p1 - T2 p2|^2 with
// J1 = diff^T dT1, J2 = -diff^T dT2 (no factor 2a), in the local
// parameterisation T <- T Exp([rho; phi]) (local_parameterization_se3.hpp:
// 21-99).  With w = P1 - P2 (world), u = [w; P1 x w] the local Jacobians are
// j1 = M1 u and j2 = -M2 u, M = [[R^T, 0], [-R^T [t]x, R^T]], so each edge only
// accumulates S = sum u u^T, h = sum r u and c = sum r^2 over its pairs.
//
// The Ceres solve is replaced by this Levenberg-Marquardt schedule:
//   D = clamp(diag(H), 1e-6, 1e32); solve (H + D/mu) d = -g;
//   rho = (F - F_new) / (-(g.d + d.H.d/2)); accept if rho > 1e-3:
//     mu /= max(1/3, 1 - (2 rho - 1)^3) (capped at 1e16), v = 2;
//   else mu /= v, v *= 2.  mu = 1e4 at the start of every solve; stop on
//   iterations_per_solve candidates, |dF|/F < 1e-5 after an accepted step,
//   ||g||inf < 1e-5 or ||d|| < 1e-5 (||x|| + 1e-5).

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cvo_device.h"
#include "cvo_internal.h"

namespace cvo {

constexpr int kBaMaxFrames = 64;
constexpr int kBaMaxFree = 20;                 // free frames: 120 dof, Cholesky in 115 KB of LDS
constexpr int kBaMaxDof = 6 * kBaMaxFree;
constexpr int kBaMaxEdges = 4096;
constexpr int kBaTerms = 28;                   // S (21, upper triangle), h (6), c (1)
constexpr int kBaContrib = 121;                // B11, B12, B22 (36 each), g1, g2 (6 each), c
constexpr int kBaRowsPerBlock = 256;           // pair pass: one row per thread
constexpr int kPairLoads = 4;                  // pair pass: loads in flight per lane
constexpr int kSolveThreads = 256;             // four waves: the Cholesky's trailing update spread over rows
static_assert(kSolveThreads == 256 && 6 * kBaMaxFree <= kSolveThreads, "the solve's reductions assume four waves");
constexpr int kChol = 8;                       // Cholesky block: columns a panel
constexpr int kTri = 8;                        // triangular solves: steps a load batch

struct BaEdge {
  double ell;
  int64_t nnz, nnz_last;
  int32_t f1, f2;
  int32_t K, last_max, iters_per_ell, on;
  int32_t n_rows, n_pair_blocks;
  int32_t pad0, pad1;
};

struct BaDev {
  int32_t n_frames, n_edges, nfree, n;
  int32_t iter, done, max_iters, iters_per_solve;
  int32_t lm_iter, lm_stop, eval_pending, eval_is_candidate;
  int32_t K_init, min_nonzeros, iter_per_ell, counter_nonzero;
  int32_t stopped_no_edges, accepted_total, evals_total, h_cur;
  float ell_min_f, decay_f, rq_coef, pad_f;
  EllPolicy pol;
  double ell_min, ell_max;
  double F, Fn, mu, v, change;
  int32_t free_idx[kBaMaxFrames];
  int32_t frame_of_free[kBaMaxFree];
  float frame_norm[kBaMaxFrames];
  double X[kBaMaxFrames * 12];     // current poses, row-major 3x4 (world <- sensor)
  double Xe[kBaMaxFrames * 12];    // poses of the pending pair pass
  double Xold[kBaMaxFrames * 12];  // poses at the start of the outer iteration
  double g[kBaMaxDof], gn[kBaMaxDof], d[kBaMaxDof];
};

// ---------------------------------------------------------------------------
// Helpers
// ---------------------------------------------------------------------------

// M = [[R^T, 0], [-R^T [t]x, R^T]] of a 3x4 pose, row-major 6x6.
__device__ inline void pose_M(const double* T, double M[36]) {
  double Rt[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Rt[r * 3 + c] = T[c * 4 + r];
  const double t0 = T[3], t1 = T[7], t2 = T[11];
  const double tx[9] = {0.0, -t2, t1, t2, 0.0, -t0, -t1, t0, 0.0};
  for (int i = 0; i < 36; ++i) M[i] = 0.0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      M[r * 6 + c] = Rt[r * 3 + c];
      M[(r + 3) * 6 + c + 3] = Rt[r * 3 + c];
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += Rt[r * 3 + k] * tx[k * 3 + c];
      M[(r + 3) * 6 + c] = -s;
    }
}

// T <- T Exp([rho; phi]), translation part first.  out may alias Tin.
__device__ inline void pose_plus(const double* Tin, const double* dlt, double* out) {
  double T[12];
  for (int i = 0; i < 12; ++i) T[i] = Tin[i];
  const double xi[6] = {dlt[3], dlt[4], dlt[5], dlt[0], dlt[1], dlt[2]};
  double dR[9], dT[3];
  se3_exp(xi, 1.0, dR, dT);
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c)
      out[r * 4 + c] = T[r * 4 + 0] * dR[0 * 3 + c] + T[r * 4 + 1] * dR[1 * 3 + c] + T[r * 4 + 2] * dR[2 * 3 + c];
    out[r * 4 + 3] = T[r * 4 + 0] * dT[0] + T[r * 4 + 1] * dT[1] + T[r * 4 + 2] * dT[2] + T[r * 4 + 3];
  }
}

// Float maps of an edge: source and target world transforms, and the query
// map from the world into the target's sensor frame.
__device__ inline void edge_maps(const double* T1, const double* T2, ProblemState* st) {
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) {
      st->As[r * 3 + c] = (float)T1[r * 4 + c];
      st->A[r * 3 + c] = (float)T2[r * 4 + c];
      st->Rq[r * 3 + c] = (float)T2[c * 4 + r];
    }
    st->bs[r] = (float)T1[r * 4 + 3];
    st->b[r] = (float)T2[r * 4 + 3];
    st->Tq[r] = (float)(-(T2[0 * 4 + r] * T2[3] + T2[1 * 4 + r] * T2[7] + T2[2 * 4 + r] * T2[11]));
  }
  st->has_src_xf = 1;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void ba_prepare_kernel(BaDev* __restrict__ ba, BaEdge* __restrict__ edges,
                                                         const AssocProblem* __restrict__ probs) {
  const int tid = threadIdx.x;
  const bool done = ba->done != 0;
  const int E = ba->n_edges;
  if (done) {
    for (int e = tid; e < E; e += blockDim.x) probs[e].st->done = 1;
    return;
  }
  for (int i = tid; i < ba->n_frames * 12; i += blockDim.x) {
    ba->Xold[i] = ba->X[i];
    ba->Xe[i] = ba->X[i];
  }
  for (int e = tid; e < E; e += blockDim.x) {
    BaEdge& ed = edges[e];
    ProblemState* st = probs[e].st;
    ed.nnz_last = ed.nnz;
    if (ed.last_max > 0) ed.K = min(ba->K_init, (int)((double)ed.last_max * 1.1));
    if (ed.K < 1) ed.K = 1;
    const double* T1 = ba->X + 12 * ed.f1;
    const double* T2 = ba->X + 12 * ed.f2;
    edge_maps(T1, T2, st);
    st->ell = (float)ed.ell;
    st->K = ed.K;
    st->iter = ba->iter;
    st->done = 0;
    st->n_overflow = 0;
    // query radius bound at this lengthscale over the world-frame source
    const double tn = sqrt(T1[3] * T1[3] + T1[7] * T1[7] + T1[11] * T1[11]);
    const float sn = ba->frame_norm[ed.f1] + (float)tn;
    const float tq = fmaxf(fabsf(st->Tq[0]), fmaxf(fabsf(st->Tq[1]), fabsf(st->Tq[2])));
    const float r_need =
        ba->rq_coef * st->ell * (1.0f + sn / 500.0f) * 1.0001f + 1e-4f * (1.0f + sn + tq) + 1e-6f;
    const bool big = policy_big(ba->pol, st->ell);
    st->group = big ? ba->pol.group_big : ba->pol.group_small;
    st->level = choose_level(probs[e].grid, r_need * (big ? ba->pol.scale_big : ba->pol.scale_small));
  }
}

// Edge switches after the association (IRLS.cpp:235-250,
// IRLS_State_GPU.cpp:17-21) and the start of the Levenberg-Marquardt solve.
__global__ __launch_bounds__(256) void ba_edges_kernel(BaDev* __restrict__ ba, BaEdge* __restrict__ edges,
                                                       const AssocProblem* __restrict__ probs) {
  if (ba->done) return;
  const int tid = threadIdx.x;
  __shared__ int s_cnt[256];
  int cnt = 0;
  for (int e = tid; e < ba->n_edges; e += blockDim.x) {
    BaEdge& ed = edges[e];
    const ProblemState* st = probs[e].st;
    ed.nnz = (int64_t)st->nnz;
    ed.last_max = (int32_t)st->max_row;
    ed.on = (ed.nnz > 0 && ed.nnz >= (int64_t)ba->min_nonzeros) ? 1 : 0;
    if (ed.nnz > (int64_t)ba->min_nonzeros) cnt++;
  }
  s_cnt[tid] = cnt;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) s_cnt[tid] += s_cnt[tid + w];
    __syncthreads();
  }
  if (tid == 0) {
    ba->counter_nonzero = s_cnt[0];
    if (s_cnt[0] == 0) {
      ba->done = 1;
      ba->stopped_no_edges = 1;
      ba->eval_pending = 0;
      return;
    }
    ba->mu = 1e4;
    ba->v = 2.0;
    ba->lm_iter = 0;
    ba->lm_stop = 0;
    ba->eval_is_candidate = 0;
    ba->eval_pending = ba->n > 0 ? 1 : 0;
  }
}

// FP64 pair pass at the poses Xe, block partials of S (21), h (6), c (1).
// Each wave takes 64 source rows of one edge (lane i owns row i and its world
// point), prefix-sums their pair counts and spreads the flattened pairs over
// its lanes, so every lane's loads are independent and the lanes stay busy
// whatever the row lengths.
__global__ __launch_bounds__(kBaRowsPerBlock) void ba_pair_kernel(const BaDev* __restrict__ ba,
                                                                  const BaEdge* __restrict__ edges,
                                                                  const AssocProblem* __restrict__ probs,
                                                                  double* __restrict__ part, int max_blocks) {
  if (ba->done || !ba->eval_pending) return;
  const int e = blockIdx.y;
  const BaEdge& ed = edges[e];
  if (!ed.on || (int)blockIdx.x >= ed.n_pair_blocks) return;
  const AssocProblem& P = probs[e];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr0 = blockIdx.x * kBaRowsPerBlock + wave * 64;
  // pairs are stored at the row stride of this iteration's cap (the
  // reference re-strides its pair matrix the same way)
  const size_t row_stride = (size_t)P.st->K;
  double T1[12], T2[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    T1[i] = ba->Xe[12 * ed.f1 + i];
    T2[i] = ba->Xe[12 * ed.f2 + i];
  }
  // the lane's own row: world point and pair count
  const int r_own = wr0 + lane;
  uint32_t cnt = 0;
  double X0 = 0.0, X1 = 0.0, X2 = 0.0;
  if (r_own < P.n_src) {
    cnt = ld_global(P.counts + r_own);
    const float4 x4 = ld_global(P.src_pos + r_own);
    const double p0 = x4.x, p1 = x4.y, p2 = x4.z;
    X0 = fma(T1[2], p2, fma(T1[1], p1, fma(T1[0], p0, T1[3])));
    X1 = fma(T1[6], p2, fma(T1[5], p1, fma(T1[4], p0, T1[7])));
    X2 = fma(T1[10], p2, fma(T1[9], p1, fma(T1[8], p0, T1[11])));
  }
  uint32_t incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  const uint32_t total = __shfl(incl, 63, 64);
  const uint32_t excl = incl - cnt;
  double acc[kBaTerms];
#pragma unroll
  for (int k = 0; k < kBaTerms; ++k) acc[k] = 0.0;
  // kPairLoads pairs a lane a trip: the loads go out unconditionally (past
  // the end a lane reloads the last pair, which it then skips) so that all
  // are in flight before the first is used, then the pairs are summed in
  // flattened order, lane by lane as one pair a trip would
  for (uint32_t base = 0; base < total; base += 64 * kPairLoads) {
    int kr[kPairLoads];
    float4 yv[kPairLoads];
#pragma unroll
    for (int u = 0; u < kPairLoads; ++u) {
      const uint32_t idx = min(base + (uint32_t)(u * 64 + lane), total - 1);
      // row of flattened pair idx: the last lane whose exclusive offset <= idx
      int k = 0;
#pragma unroll
      for (int s = 32; s > 0; s >>= 1) {
        const uint32_t ex = __shfl(excl, k + s, 64);
        if (ex <= idx) k += s;
      }
      const uint32_t slot = idx - __shfl(excl, k, 64);
      kr[u] = k;
      yv[u] = ld_global(P.pairs4 + (size_t)(wr0 + k) * row_stride + slot);
    }
#pragma unroll
    for (int t = 0; t < kPairLoads; ++t) {
      const int k = kr[t];
      const double Y0 = __shfl(X0, k, 64), Y1 = __shfl(X1, k, 64), Y2 = __shfl(X2, k, 64);
      if (base + (uint32_t)(t * 64 + lane) < total) {
        // FP64 with explicit fused multiply-adds (the library is built with
        // contraction off for the FP32 gates)
        const float4 y4 = yv[t];
        const double a = (double)y4.w;
        const double q0 = y4.x, q1 = y4.y, q2 = y4.z;
        const double w0 = Y0 - fma(T2[2], q2, fma(T2[1], q1, fma(T2[0], q0, T2[3])));
        const double w1 = Y1 - fma(T2[6], q2, fma(T2[5], q1, fma(T2[4], q0, T2[7])));
        const double w2 = Y2 - fma(T2[10], q2, fma(T2[9], q1, fma(T2[8], q0, T2[11])));
        const double res = a * fma(w2, w2, fma(w1, w1, w0 * w0));
        const double u[6] = {w0, w1, w2, fma(Y1, w2, -Y2 * w1), fma(Y2, w0, -Y0 * w2), fma(Y0, w1, -Y1 * w0)};
        int q = 0;
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
          for (int j = i; j < 6; ++j) {
            acc[q] = fma(u[i], u[j], acc[q]);
            ++q;
          }
#pragma unroll
        for (int i = 0; i < 6; ++i) acc[21 + i] = fma(res, u[i], acc[21 + i]);
        acc[27] = fma(res, res, acc[27]);
      }
    }
  }
  // fixed-order block partial: wave shuffles, then the four waves in LDS
  __shared__ double s_w[kBaTerms][kBaRowsPerBlock / 64];
#pragma unroll
  for (int k = 0; k < kBaTerms; ++k) {
    double v = acc[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) s_w[k][wave] = v;
  }
  __syncthreads();
  if (tid < kBaTerms) {
    double v = 0.0;
    for (int w = 0; w < kBaRowsPerBlock / 64; ++w) v += s_w[tid][w];
    part[((size_t)e * max_blocks + blockIdx.x) * kBaTerms + tid] = v;
  }
}

// Per-edge totals of the pair pass and the edge's blocks of the normal
// equations at the poses Xe: B11 = M1 S M1^T, B12 = -M1 S M2^T,
// B22 = M2 S M2^T, g1 = M1 h, g2 = -M2 h.
__global__ __launch_bounds__(256) void ba_edge_reduce_kernel(const BaDev* __restrict__ ba,
                                                             const BaEdge* __restrict__ edges,
                                                             const double* __restrict__ part, int max_blocks,
                                                             double* __restrict__ contrib) {
  if (ba->done || !ba->eval_pending) return;
  const int e = blockIdx.x;
  const BaEdge& ed = edges[e];
  if (!ed.on) return;
  const int tid = threadIdx.x;
  __shared__ double s_p[kBaTerms][8];
  __shared__ double s_t[kBaTerms];
  __shared__ double s_S[36], s_M1[36], s_M2[36], s_T1[36], s_T2[36];
  if (tid < kBaTerms * 8) {
    const int k = tid >> 3, l = tid & 7;
    double v = 0.0;
    for (int b = l; b < ed.n_pair_blocks; b += 8) v += part[((size_t)e * max_blocks + b) * kBaTerms + k];
    s_p[k][l] = v;
  }
  if (tid == 0) pose_M(ba->Xe + 12 * ed.f1, s_M1);
  if (tid == 1) pose_M(ba->Xe + 12 * ed.f2, s_M2);
  __syncthreads();
  if (tid < kBaTerms) {
    double v = 0.0;
    for (int l = 0; l < 8; ++l) v += s_p[tid][l];
    s_t[tid] = v;
  }
  __syncthreads();
  if (tid < 36) {
    // full symmetric S from the packed upper triangle
    const int i = tid / 6, j = tid % 6;
    const int a = min(i, j), b = max(i, j);
    const int k = a * 6 - a * (a - 1) / 2 + (b - a);
    s_S[tid] = s_t[k];
  }
  __syncthreads();
  if (tid < 72) {
    const int which = tid / 36, idx = tid % 36, i = idx / 6, j = idx % 6;
    const double* M = which ? s_M2 : s_M1;
    double v = 0.0;
    for (int k = 0; k < 6; ++k) v += M[i * 6 + k] * s_S[k * 6 + j];
    (which ? s_T2 : s_T1)[idx] = v;
  }
  __syncthreads();
  double* out = contrib + (size_t)e * kBaContrib;
  if (tid < 108) {
    const int blk = tid / 36, idx = tid % 36, i = idx / 6, j = idx % 6;
    const double* Tm = (blk == 2) ? s_T2 : s_T1;
    const double* M = (blk == 0) ? s_M1 : s_M2;
    double v = 0.0;
    for (int k = 0; k < 6; ++k) v += Tm[i * 6 + k] * M[j * 6 + k];
    out[tid] = (blk == 1) ? -v : v;
  } else if (tid < 120) {
    const int which = (tid - 108) / 6, i = (tid - 108) % 6;
    const double* M = which ? s_M2 : s_M1;
    double v = 0.0;
    for (int k = 0; k < 6; ++k) v += M[i * 6 + k] * s_t[21 + k];
    out[tid] = which ? -v : v;
  } else if (tid == 120) {
    out[120] = s_t[27];
  }
}

// Sum and maximum over the 64 lanes of a wave (butterfly; every lane gets the
// result, in a fixed order).
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Flat element e of the trailing lower triangle of an n x n factorisation,
// ordered by column descending: column k = n - 1 - m holds entries
// m (m + 1) / 2 ... m (m + 1) / 2 + m, rows k ... n - 1.  So the elements of
// the columns right of j are exactly e < (n - j - 1)(n - j) / 2.
__device__ inline void trailing_index(int e, int n, int& i, int& k) {
  int m = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
  while ((m + 1) * (m + 2) / 2 <= e) ++m;
  while (m * (m + 1) / 2 > e) --m;
  k = n - 1 - m;
  i = k + (e - m * (m + 1) / 2);
}

// Orders one wave's LDS accesses across the point (other lanes' writes before
// it are seen by reads after it); the compiler moves no LDS access over it.
__device__ __forceinline__ void panel_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane l's value of v (l uniform), moved through scalar registers.
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum and maximum over the four waves of a solve block (waves in order);
// every thread calls them and gets the result.  The leading barrier keeps a
// previous call's readers ahead of this call's writers.
__device__ inline double block_sum4(double v, double* s_red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  return ((s_red[0] + s_red[1]) + s_red[2]) + s_red[3];
}
__device__ inline double block_max4(double v, double* s_red) {
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  return fmax(fmax(s_red[0], s_red[1]), fmax(s_red[2], s_red[3]));
}

// H, g and F at the poses Xe from the edge blocks: one workgroup per 6x6 block
// (fi, fj) of H, then one per 6-block of g (the first also sums F); every sum
// runs over the edges in index order.  H goes to the buffer that is not the
// current linearisation.
__global__ __launch_bounds__(64) void ba_assemble_kernel(BaDev* __restrict__ ba, const BaEdge* __restrict__ edges,
                                                         const double* __restrict__ contrib,
                                                         double* __restrict__ Hbuf) {
  if (ba->done || !ba->eval_pending) return;
  const int nf = ba->nfree, n = ba->n, E = ba->n_edges;
  const int b = blockIdx.x, t = threadIdx.x;
  double* Hn = Hbuf + (size_t)(1 - ba->h_cur) * kBaMaxDof * kBaMaxDof;
  __shared__ int s_o1[1024], s_o2[1024];
  const bool hblock = b < nf * nf;
  const int fi = hblock ? b / nf : b - nf * nf;
  const int fj = hblock ? b - (b / nf) * nf : 0;
  const int a = t / 6, c = t % 6;
  double s = 0.0;
  for (int e0 = 0; e0 < E; e0 += 1024) {
    const int ne = min(1024, E - e0);
    __syncthreads();
    for (int q = t; q < ne; q += 64) {
      const BaEdge& ed = edges[e0 + q];
      s_o1[q] = ed.on ? ba->free_idx[ed.f1] : -2;
      s_o2[q] = ed.on ? ba->free_idx[ed.f2] : -2;
    }
    __syncthreads();
    for (int q = 0; q < ne; ++q) {
      const int o1 = s_o1[q], o2 = s_o2[q];
      if (o1 == -2) continue;
      const double* C = contrib + (size_t)(e0 + q) * kBaContrib;
      if (hblock) {
        if (t < 36) {
          if (o1 == fi && o1 == fj) s += C[a * 6 + c];
          if (o2 == fi && o2 == fj) s += C[72 + a * 6 + c];
          if (o1 == fi && o2 == fj) s += C[36 + a * 6 + c];
          if (o2 == fi && o1 == fj) s += C[36 + c * 6 + a];
        }
      } else if (t < 6) {
        if (o1 == fi) s += C[108 + t];
        if (o2 == fi) s += C[114 + t];
      } else if (t == 6 && fi == 0) {
        s += 0.5 * C[120];
      }
    }
  }
  if (hblock) {
    if (t < 36) Hn[(size_t)(6 * fi + a) * n + 6 * fj + c] = s;
  } else if (t < 6) {
    ba->gn[6 * fi + t] = s;
  } else if (t == 6 && fi == 0) {
    ba->Fn = s;
  }
}

// One block of four waves: the Levenberg-Marquardt decision on an evaluated
// candidate, then the next candidate (or the end of the solve).  Scalars of
// the schedule are kept uniform in every thread (block-wide reductions) and
// written back by thread 0; thread i < n holds entry i of the gradient; the
// damped system is factored in LDS (right-looking Cholesky, the trailing
// update over all four waves), the triangular solves run in wave 0 with the
// right-hand side in registers (rows lane and lane + 64).
__global__ __launch_bounds__(kSolveThreads) void ba_solve_kernel(BaDev* __restrict__ ba, double* __restrict__ Hbuf) {
  if (ba->done || !ba->eval_pending) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = ba->n;
  const int ld = n | 1;  // odd row stride in LDS
  __shared__ double s_A[kBaMaxDof * (kBaMaxDof + 1)];
  __shared__ double s_d[kBaMaxDof];
  __shared__ double s_g[kBaMaxDof];
  __shared__ double s_inv[kBaMaxDof];  // 1 / L(j, j)
  __shared__ double s_red[kSolveThreads / 64];
  __shared__ int s_fail;
  const bool candidate = ba->eval_is_candidate != 0;
  int cur = ba->h_cur;
  double mu = ba->mu, v = ba->v, F = ba->F;
  const double Fn = ba->Fn;
  int lm_iter = ba->lm_iter, lm_stop = ba->lm_stop;
  int accepted_total = ba->accepted_total, evals_total = ba->evals_total;
  const int max_lm = ba->iters_per_solve;
  bool take = !candidate;
  if (candidate) {
    // predicted decrease -(g.d + d.H.d / 2) at the current linearisation,
    // H read once, coalesced
    const double* H = Hbuf + (size_t)cur * kBaMaxDof * kBaMaxDof;
    if (tid < n) s_d[tid] = ba->d[tid];
    __syncthreads();
    double part = 0.0;
    // eight loads in flight per thread
#pragma unroll 8
    for (int idx = tid; idx < n * n; idx += kSolveThreads) {
      const int i = idx / n, j = idx - i * n;
      part += s_d[i] * H[idx] * s_d[j];
    }
    part *= 0.5;
    if (tid < n) part += s_d[tid] * ba->g[tid];
    const double pred = -block_sum4(part, s_red);
    const double rho = pred > 0.0 ? (F - Fn) / pred : -1.0;
    evals_total++;
    if (rho > 1e-3) {
      const double rel = fabs(F - Fn) / fmax(F, 1e-300);
      mu = fmin(mu / fmax(1.0 / 3.0, 1.0 - pow(2.0 * rho - 1.0, 3.0)), 1e16);
      v = 2.0;
      accepted_total++;
      if (rel < 1e-5) lm_stop = 1;
      take = true;
    } else {
      mu /= v;
      v *= 2.0;
    }
    lm_iter++;
  }
  // current linearisation: thread i's gradient entry, poses X
  const double* Xs = (take && candidate) ? ba->Xe : ba->X;
  double g = 0.0;
  if (take) {
    if (tid < n) {
      g = ba->gn[tid];
      ba->g[tid] = g;
    }
    F = Fn;
    cur = 1 - cur;
  } else {
    if (tid < n) g = ba->g[tid];
  }
  if (tid < n) s_g[tid] = g;
  double xs2 = 0.0;
  for (int i = tid; i < ba->n_frames * 12; i += kSolveThreads) {
    const double x = Xs[i];
    xs2 += x * x;
    if (take && candidate) ba->X[i] = x;
  }
  const double xn = sqrt(block_sum4(xs2, s_red));
  const double* H = Hbuf + (size_t)cur * kBaMaxDof * kBaMaxDof;
  int pending = 0;
  // every thread holds the same control values (block-wide reductions), so
  // the loop and its barriers are uniform
  while (true) {
    if (lm_stop || lm_iter >= max_lm) break;
    const double gmax = block_max4(fabs(g), s_red);
    if (gmax < 1e-5) {
      lm_stop = 1;
      break;
    }
#pragma unroll 8
    for (int idx = tid; idx < n * n; idx += kSolveThreads) {
      const int i = idx / n, j = idx - i * n;
      double x = H[idx];
      if (i == j) x += fmin(fmax(x, 1e-6), 1e32) / mu;
      s_A[i * ld + j] = x;
    }
    __syncthreads();
    // blocked right-looking Cholesky, kChol columns a block: wave 0 factors
    // the block's panel (rows below the block, its columns) alone -- a wave's
    // LDS operations complete in order, so wave-scope fences order it --
    // then all four waves subtract the block's columns from the trailing
    // triangle, per element in the same order as column by column (the
    // result is the unblocked one bit for bit); the diagonal's reciprocal
    // goes to s_inv
    bool fail = false;
    for (int j0 = 0; j0 < n; j0 += kChol) {
      const int jb = min(kChol, n - j0);
      if (wave == 0) {
        int f = 0;
        for (int c = 0; c < jb; ++c) {
          const int j = j0 + c;
          const double ajj = s_A[j * ld + j];
          if (!(ajj > 0.0)) {
            f = 1;
            break;
          }
          const double inv = 1.0 / sqrt(ajj);
          for (int i = j + 1 + lane; i < n; i += 64) s_A[i * ld + j] *= inv;
          if (lane == 0) s_inv[j] = inv;
          panel_lds_sync();  // the scaled column, read across lanes below
          for (int i = j + 1 + lane; i < n; i += 64) {
            const double lij = s_A[i * ld + j];
            for (int k = j + 1; k < j0 + jb && k <= i; ++k) s_A[i * ld + k] -= lij * s_A[k * ld + j];
          }
          panel_lds_sync();  // the next column's diagonal and entries
        }
        if (lane == 0) s_fail = f;
      }
      __syncthreads();
      if (s_fail) {  // uniform
        fail = true;
        break;
      }
      // the trailing triangle {(i, k): j0 + jb <= k <= i < n} as M flat
      // elements (the columns right of column j0 + jb - 1), four per thread
      // a trip with all reads issued before the writes
      const int rem = n - j0 - jb;
      const int M = rem * (rem + 1) / 2;
      for (int e0 = tid; e0 < M; e0 += 4 * kSolveThreads) {
        int ii[4], kk[4];
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = min(e0 + u * kSolveThreads, M - 1);
          trailing_index(e, n, ii[u], kk[u]);
          double a = s_A[ii[u] * ld + kk[u]];
          for (int c = 0; c < jb; ++c) a -= s_A[ii[u] * ld + j0 + c] * s_A[kk[u] * ld + j0 + c];
          x[u] = a;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (e0 + u * kSolveThreads < M) s_A[ii[u] * ld + kk[u]] = x[u];
      }
      __syncthreads();
    }
    if (fail) {
      mu /= v;
      v *= 2.0;
      lm_iter++;
      __syncthreads();
      continue;
    }
    // L y = -g, then L^T d = y: one wave, rows r and r + 64 per lane; the
    // solved entry moves by readlane, and the L entries and reciprocals of
    // kTri steps are read from LDS ahead of the steps that use them (rows
    // clamped to n - 1 for the loads; those values are not used)
    if (wave == 0) {
      const int r0 = lane, r1 = lane + 64;
      const int c0 = min(r0, n - 1), c1 = min(r1, n - 1);
      double b0 = r0 < n ? -s_g[r0] : 0.0, b1 = r1 < n ? -s_g[r1] : 0.0;
      for (int i0 = 0; i0 < n; i0 += kTri) {
        double l0[kTri], l1[kTri], sv[kTri];
#pragma unroll
        for (int u = 0; u < kTri; ++u) {
          const int i = min(i0 + u, n - 1);
          l0[u] = s_A[c0 * ld + i];
          l1[u] = s_A[c1 * ld + i];
          sv[u] = s_inv[i];
        }
#pragma unroll
        for (int u = 0; u < kTri; ++u) {
          const int i = i0 + u;
          if (i >= n) break;  // uniform
          const double bi = i < 64 ? readlane_f64(b0, i) : readlane_f64(b1, i - 64);
          const double yi = bi * sv[u];
          if (r0 == i) b0 = yi;
          if (r1 == i) b1 = yi;
          if (r0 > i && r0 < n) b0 -= l0[u] * yi;
          if (r1 > i && r1 < n) b1 -= l1[u] * yi;
        }
      }
      for (int i0 = n - 1; i0 >= 0; i0 -= kTri) {
        double l0[kTri], l1[kTri], sv[kTri];
#pragma unroll
        for (int u = 0; u < kTri; ++u) {
          const int i = max(i0 - u, 0);
          l0[u] = s_A[i * ld + c0];
          l1[u] = s_A[i * ld + c1];
          sv[u] = s_inv[i];
        }
#pragma unroll
        for (int u = 0; u < kTri; ++u) {
          const int i = i0 - u;
          if (i < 0) break;  // uniform
          const double bi = i < 64 ? readlane_f64(b0, i) : readlane_f64(b1, i - 64);
          const double xi = bi * sv[u];
          if (r0 == i) b0 = xi;
          if (r1 == i) b1 = xi;
          if (r0 < i) b0 -= l0[u] * xi;
          if (r1 < i) b1 -= l1[u] * xi;
        }
      }
      if (r0 < n) s_d[r0] = b0;
      if (r1 < n) s_d[r1] = b1;
    }
    __syncthreads();
    const double dd = tid < n ? s_d[tid] * s_d[tid] : 0.0;
    const double dn = sqrt(block_sum4(dd, s_red));
    if (dn < 1e-5 * (xn + 1e-5)) {
      lm_stop = 1;
      break;
    }
    if (tid < n) ba->d[tid] = s_d[tid];
    for (int f = tid; f < ba->n_frames; f += kSolveThreads) {
      const int o = ba->free_idx[f];
      if (o >= 0) pose_plus(Xs + 12 * f, s_d + 6 * o, ba->Xe + 12 * f);
      else
        for (int i = 0; i < 12; ++i) ba->Xe[12 * f + i] = Xs[12 * f + i];
    }
    pending = 1;
    break;
  }
  if (tid == 0) {
    ba->h_cur = cur;
    ba->mu = mu;
    ba->v = v;
    ba->F = F;
    ba->lm_iter = lm_iter;
    ba->lm_stop = lm_stop;
    ba->accepted_total = accepted_total;
    ba->evals_total = evals_total;
    ba->eval_pending = pending;
    ba->eval_is_candidate = pending;
  }
}

// Per-edge lengthscale (BinaryStateGPU::update_ell, IRLS_State_GPU.cpp:
// 117-146, mode 0), pose change (IRLS.cpp:99-108) and convergence
// (IRLS.cpp:325-339).  One wave: lanes over edges and frames.
__global__ __launch_bounds__(64) void ba_finish_kernel(BaDev* __restrict__ ba, BaEdge* __restrict__ edges) {
  if (ba->done) return;
  const int lane = threadIdx.x;
  int not_min = 0;
  for (int e = lane; e < ba->n_edges; e += 64) {
    BaEdge& ed = edges[e];
    ed.iters_per_ell++;
    if (ed.nnz_last > 0 && ed.nnz <= ed.nnz_last && ed.iters_per_ell / ba->iter_per_ell > 0) {
      ed.ell = ed.ell * (double)ba->decay_f;
      ed.iters_per_ell = 0;
    }
    if (ed.ell < ba->ell_min) ed.ell = ba->ell_min;
    else if (ed.ell > ba->ell_max) ed.ell = ba->ell_max;
    if (!(fabsf((float)ed.ell - ba->ell_min_f) < 1e-4f)) not_min = 1;
  }
  double change = 0.0;
  for (int f = lane; f < ba->n_frames; f += 64) {
    const double* A = ba->Xold + 12 * f;
    const double* B = ba->X + 12 * f;
    double R[9], t[3];
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c)
        R[r * 3 + c] = A[0 * 4 + r] * B[0 * 4 + c] + A[1 * 4 + r] * B[1 * 4 + c] + A[2 * 4 + r] * B[2 * 4 + c];
      t[r] = A[0 * 4 + r] * (B[3] - A[3]) + A[1 * 4 + r] * (B[7] - A[7]) + A[2 * 4 + r] * (B[11] - A[11]);
    }
    change += se3_log_norm(R, t);
  }
  change = wave_sum(change);
  const bool all_min = wave_max((double)not_min) == 0.0;
  if (lane == 0) {
    ba->change = change;
    ba->iter++;
    if ((all_min && change < 1e-5) || ba->iter >= ba->max_iters) ba->done = 1;
  }
}

// ---------------------------------------------------------------------------
// Host
// ---------------------------------------------------------------------------

template <class T>
static T* as(DevBuf& b) {
  return reinterpret_cast<T*>(b.p);
}

enum {
  BA_PAIRS = 0, BA_COUNTS, BA_OVF, BA_STATE, BA_PROBS, BA_PART_A, BA_PART_OVF, BA_PART_CNT,
  BA_DEV, BA_EDGES, BA_H, BA_HN, BA_PAIR_PART, BA_CONTRIB, BA_FLAG, BA_PAIR_IDS
};

// Six buffers per frame: sorted source (pos, feat, sem, geo), grid store and
// grid header.
constexpr int kFrameBufs = 6;

static cvo_status validate_frames(cvo_ctx* ctx, int n_frames, const cvo_dcloud* const* frames, const double* poses,
                                  const uint8_t* hold_const, int n_edges, const int32_t* edges, int* nfree_out) {
  const cvo_params& p = ctx->params;
  if (n_frames <= 0 || !frames || !poses) return set_err(ctx, CVO_ERR_INVALID_ARG, "no frames");
  if (n_frames > kBaMaxFrames) return set_err(ctx, CVO_ERR_LIMIT, "at most 64 frames");
  if (n_edges < 0 || n_edges > kBaMaxEdges) return set_err(ctx, CVO_ERR_LIMIT, "edge count out of [0, 4096]");
  if (n_edges > 0 && !edges) return set_err(ctx, CVO_ERR_INVALID_ARG, "null edge list");
  for (int f = 0; f < n_frames; ++f) {
    if (!frames[f]) return set_err(ctx, CVO_ERR_INVALID_ARG, "null frame");
    if (frames[f]->n == 0) return set_err(ctx, CVO_ERR_EMPTY, "empty frame");
    if (frames[f]->F != frames[0]->F || frames[f]->K != frames[0]->K)
      return set_err(ctx, CVO_ERR_INVALID_ARG, "frames differ in feature or class dimension");
    for (int i = 0; i < 12; ++i)
      if (!std::isfinite(poses[12 * f + i])) return set_err(ctx, CVO_ERR_INVALID_ARG, "non-finite pose");
  }
  int nfree = 0;
  for (int f = 0; f < n_frames; ++f)
    if (!hold_const || !hold_const[f]) nfree++;
  if (nfree > kBaMaxFree) return set_err(ctx, CVO_ERR_LIMIT, "at most 20 frames may be free");
  for (int e = 0; e < n_edges; ++e) {
    const int a = edges[2 * e], b = edges[2 * e + 1];
    if (a < 0 || a >= n_frames || b < 0 || b >= n_frames || a == b)
      return set_err(ctx, CVO_ERR_INVALID_ARG, "edge frame index out of range or self edge");
    CVO_TRY(validate_pair(ctx, frames[a], frames[b]));
  }
  if (p.multiframe_num_neighbors < 1 || p.multiframe_num_neighbors > 1024)
    return set_err(ctx, CVO_ERR_LIMIT, "multiframe_num_neighbors must be in [1, 1024]");
  // the K-nearest search keeps at most KDTREE_K_SIZE per row (cukdtree.h:12,40)
  if (p.is_using_kdtree && p.multiframe_num_neighbors > kKnnMaxK)
    return set_err(ctx, CVO_ERR_LIMIT, "is_using_kdtree needs multiframe_num_neighbors <= 512");
  if (p.multiframe_iter_per_ell <= 0) return set_err(ctx, CVO_ERR_INVALID_ARG, "multiframe_iter_per_ell must be > 0");
  if (!(p.multiframe_ell_init > 0.0f) || !(p.multiframe_ell_min > 0.0f))
    return set_err(ctx, CVO_ERR_INVALID_ARG, "multiframe lengthscales must be positive");
  *nfree_out = nfree;
  return CVO_OK;
}


struct BaLaunch {
  BaDev* ba;
  BaEdge* edges;
  AssocProblem* probs;
  int n_edges, total_blocks, max_pair_blocks, lm_rounds, n_asm_blocks;
  double* part;
  double* contrib;
  double* H;  // two n x n halves: current linearisation and evaluation
  KConsts kc;
};

static void ba_assoc_phase(const BaLaunch& L, hipStream_t s) {
  hipLaunchKernelGGL(ba_prepare_kernel, dim3(1), dim3(256), 0, s, L.ba, L.edges, L.probs);
  if (L.n_edges > 0) {
    launch_assoc(L.probs, L.n_edges, L.total_blocks, L.kc, s);
    launch_overflow(L.probs, L.n_edges, L.kc, s);
  }
  hipLaunchKernelGGL(ba_edges_kernel, dim3(1), dim3(256), 0, s, L.ba, L.edges, L.probs);
}

static void ba_lm_round(const BaLaunch& L, hipStream_t s) {
  hipLaunchKernelGGL(ba_pair_kernel, dim3(L.max_pair_blocks, L.n_edges), dim3(kBaRowsPerBlock), 0, s, L.ba,
                     L.edges, L.probs, L.part, L.max_pair_blocks);
  hipLaunchKernelGGL(ba_edge_reduce_kernel, dim3(L.n_edges), dim3(256), 0, s, L.ba, L.edges, L.part,
                     L.max_pair_blocks, L.contrib);
  hipLaunchKernelGGL(ba_assemble_kernel, dim3(L.n_asm_blocks), dim3(64), 0, s, L.ba, L.edges, L.contrib, L.H);
  hipLaunchKernelGGL(ba_solve_kernel, dim3(1), dim3(kSolveThreads), 0, s, L.ba, L.H);
}

static void ba_iteration(const BaLaunch& L, hipStream_t s) {
  ba_assoc_phase(L, s);
  if (L.n_edges > 0)
    for (int k = 0; k < L.lm_rounds; ++k) ba_lm_round(L, s);
  hipLaunchKernelGGL(ba_finish_kernel, dim3(1), dim3(64), 0, s, L.ba, L.edges);
}

}  // namespace cvo

using namespace cvo;

extern "C" cvo_status cvo_align_multiframe(cvo_ctx* ctx, int32_t n_frames, const cvo_dcloud* const* frames,
                                           double* poses, const uint8_t* hold_const, int32_t n_edges,
                                           const int32_t* edges, cvo_assoc* per_edge, double* seconds,
                                           int32_t* iters, double* per_iter_seconds, int32_t max_iter_records) {
  if (!ctx) return CVO_ERR_INVALID_ARG;
  if (per_edge)
    for (int e = 0; e < n_edges; ++e) std::memset(&per_edge[e], 0, sizeof(cvo_assoc));
  if (seconds) *seconds = 0.0;
  if (iters) *iters = 0;
  int nfree = 0;
  CVO_TRY(validate_frames(ctx, n_frames, frames, poses, hold_const, n_edges, edges, &nfree));
  const cvo_params& p = ctx->params;
  hipStream_t st = ctx->stream;
  const auto t_start = std::chrono::steady_clock::now();
  KConsts kc = make_kconsts(p, frames[0]->F, frames[0]->K, MODE_BA);
  apply_tuning(ctx, kc, true);  // one batched launch over the edges
  const int rb = rows_per_block(kc.max_group);

  // frames that are sources / targets of some edge
  std::vector<char> is_src(n_frames, 0), is_tgt(n_frames, 0);
  for (int e = 0; e < n_edges; ++e) {
    is_src[edges[2 * e]] = 1;
    is_tgt[edges[2 * e + 1]] = 1;
  }
  float max_world = 0.0f, max_t = 0.0f;
  for (int f = 0; f < n_frames; ++f) {
    const double* T = poses + 12 * f;
    const float tn = (float)std::sqrt(T[3] * T[3] + T[7] * T[7] + T[11] * T[11]);
    max_world = std::max(max_world, frames[f]->max_norm + tn);
    max_t = std::max(max_t, tn);
  }
  if (ctx->b_frames.size() < (size_t)kFrameBufs * n_frames) ctx->b_frames.resize((size_t)kFrameBufs * n_frames);
  std::vector<GridHost> grids(n_frames);
  const float ell_hi = std::max(p.multiframe_ell_init * 1.2f, p.multiframe_ell_min);
  const float ell_lo = std::min(p.multiframe_ell_min, p.multiframe_ell_init);
  const std::vector<float> hs = level_sizes(kc, ell_hi, ell_lo, max_world * 1.05f, max_t * 1.05f + 1.0f, kc.min_scale);
  for (int f = 0; f < n_frames; ++f) {
    DevBuf* fb = &ctx->b_frames[(size_t)kFrameBufs * f];
    if (is_src[f]) CVO_TRY(prepare_source_into(ctx, frames[f], fb[0], fb[1], fb[2], fb[3]));
    if (is_tgt[f]) CVO_TRY(build_grid_into(ctx, frames[f], hs, grids[f], fb[4], fb[5]));
  }

  // per-edge association problems
  const int cap = p.multiframe_num_neighbors;
  size_t tot_rows = 0;
  int total_blocks = 0, max_pair_blocks = 1;
  std::vector<size_t> row_off(n_edges + 1, 0);
  for (int e = 0; e < n_edges; ++e) {
    const int n = frames[edges[2 * e]]->n;
    row_off[e + 1] = row_off[e] + (size_t)n;
    total_blocks += (n + rb - 1) / rb;
    max_pair_blocks = std::max(max_pair_blocks, (n + kBaRowsPerBlock - 1) / kBaRowsPerBlock);
  }
  tot_rows = row_off[n_edges];
  const size_t E1 = (size_t)std::max(n_edges, 1);
  DevBuf* B = ctx->b_ba;
  CVO_TRY(ensure(ctx, B[BA_PAIRS], sizeof(float4) * std::max<size_t>(tot_rows, 1) * cap));
  CVO_TRY(ensure(ctx, B[BA_COUNTS], sizeof(uint32_t) * std::max<size_t>(tot_rows, 1)));
  CVO_TRY(ensure(ctx, B[BA_OVF], sizeof(uint32_t) * std::max<size_t>(tot_rows, 1)));
  CVO_TRY(ensure(ctx, B[BA_STATE], sizeof(ProblemState) * E1));
  CVO_TRY(ensure(ctx, B[BA_PROBS], sizeof(AssocProblem) * E1));
  CVO_TRY(ensure(ctx, B[BA_PART_A], sizeof(double) * 8 * (size_t)std::max(total_blocks, 1)));
  CVO_TRY(ensure(ctx, B[BA_PART_OVF], sizeof(double) * 10 * kOvfBlocks * E1));
  CVO_TRY(ensure(ctx, B[BA_PART_CNT], sizeof(uint32_t) * 2 * (size_t)std::max(total_blocks, 1)));
  CVO_TRY(ensure(ctx, B[BA_DEV], sizeof(BaDev)));
  CVO_TRY(ensure(ctx, B[BA_EDGES], sizeof(BaEdge) * E1));
  CVO_TRY(ensure(ctx, B[BA_H], sizeof(double) * 2 * kBaMaxDof * kBaMaxDof));
  CVO_TRY(ensure(ctx, B[BA_PAIR_PART], sizeof(double) * kBaTerms * (size_t)max_pair_blocks * E1));
  CVO_TRY(ensure(ctx, B[BA_CONTRIB], sizeof(double) * kBaContrib * E1));
  // the export needs each stored pair's target (its sorted position)
  if (per_edge) CVO_TRY(ensure(ctx, B[BA_PAIR_IDS], sizeof(uint32_t) * std::max<size_t>(tot_rows, 1) * cap));
  CVO_HIP(ctx, hipMemsetAsync(B[BA_OVF].p, 0, sizeof(uint32_t) * std::max<size_t>(tot_rows, 1), st));
  CVO_HIP(ctx, hipMemsetAsync(B[BA_STATE].p, 0, sizeof(ProblemState) * E1, st));

  std::vector<AssocProblem> hp(E1);
  std::vector<BaEdge> he(E1);
  int blk = 0;
  for (int e = 0; e < n_edges; ++e) {
    const int f1 = edges[2 * e], f2 = edges[2 * e + 1];
    const cvo_dcloud* s = frames[f1];
    const cvo_dcloud* t = frames[f2];
    DevBuf* sb = &ctx->b_frames[(size_t)kFrameBufs * f1];
    const int n = s->n;
    const int nb = (n + rb - 1) / rb;
    AssocProblem& P = hp[e];
    std::memset(&P, 0, sizeof(P));
    P.src_pos = as<float4>(sb[0]);
    P.src_feat = s->feat ? as<float>(sb[1]) : nullptr;
    P.src_sem = s->sem ? as<float>(sb[2]) : nullptr;
    P.src_geo = s->geo ? as<float>(sb[3]) : nullptr;
    P.n_src = n;
    P.n_tgt = t->n;
    P.grid = grids[f2].dev_ptr;
    P.tgt_xyz = t->xyz;
    P.tgt_feat = t->feat;
    P.tgt_sem = t->sem;
    P.tgt_geo = t->geo;
    P.st = as<ProblemState>(B[BA_STATE]) + e;
    P.pairs = nullptr;
    P.pairs4 = as<float4>(B[BA_PAIRS]) + row_off[e] * cap;
    P.pair_ids = per_edge ? as<uint32_t>(B[BA_PAIR_IDS]) + row_off[e] * cap : nullptr;
    P.counts = as<uint32_t>(B[BA_COUNTS]) + row_off[e];
    P.ovf_stamp = as<uint32_t>(B[BA_OVF]) + row_off[e];
    P.cap_alloc = cap;
    P.block_begin = blk;
    P.n_blocks = nb;
    P.part_a = as<double>(B[BA_PART_A]) + (size_t)blk * 8;
    P.part_ovf = as<double>(B[BA_PART_OVF]) + (size_t)e * kOvfBlocks * 10;
    P.part_s = nullptr;
    P.part_cnt = as<uint32_t>(B[BA_PART_CNT]) + (size_t)blk * 2;
    blk += nb;
    BaEdge& ed = he[e];
    std::memset(&ed, 0, sizeof(ed));
    ed.ell = (double)p.multiframe_ell_init;
    ed.f1 = f1;
    ed.f2 = f2;
    ed.K = cap;
    ed.n_rows = n;
    ed.n_pair_blocks = (n + kBaRowsPerBlock - 1) / kBaRowsPerBlock;
  }
  if (n_edges > 0) {
    CVO_HIP(ctx, hipMemcpyAsync(B[BA_PROBS].p, hp.data(), sizeof(AssocProblem) * n_edges, hipMemcpyHostToDevice, st));
    CVO_HIP(ctx, hipMemcpyAsync(B[BA_EDGES].p, he.data(), sizeof(BaEdge) * n_edges, hipMemcpyHostToDevice, st));
  }
  BaDev* hb = new BaDev;
  std::memset(hb, 0, sizeof(BaDev));
  hb->n_frames = n_frames;
  hb->n_edges = n_edges;
  hb->nfree = nfree;
  hb->n = 6 * nfree;
  hb->max_iters = p.multiframe_max_iters;
  hb->iters_per_solve = p.multiframe_iterations_per_solve;
  hb->K_init = cap;
  hb->min_nonzeros = p.multiframe_min_nonzeros;
  hb->iter_per_ell = p.multiframe_iter_per_ell;
  hb->ell_min_f = p.multiframe_ell_min;
  hb->decay_f = p.multiframe_ell_decay_rate;
  hb->rq_coef = kc.rq_coef;
  hb->pol = kc.pol;
  hb->ell_min = (double)p.multiframe_ell_min;
  hb->ell_max = (double)p.multiframe_ell_init * 1.2;
  int o = 0;
  for (int f = 0; f < n_frames; ++f) {
    hb->free_idx[f] = (!hold_const || !hold_const[f]) ? o++ : -1;
    if (hb->free_idx[f] >= 0) hb->frame_of_free[hb->free_idx[f]] = f;
    hb->frame_norm[f] = frames[f]->max_norm;
    for (int i = 0; i < 12; ++i) hb->X[12 * f + i] = poses[12 * f + i];
  }
  hb->done = p.multiframe_max_iters <= 0 ? 1 : 0;
  hipError_t ce = hipMemcpyAsync(B[BA_DEV].p, hb, sizeof(BaDev), hipMemcpyHostToDevice, st);
  hipError_t se = hipStreamSynchronize(st);  // hb is pageable host memory
  delete hb;
  if (ce != hipSuccess) return hip_check(ctx, ce, "BA state upload");
  if (se != hipSuccess) return hip_check(ctx, se, "BA state upload");

  BaLaunch L;
  L.ba = as<BaDev>(B[BA_DEV]);
  L.edges = as<BaEdge>(B[BA_EDGES]);
  L.probs = as<AssocProblem>(B[BA_PROBS]);
  L.n_edges = n_edges;
  L.total_blocks = total_blocks;
  L.max_pair_blocks = max_pair_blocks;
  L.lm_rounds = std::max(0, p.multiframe_iterations_per_solve) + 1;
  L.n_asm_blocks = std::max(1, nfree * nfree + nfree);
  L.part = as<double>(B[BA_PAIR_PART]);
  L.contrib = as<double>(B[BA_CONTRIB]);
  L.H = as<double>(B[BA_H]);
  L.kc = kc;

  CVO_TRY(pinned(ctx, 64));
  volatile int32_t* flag = reinterpret_cast<volatile int32_t*>(ctx->pinned);
  const char* ng = std::getenv("CVO_NO_GRAPH");
  const bool use_graph = !(ng && ng[0] && ng[0] != '0') && !ctx->profiling;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (use_graph) {
    CVO_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    ba_iteration(L, st);
    hipError_t e1 = hipStreamEndCapture(st, &graph);
    if (e1 != hipSuccess) return hip_check(ctx, e1, "hipStreamEndCapture");
    hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e2 != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return hip_check(ctx, e2, "hipGraphInstantiate");
    }
  }
  CVO_TRY(ensure_events(ctx, 24));
  hipEvent_t ev[6];
  for (int i = 0; i < 6; ++i) ev[i] = ctx->ev_pool[16 + i];
  cvo_status status = CVO_OK;
  const int max_outer = std::max(0, p.multiframe_max_iters) + 1;
  for (int it = 0; it < max_outer; ++it) {
    const auto ti0 = std::chrono::steady_clock::now();
    hipError_t le = hipSuccess;
    if (use_graph) {
      le = hipGraphLaunch(exec, st);
    } else if (ctx->profiling) {
      (void)hipEventRecord(ev[0], st);
      hipLaunchKernelGGL(ba_prepare_kernel, dim3(1), dim3(256), 0, st, L.ba, L.edges, L.probs);
      // association kernel timed alone (events at its own start and end)
      if (n_edges > 0) launch_assoc(L.probs, n_edges, total_blocks, kc, st, nullptr, ev[1], ev[2]);
      else (void)hipEventRecord(ev[1], st), (void)hipEventRecord(ev[2], st);
      if (n_edges > 0) {
        launch_overflow(L.probs, n_edges, kc, st);
      }
      hipLaunchKernelGGL(ba_edges_kernel, dim3(1), dim3(256), 0, st, L.ba, L.edges, L.probs);
      (void)hipEventRecord(ev[3], st);
      // algorithmic bytes: association reads both clouds' points and writes
      // the pairs; a pair pass reads the stored pairs, the target positions
      // they name and the source rows
      std::vector<BaEdge> eh(n_edges);
      if (n_edges > 0)
        (void)hipMemcpyAsync(eh.data(), L.edges, sizeof(BaEdge) * n_edges, hipMemcpyDeviceToHost, st);
      (void)hipStreamSynchronize(st);
      float m_assoc = 0, m_ovf = 0;
      (void)hipEventElapsedTime(&m_assoc, ev[1], ev[2]);
      (void)hipEventElapsedTime(&m_ovf, ev[2], ev[3]);
      double pair_bytes = 0.0;
      const double per_pt = 12.0 + (kc.use_col ? 4.0 * kc.F : 0.0) + (kc.use_sem ? 4.0 * kc.Ks : 0.0);
      for (int e = 0; e < n_edges; ++e) {
        const double S = (double)eh[e].nnz;
        ctx->stats.bytes[CVO_KSTAT_ASSOC] += per_pt * (frames[eh[e].f1]->n + frames[eh[e].f2]->n) + 16.0 * S +
                                             4.0 * frames[eh[e].f1]->n;
        ctx->stats.pairs_stored += S;
        if (eh[e].on) pair_bytes += 16.0 * S + 20.0 * eh[e].n_rows;
      }
      float ms_pair = 0.0f, ms_solve = 0.0f;
      int pair_launches = 0;
      for (int k = 0; k < (n_edges > 0 ? L.lm_rounds : 0); ++k) {
        int32_t pend[2] = {0, 0};
        (void)hipMemcpyAsync(pend, &L.ba->eval_pending, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, st);
        (void)hipStreamSynchronize(st);
        const bool work = pend[0] != 0;
        // the pair kernel timed alone
        hipExtLaunchKernelGGL(ba_pair_kernel, dim3(L.max_pair_blocks, n_edges), dim3(kBaRowsPerBlock), 0, st, ev[4],
                              ev[5], 0, L.ba, L.edges, L.probs, L.part, L.max_pair_blocks);
        hipLaunchKernelGGL(ba_edge_reduce_kernel, dim3(n_edges), dim3(256), 0, st, L.ba, L.edges, L.part,
                           L.max_pair_blocks, L.contrib);
        hipLaunchKernelGGL(ba_assemble_kernel, dim3(L.n_asm_blocks), dim3(64), 0, st, L.ba, L.edges, L.contrib,
                           L.H);
        hipLaunchKernelGGL(ba_solve_kernel, dim3(1), dim3(kSolveThreads), 0, st, L.ba, L.H);
        (void)hipEventRecord(ev[3], st);
        (void)hipEventSynchronize(ev[3]);
        float a = 0, b = 0;
        (void)hipEventElapsedTime(&a, ev[4], ev[5]);
        (void)hipEventElapsedTime(&b, ev[5], ev[3]);
        if (work) {
          ms_pair += a;
          ms_solve += b;
          pair_launches++;
          ctx->stats.bytes[CVO_KSTAT_BA_PAIR] += pair_bytes;
        }
      }
      hipLaunchKernelGGL(ba_finish_kernel, dim3(1), dim3(64), 0, st, L.ba, L.edges);
      ctx->stats.ms[CVO_KSTAT_ASSOC] += m_assoc;
      ctx->stats.launches[CVO_KSTAT_ASSOC] += 1;
      ctx->stats.ms[CVO_KSTAT_OVERFLOW] += m_ovf;
      ctx->stats.launches[CVO_KSTAT_OVERFLOW] += 1;
      ctx->stats.ms[CVO_KSTAT_BA_PAIR] += ms_pair;
      ctx->stats.launches[CVO_KSTAT_BA_PAIR] += pair_launches;
      ctx->stats.ms[CVO_KSTAT_BA_SOLVE] += ms_solve;
      ctx->stats.launches[CVO_KSTAT_BA_SOLVE] += pair_launches;
      le = hipGetLastError();
    } else {
      ba_iteration(L, st);
      le = hipGetLastError();
    }
    if (le != hipSuccess) {
      status = hip_check(ctx, le, "BA iteration launch");
      break;
    }
    // BaDev::iter and BaDev::done are adjacent
    hipError_t me = hipMemcpyAsync((void*)flag, &L.ba->iter, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, st);
    if (me == hipSuccess) me = hipStreamSynchronize(st);
    if (me != hipSuccess) {
      status = hip_check(ctx, me, "BA flag readback");
      break;
    }
    const auto ti1 = std::chrono::steady_clock::now();
    if (per_iter_seconds && it < max_iter_records) per_iter_seconds[it] = std::chrono::duration<double>(ti1 - ti0).count();
    if (flag[1]) break;
  }
  if (use_graph) {
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
  }
  if (status != CVO_OK) return status;
  BaDev* out = new BaDev;
  hipError_t re = hipMemcpy(out, L.ba, sizeof(BaDev), hipMemcpyDeviceToHost);
  if (re != hipSuccess) {
    delete out;
    return hip_check(ctx, re, "BA state readback");
  }
  for (int f = 0; f < n_frames; ++f)
    for (int i = 0; i < 12; ++i) poses[12 * f + i] = out->X[12 * f + i];
  if (iters) *iters = out->iter;
  ctx->ba_last_accepted = out->accepted_total;
  ctx->ba_last_evals = out->evals_total;
  delete out;
  // with multiframe_max_iters > 0 the first outer iteration associates
  const bool any_pass = p.multiframe_max_iters > 0;
  // each edge's pairs of the last association pass (BinaryStateGPU::
  // export_association, IRLS_State_GPU.cu:345-351; CvoGPU.cu:1990-1996)
  if (per_edge && any_pass)
    for (int e = 0; e < n_edges; ++e)
      CVO_TRY(export_problem(ctx, L.probs + e, frames[edges[2 * e]]->n, frames[edges[2 * e + 1]]->n, &per_edge[e]));
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  return CVO_OK;
}
