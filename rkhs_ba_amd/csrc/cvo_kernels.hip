// gfx950 kernels of the CVO two-view loop.
//
//   grid_cell_count / grid_gather : uniform-grid levels over a target cloud
//   src_sort_keys / src_gather    : spatially sorted copy of the source rows
//   assoc_kernel                  : ell-ball search + product kernel + row store
//                                   (first K by target index) + fused flow
//                                   (fill_in_A_mat_gpu CvoGPU.cu:951-1099 and
//                                   compute_flow_gpu_no_eigen :2001-2062)
//   overflow_kernel               : the pass totals, plus the rows whose
//                                   passing pairs exceed the cap K (their K
//                                   smallest passing target indices)
//   step_kernel                   : Taylor step terms (:333-462); its last
//                                   block runs the device controller
//                                   (:464-544, :1289-1433)
//
// Rows are processed by groups of G = 8 or 16 lanes (256 / G rows per
// 256-thread block), G chosen per iteration from the lengthscale.  Reductions
// are block partials summed in fixed order, so results are bitwise
// reproducible run to run.

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "cvo_internal.h"
#include "cvo_device.h"
#include "cvo_ball.h"

namespace cvo {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

// Workgroups are dispatched to the eight XCDs round robin by id (XCD = id % 8).
// This maps a dispatch id to a logical block so that every XCD takes one
// contiguous range of logical blocks: spatially sorted rows whose overlapping
// target neighbourhoods its own L2 then serves.  A bijection of [0, nblk).
__device__ __forceinline__ int xcd_block(int bid, int nblk) {
  const int q = nblk >> 3, r = nblk & 7;
  const int x = bid & 7, k = bid >> 3;
  return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Cross-block publication without cache maintenance.  Partials are written
// with device-scope stores (sc1: written through to the device coherence
// point) by lanes of wave 0, which then waits for its stores (vmcnt(0))
// before lane 0 counts the block in with a device-scope atomic; the block
// that counts last reads the partials with device-scope loads.  An agent
// fence would instead write back and invalidate the whole XCD L2 in every
// block, which costs the association kernel its cached target neighbourhoods.
__device__ __forceinline__ void st_dev(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_dev(double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_dev(uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t count_in(uint32_t* counter) {
  return __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima), between two
// compiler-only fences: the builtin orders the hardware, the fences keep the
// compiler from moving the partial stores past it or the completion atomic
// before it (neither fence emits an instruction or a cache operation).
__device__ __forceinline__ void wait_stores() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Exact evaluation of a pair's kernel value with FP64 exponentials rounded to
// float, the reference's rounding points (CvoGPU.cu:1020-1054).  Used only for
// pairs whose fast FP32 value lies within 2e-5 relative of sp_thres, so the
// stored set equals the reference's.
__device__ __noinline__ float exact_value(float sigma2, float c_sigma2, float s_sigma2, float c2, float s_ell,
                                         int use_geo, int use_col, int use_sem, float d2, float l, float d2c,
                                         float d2s, float geo_sim) {
  float k = 1.0f, ck = 1.0f, sk = 1.0f;
  if (use_geo) k = (float)((double)sigma2 * exp((double)(-d2) / (2.0 * (double)l * (double)l)));
  if (use_col) ck = (float)((double)c_sigma2 * exp((double)(-d2c) / (2.0 * (double)c2)));
  if (use_sem) sk = (float)((double)s_sigma2 * exp((double)(-d2s) / (2.0 * (double)s_ell * (double)s_ell)));
  return ck * k * sk * geo_sim;
}

__device__ __forceinline__ float geo_similarity(const float* ga, const float* gb) {
  float n2a = fmaf(ga[1], ga[1], ga[0] * ga[0]);
  float n2b = fmaf(gb[1], gb[1], gb[0] * gb[0]);
  float dab = fmaf(ga[1], gb[1], ga[0] * gb[0]);
  return dab * dab / (n2a * n2b);
}

// |y - x|^2 with the association's FMA order.
__device__ __forceinline__ float sq_dist(float x0, float x1, float x2, float y0, float y1, float y2) {
  const float dx = y0 - x0, dy = y1 - x1, dz = y2 - x2;
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// Locate the problem of a flat block index.
__device__ __forceinline__ int find_problem(const AssocProblem* probs, int n_probs, int blk) {
  int lo = 0, hi = n_probs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (probs[mid].block_begin <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// Grid construction
// ---------------------------------------------------------------------------

__global__ void grid_cell_count(const float* __restrict__ xyz, int n, float ox, float oy, float oz, float inv_h,
                                int nx, int ny, int nz, uint32_t* __restrict__ key, uint32_t* __restrict__ count) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int cx = (int)floorf((xyz[3 * j + 0] - ox) * inv_h);
  int cy = (int)floorf((xyz[3 * j + 1] - oy) * inv_h);
  int cz = (int)floorf((xyz[3 * j + 2] - oz) * inv_h);
  cx = min(max(cx, 0), nx - 1);
  cy = min(max(cy, 0), ny - 1);
  cz = min(max(cz, 0), nz - 1);
  uint32_t c = ((uint32_t)cz * (uint32_t)ny + (uint32_t)cy) * (uint32_t)nx + (uint32_t)cx;
  key[j] = c;
  atomicAdd(&count[c], 1u);
}

// Zero fill, in place of a memset node (which a caller's graph capture
// should not record).
__global__ void zero_u32(uint32_t* __restrict__ v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) v[i] = 0u;
}

__global__ void iota_u32(uint32_t* __restrict__ v, int n) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) v[j] = (uint32_t)j;
}

// Gathers the cloud into sorted order: pos = (x, y, z, original id bits).
__global__ void gather_sorted(const uint32_t* __restrict__ order, int n, const float* __restrict__ xyz,
                              const float* __restrict__ feat, int F, const float* __restrict__ sem, int K,
                              const float* __restrict__ geo, float4* __restrict__ pos, float* __restrict__ feat_s,
                              float* __restrict__ sem_s, float* __restrict__ geo_s, uint32_t* __restrict__ inv_perm) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint32_t j = order[s];
  pos[s] = make_float4(xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2], __uint_as_float(j));
  if (feat_s)
    for (int f = 0; f < F; ++f) feat_s[(size_t)s * F + f] = feat[(size_t)j * F + f];
  if (sem_s)
    for (int k = 0; k < K; ++k) sem_s[(size_t)s * K + k] = sem[(size_t)j * K + k];
  if (geo_s) {
    geo_s[2 * (size_t)s] = geo[2 * (size_t)j];
    geo_s[2 * (size_t)s + 1] = geo[2 * (size_t)j + 1];
  }
  if (inv_perm) inv_perm[j] = (uint32_t)s;
}

// 30-bit Morton key of a point in its cloud's bounding box (source row order).
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ void morton_keys(const float* __restrict__ xyz, int n, float ox, float oy, float oz, float sx, float sy,
                            float sz, uint32_t* __restrict__ key) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t a = (uint32_t)fminf(fmaxf((xyz[3 * j] - ox) * sx, 0.0f), 1023.0f);
  uint32_t b = (uint32_t)fminf(fmaxf((xyz[3 * j + 1] - oy) * sy, 0.0f), 1023.0f);
  uint32_t c = (uint32_t)fminf(fmaxf((xyz[3 * j + 2] - oz) * sz, 0.0f), 1023.0f);
  key[j] = spread10(a) | (spread10(b) << 1) | (spread10(c) << 2);
}

// ---------------------------------------------------------------------------
// Association + fused flow
// ---------------------------------------------------------------------------

// Loads a scalar feature (compile-time or runtime dimension).
template <int N>
struct Dim {
  __device__ static constexpr int get(int) { return N; }
};
template <>
struct Dim<-1> {
  __device__ static int get(int rt) { return rt; }
};

// Float maps of one iteration, loaded once per block from the problem state.
struct Maps {
  float A[9], b[3];    // target: y' = A y + b
  float Rq[9], Tq[3];  // query into the target frame
  float As[9], bs[3];  // source map
  bool src_xf;
};

__device__ __forceinline__ void load_maps(const ProblemState* st, Maps& m) {
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    m.A[i] = st->A[i];
    m.Rq[i] = st->Rq[i];
    m.As[i] = st->As[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    m.b[i] = st->b[i];
    m.Tq[i] = st->Tq[i];
    m.bs[i] = st->bs[i];
  }
  m.src_xf = st->has_src_xf != 0;
}

// Scalars of the exact path (passed by value: no private copy of KConsts).
struct ExactK {
  float sigma2, c_sigma2, s_sigma2, c2, s_ell;
  int use_geo, use_col, use_sem;
};

__device__ __forceinline__ ExactK exact_consts(const KConsts& kc) {
  ExactK e;
  e.sigma2 = kc.sigma2;
  e.c_sigma2 = kc.c_sigma2;
  e.s_sigma2 = kc.s_sigma2;
  e.c2 = kc.c2;
  e.s_ell = kc.s_ell;
  e.use_geo = kc.use_geo && !kc.no_geo_factor;
  e.use_col = kc.use_col;
  e.use_sem = kc.use_sem;
  return e;
}

// The kernel value of a candidate pair: gates, then the product kernel.
// Returns false when a gate rejects or a <= sp_thres.  gmin: the geometric
// type similarity below which a pair is rejected (0.01 in fill_in_A_mat_gpu,
// CvoGPU.cu:1015; 0.001 in fill_in_A_mat_cukdtree, :898).
template <int FC, int KC>
__device__ __forceinline__ bool pair_kernel(const KConsts& kc, const ExactK& ek, float x0, float x1, float x2,
                                            float y0, float y1, float y2, float d2t, float qg, float l,
                                            const float* xf, const float* yf, const float* xs, const float* ys,
                                            const float* xg, const float* yg, float& a, float gmin = 0.01f) {
  const float dx = y0 - x0, dy = y1 - x1, dz = y2 - x2;
  const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  if (kc.use_geo && !(d2 < d2t)) return false;
  float geo_sim = 1.0f;
  if (kc.use_gtype) {
    // yg always addresses the target cloud's type vectors in global memory
    const float ygl[2] = {ld_global(yg), ld_global(yg + 1)};
    geo_sim = geo_similarity(xg, ygl);
    if (geo_sim < gmin) return false;
  }
  float d2c = 0.0f, d2s = 0.0f;
  if (FC != 0) {
    const int F = Dim<FC>::get(kc.F);
#pragma unroll
    for (int f = 0; f < F; ++f) {
      float t = xf[f] - yf[f];
      d2c = fmaf(t, t, d2c);
    }
    if (!(d2c < kc.d2_c_thres)) return false;
  }
  if (KC != 0) {
    const int KS = Dim<KC>::get(kc.Ks);
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      float t = xs[k] - ys[k];
      d2s = fmaf(t, t, d2s);
    }
    if (!(d2s < kc.d2_s_thres)) return false;
  }
  float ex = (kc.use_geo && !kc.no_geo_factor) ? d2 * qg : 0.0f;
  if (FC != 0) ex = fmaf(d2c, kc.q_c, ex);
  if (KC != 0) ex = fmaf(d2s, kc.q_s, ex);
  a = kc.amp * expf(-ex) * geo_sim;
  if (fabsf(a - kc.sp) <= 2e-5f * kc.sp)
    a = exact_value(ek.sigma2, ek.c_sigma2, ek.s_sigma2, ek.c2, ek.s_ell, ek.use_geo, ek.use_col, ek.use_sem, d2, l,
                    d2c, d2s, geo_sim);
  return a > kc.sp;
}

// Stores one kept pair of row r in slot `slot`: the bundle adjustment keeps
// the target's own coordinates and a at the iteration's row stride K (and,
// when asked, the target's sorted position for an association export);
// the other modes keep (sorted target position, a) at the allocated stride,
// unless the pass only sums (no pair array).
// NOBA: the launch never carries a bundle-adjustment problem (single-problem
// kernels), so that store is compiled out.
template <bool NOBA = false>
__device__ __forceinline__ void store_pair(const AssocProblem& P, int mode, int r, uint32_t K, uint32_t slot,
                                           uint32_t pidx, const float4& y4, float a) {
  if (!NOBA && mode == MODE_BA) {
    st_global(P.pairs4 + (size_t)r * K + slot, make_float4(y4.x, y4.y, y4.z, a));
    if (P.pair_ids) st_global(P.pair_ids + (size_t)r * K + slot, pidx);
  } else if (P.pairs) {
    st_global(P.pairs + (size_t)r * P.cap_alloc + slot, make_uint2(pidx, __float_as_uint(a)));
  }
}

struct RowResult {
  double acc[7];  // omega(3), v(3), ip
  uint32_t stored;
  bool overflow;
};

// One row handled by a group of G lanes.  FC/KC: color/semantic dimension when the
// gate is on (0 off, -1 runtime).  MODE_ALIGN accumulates the flow, MODE_IP and
// MODE_BA only the sum of a.
template <int FC, int KC, int G, bool NOBA>
__device__ __forceinline__ void assoc_row(const AssocProblem& P, const KConsts& kc, const ExactK& ek,
                                          const GridLevel& GL, const Maps& mp, int r, int g, int mode, float ell,
                                          int K, const float* xf, const float* xs, RowResult& out) {
#pragma unroll
  for (int c = 0; c < 7; ++c) out.acc[c] = 0.0;
  out.stored = 0;
  out.overflow = false;
  const uint32_t Ku = (uint32_t)K;
  if (!kc.use_geo) {
    // no distance gate: every target is a candidate, so the row takes the
    // exact full scan of the capped-row kernel
    out.stored = Ku + 1u;
    out.overflow = true;
    return;
  }
  const int F = Dim<FC>::get(kc.F);
  const int KS = Dim<KC>::get(kc.Ks);
  float4 x4 = ld_global(P.src_pos + r);
  float x0 = x4.x, x1 = x4.y, x2 = x4.z;
  if (mp.src_xf) {
    float t0, t1, t2;
    affine3(mp.As, mp.bs, x0, x1, x2, t0, t1, t2);
    x0 = t0; x1 = t1; x2 = t2;
  }
  // the adaptive-lengthscale association uses ell as is (AdaptiveCvoGPU.cu:281)
  const float l = kc.fixed_ell ? ell : range_ell(ell, x0, x1, x2);
  const float d2t = (float)(-2.0 * (double)l * (double)l * (double)kc.lg);
  const float qg = (float)(1.0 / (2.0 * (double)l * (double)l));
  float xg[2] = {0.0f, 0.0f};
  if (kc.use_gtype) {
    xg[0] = ld_global(P.src_geo + 2 * (size_t)r);
    xg[1] = ld_global(P.src_geo + 2 * (size_t)r + 1);
  }
  // query box in the target frame
  float q0, q1, q2;
  affine3(mp.Rq, mp.Tq, x0, x1, x2, q0, q1, q2);
  const float rq = sqrtf(d2t) * 1.001f + 1e-4f * (1.0f + fmaxf(fabsf(q0), fmaxf(fabsf(q1), fabsf(q2))));
  int lo0 = cell_floor((q0 - rq - GL.ox) * GL.inv_h, GL.nx), hi0 = cell_floor((q0 + rq - GL.ox) * GL.inv_h, GL.nx);
  int lo1 = cell_floor((q1 - rq - GL.oy) * GL.inv_h, GL.ny), hi1 = cell_floor((q1 + rq - GL.oy) * GL.inv_h, GL.ny);
  int lo2 = cell_floor((q2 - rq - GL.oz) * GL.inv_h, GL.nz), hi2 = cell_floor((q2 + rq - GL.oz) * GL.inv_h, GL.nz);
  int nruns = 0, span1 = 0;
  if (!(hi0 < 0 || lo0 >= GL.nx || hi1 < 0 || lo1 >= GL.ny || hi2 < 0 || lo2 >= GL.nz)) {
    lo0 = max(lo0, 0); hi0 = min(hi0, GL.nx - 1);
    lo1 = max(lo1, 0); hi1 = min(hi1, GL.ny - 1);
    lo2 = max(lo2, 0); hi2 = min(hi2, GL.nz - 1);
    span1 = hi1 - lo1 + 1;
    nruns = span1 * (hi2 - lo2 + 1);
  }
  uint32_t stored = 0;
  bool stop = false;
  // batched launches (bundle adjustment, multi-start): a run's cells beyond
  // the query radius hold no passing target, so the run is clipped to the
  // ball's chord over its (y, z) cell square, or skipped; the slack covers
  // the float rounding of the points' cell assignment, and rq itself exceeds
  // the distance gate's radius, so no passing target is lost and the passing
  // targets keep their enumeration order.  (Single-problem launches do not
  // clip: their fine cells leave little to cut, and the extra arithmetic on
  // the row's dependent chain costs more.)
  constexpr bool CLIP = !NOBA;
  const float rq2 = rq * rq, slack = 1e-3f * GL.h;
  (void)rq2;
  (void)slack;
  for (int rbase = 0; rbase < nruns && !stop; rbase += G) {
    const int myrun = rbase + g;
    uint32_t rs = 0, rl = 0;
    if (myrun < nruns) {
      const int iz = myrun / span1, iy = myrun - iz * span1;
      if constexpr (CLIP) {
        const int cy = lo1 + iy, cz = lo2 + iz;
        const float ys = GL.oy + (float)cy * GL.h, zs = GL.oz + (float)cz * GL.h;
        const float dy = fmaxf(0.0f, fmaxf(ys - q1, q1 - (ys + GL.h)) - slack);
        const float dz = fmaxf(0.0f, fmaxf(zs - q2, q2 - (zs + GL.h)) - slack);
        const float rem = rq2 - (dy * dy + dz * dz);
        if (rem >= 0.0f) {  // else the run lies beyond the radius
          const float ch = sqrtf(rem) + slack;
          const int lo = max(lo0, cell_floor((q0 - ch - GL.ox) * GL.inv_h, GL.nx));
          const int hi = min(hi0, cell_floor((q0 + ch - GL.ox) * GL.inv_h, GL.nx));
          if (lo <= hi) {
            const uint32_t c0 = ((uint32_t)cz * (uint32_t)GL.ny + (uint32_t)cy) * (uint32_t)GL.nx + (uint32_t)lo;
            rs = ld_global(GL.cell_start + c0);
            rl = ld_global(GL.cell_start + c0 + (uint32_t)(hi - lo) + 1u) - rs;
          }
        }
      } else {
        const uint32_t c0 =
            ((uint32_t)(lo2 + iz) * (uint32_t)GL.ny + (uint32_t)(lo1 + iy)) * (uint32_t)GL.nx + (uint32_t)lo0;
        rs = ld_global(GL.cell_start + c0);
        rl = ld_global(GL.cell_start + c0 + (uint32_t)(hi0 - lo0) + 1u) - rs;
      }
    }
    uint32_t incl = rl;
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
      uint32_t t = __shfl_up(incl, o, G);
      if (g >= o) incl += t;
    }
    const uint32_t total = __shfl(incl, G - 1, G);
    const uint32_t excl = incl - rl;
    for (uint32_t base = 0; base < total; base += G) {
      const uint32_t idx = base + (uint32_t)g;
      // the run holding candidate idx: the last run whose prefix is <= idx
      int kk = 0;
#pragma unroll
      for (int s = G / 2; s > 0; s >>= 1) {
        uint32_t e = __shfl(excl, kk + s, G);
        if (e <= idx) kk += s;
      }
      const uint32_t run_start = __shfl(rs, kk, G);
      const uint32_t run_excl = __shfl(excl, kk, G);
      bool pass = false;
      float a = 0.0f, y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
      uint32_t pidx = 0;
      float4 y4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (idx < total) {
        pidx = run_start + (idx - run_excl);
        y4 = ld_global(GL.pos + pidx);
        // with compile-time dimensions the candidate's features and class
        // vector are loaded together with its position, before the gates
        // that may not need them: one memory round trip instead of three
        float yfr[FC > 0 ? FC : 1], ysr[KC > 0 ? KC : 1];
        if (FC > 0) {
#pragma unroll
          for (int f = 0; f < (FC > 0 ? FC : 1); ++f) yfr[f] = ld_global(GL.feat + (size_t)pidx * F + f);
        }
        if (KC > 0) {
#pragma unroll
          for (int k = 0; k < (KC > 0 ? KC : 1); ++k) ysr[k] = ld_global(GL.sem + (size_t)pidx * KS + k);
        }
        affine3(mp.A, mp.b, y4.x, y4.y, y4.z, y0, y1, y2);
        pass = pair_kernel<FC, KC>(kc, ek, x0, x1, x2, y0, y1, y2, d2t, qg, l, xf,
                                   FC > 0 ? yfr : (FC != 0 ? GL.feat + (size_t)pidx * F : nullptr), xs,
                                   KC > 0 ? ysr : (KC != 0 ? GL.sem + (size_t)pidx * KS : nullptr), xg,
                                   kc.use_gtype ? GL.geo + 2 * (size_t)pidx : nullptr, a);
      }
      const uint64_t m = __ballot(pass);
      constexpr uint32_t kGroupMask = (uint32_t)((1ull << G) - 1ull);
      const uint32_t gm = (uint32_t)(m >> (G * ((threadIdx.x & 63) / G))) & kGroupMask;
      const uint32_t slot = stored + (uint32_t)__popc(gm & ((1u << g) - 1u));
      if (pass && (slot < Ku || kc.uncapped)) {
        if (slot < Ku) store_pair<NOBA>(P, mode, r, Ku, slot, pidx, y4, a);
        if (mode == MODE_ALIGN) {
          const float c0 = x1 * y2 - x2 * y1, c1 = x2 * y0 - x0 * y2, c2 = x0 * y1 - x1 * y0;
          out.acc[0] += (double)(c0 * a);
          out.acc[1] += (double)(c1 * a);
          out.acc[2] += (double)(c2 * a);
          out.acc[3] += (double)((y0 - x0) * a);
          out.acc[4] += (double)((y1 - x1) * a);
          out.acc[5] += (double)((y2 - x2) * a);
        }
        out.acc[6] += (double)(kc.ip_ad2 ? sq_dist(x0, x1, x2, y0, y1, y2) * a : a);
      }
      stored += (uint32_t)__popc(gm);
      if (stored > Ku && !kc.uncapped) {
        stop = true;
        break;
      }
    }
  }
  // uncapped: every passing pair is summed, the first K stored
  out.stored = kc.uncapped ? min(stored, Ku) : stored;
  out.overflow = stored > Ku && !kc.uncapped;
}

// ---------------------------------------------------------------------------
// K-nearest association (is_using_kdtree: se_kernel_kdtree, CvoGPU.cu:1104-1154,
// fill_in_A_mat_cukdtree :832-949).  A row's candidates are its K nearest
// targets, nearest first; each must then pass the gates, the distance gate at
// the fixed lengthscale (no range scaling, :863).  A target outside the
// distance gate's ball can never be stored, and a target inside it is among
// the K nearest exactly when fewer than K ball targets are nearer, so the
// selection runs over the ball alone.  Nearness is the key (float d2 bits,
// target id): ties by ascending id.  A row counts its ball targets; with more
// than K, a most-significant-digit radix select over the keys (8 bits a pass,
// one 256-bin LDS histogram per row) finds the K-th smallest key.  The last
// pass applies the gates to the selected targets and stores the passing ones.
// ---------------------------------------------------------------------------

template <int FC, int KC, int G>
__device__ __forceinline__ void knn_row(const AssocProblem& P, const KConsts& kc, const ExactK& ek,
                                        const GridLevel& GL, const Maps& mp, int r, int g, int mode, float ell, int K,
                                        const float* xf, const float* xs, uint32_t* hist, RowResult& out) {
#pragma unroll
  for (int c = 0; c < 7; ++c) out.acc[c] = 0.0;
  out.stored = 0;
  out.overflow = false;
  const uint32_t Ku = (uint32_t)max(K, 0);
  if (Ku == 0) return;
  const int F = Dim<FC>::get(kc.F);
  const int KS = Dim<KC>::get(kc.Ks);
  float4 x4 = ld_global(P.src_pos + r);
  float x0 = x4.x, x1 = x4.y, x2 = x4.z;
  if (mp.src_xf) {
    float t0, t1, t2;
    affine3(mp.As, mp.bs, x0, x1, x2, t0, t1, t2);
    x0 = t0; x1 = t1; x2 = t2;
  }
  const float l = ell;
  const float d2t = (float)(-2.0 * (double)l * (double)l * (double)kc.lg);
  const float qg = (float)(1.0 / (2.0 * (double)l * (double)l));
  float xg[2] = {0.0f, 0.0f};
  if (kc.use_gtype) {
    xg[0] = ld_global(P.src_geo + 2 * (size_t)r);
    xg[1] = ld_global(P.src_geo + 2 * (size_t)r + 1);
  }
  float q0, q1, q2;
  affine3(mp.Rq, mp.Tq, x0, x1, x2, q0, q1, q2);
  const float rq = sqrtf(d2t) * 1.001f + 1e-4f * (1.0f + fmaxf(fabsf(q0), fmaxf(fabsf(q1), fabsf(q2))));
  const CellBox B = cell_box(GL, q0, q1, q2, rq);
  // key of a target, ~0 outside the ball (no ball key reaches it: d2 bits of
  // a finite d2 < d2t stay below 0x7f800000)
  auto key_of = [&](uint32_t pidx, float4& y4, float& y0, float& y1, float& y2) -> uint64_t {
    y4 = ld_global(GL.pos + pidx);
    affine3(mp.A, mp.b, y4.x, y4.y, y4.z, y0, y1, y2);
    const float dx = y0 - x0, dy = y1 - x1, dz = y2 - x2;
    const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    if (!(d2 < d2t)) return ~0ull;
    return ((uint64_t)__float_as_uint(d2) << 32) | (uint64_t)__float_as_uint(y4.w);
  };
  // (1)-(2) the K nearest of the ball: thr / tsh bound their keys
  uint64_t thr;
  int tsh;
  select_smallest<G>(
      [&](auto&& fn) { for_each_in_box<G>(GL, B, g, fn); }, g, Ku, hist,
      [&](uint32_t pidx) {
        float4 y4;
        float y0, y1, y2;
        return key_of(pidx, y4, y0, y1, y2);
      },
      thr, tsh);
  // (3) gates on the selected targets, stored in slots by ballot
  uint32_t stored = 0;
  for_each_in_box<G>(GL, B, g, [&](bool valid, uint32_t pidx) {
    bool pass = false;
    float a = 0.0f, y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
    float4 y4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (valid) {
      const uint64_t key = key_of(pidx, y4, y0, y1, y2);
      if (selected(key, thr, tsh)) {
        // compile-time dimensions: the target's features and class vector
        // loaded into registers through global-address loads
        float yfr[FC > 0 ? FC : 1], ysr[KC > 0 ? KC : 1];
        if (FC > 0) {
#pragma unroll
          for (int f = 0; f < (FC > 0 ? FC : 1); ++f) yfr[f] = ld_global(GL.feat + (size_t)pidx * F + f);
        }
        if (KC > 0) {
#pragma unroll
          for (int k = 0; k < (KC > 0 ? KC : 1); ++k) ysr[k] = ld_global(GL.sem + (size_t)pidx * KS + k);
        }
        pass = pair_kernel<FC, KC>(kc, ek, x0, x1, x2, y0, y1, y2, d2t, qg, l, xf,
                                   FC > 0 ? yfr : (FC != 0 ? GL.feat + (size_t)pidx * F : nullptr), xs,
                                   KC > 0 ? ysr : (KC != 0 ? GL.sem + (size_t)pidx * KS : nullptr), xg,
                                   kc.use_gtype ? GL.geo + 2 * (size_t)pidx : nullptr, a, 0.001f);
      }
    }
    const uint32_t gm = group_bits<G>(__ballot(pass));
    const uint32_t slot = stored + (uint32_t)__popc(gm & ((1u << g) - 1u));
    if (pass && slot < Ku) {
      store_pair(P, mode, r, Ku, slot, pidx, y4, a);
      if (mode == MODE_ALIGN) {
        const float c0 = x1 * y2 - x2 * y1, c1 = x2 * y0 - x0 * y2, c2 = x0 * y1 - x1 * y0;
        out.acc[0] += (double)(c0 * a);
        out.acc[1] += (double)(c1 * a);
        out.acc[2] += (double)(c2 * a);
        out.acc[3] += (double)((y0 - x0) * a);
        out.acc[4] += (double)((y1 - x1) * a);
        out.acc[5] += (double)((y2 - x2) * a);
      }
      out.acc[6] += (double)a;
    }
    stored += (uint32_t)__popc(gm);
  });
  out.stored = min(stored, Ku);
}

// Block partials of the association pass: part_a[block] = flow6, ip, -;
// part_cnt[block] = stored pairs, max row; summed into the problem state by
// the capped-row kernel that follows.

// LDS of one association block, sized for the most rows a block of the
// variant can hold (ROWS: 64 with the four-lane path, else 32; the K-nearest
// variant's histograms leave no room for more).
template <int FC, int KC, int ROWS>
struct AssocLds {
  static constexpr int FL = FC > 0 ? FC : (FC == 0 ? 1 : 64);
  static constexpr int KL = KC > 0 ? KC : (KC == 0 ? 1 : 64);
  static constexpr int kRows = ROWS;
  double row[ROWS][8];
  uint32_t cnt[ROWS];
  float feat[ROWS * FL];  // the block's source rows' features
  float sem[ROWS * KL];   // and class vectors
};

// Logical block lb of a problem with G lanes per row (256 / G rows); KNN:
// the K-nearest rows (hist: 256 bins per row).
template <int FC, int KC, int G, bool KNN, int ROWS>
__device__ __forceinline__ void assoc_block(const AssocProblem& P, const KConsts& kc, int lb,
                                            AssocLds<FC, KC, ROWS>& L, uint32_t* hist) {
  constexpr int RB = rows_per_block(G);
  static_assert(RB <= ROWS, "the block's rows exceed its LDS");
  const ProblemState* st = P.st;
  const int tid = threadIdx.x;
  const int g = tid & (G - 1);
  const int grp = tid / G;
  const int r = lb * RB + grp;
  const int mode = kc.mode;
  const float ell = st->ell;
  const int K = st->K;
  const int iter = st->iter;
  Maps mp;
  load_maps(st, mp);
  const ExactK ek = exact_consts(kc);
  // a copy: the pair stores below cannot alias it, so its fields stay in
  // registers instead of being reloaded after every store
  const GridLevel GL = P.grid->lv[st->level];
  const int row0 = lb * RB;
  const int nrows = min(RB, P.n_src - row0);
  if (FC != 0) {
    const int F = Dim<FC>::get(kc.F);
    for (int t = tid; t < nrows * F; t += kBlock) L.feat[t] = ld_global(P.src_feat + (size_t)row0 * F + t);
  }
  if (KC != 0) {
    const int KS = Dim<KC>::get(kc.Ks);
    for (int t = tid; t < nrows * KS; t += kBlock) L.sem[t] = ld_global(P.src_sem + (size_t)row0 * KS + t);
  }
  __syncthreads();
  RowResult res;
#pragma unroll
  for (int c = 0; c < 7; ++c) res.acc[c] = 0.0;
  res.stored = 0;
  res.overflow = false;
  if (r < P.n_src) {
    if constexpr (KNN)
      knn_row<FC, KC, G>(P, kc, ek, GL, mp, r, g, mode, ell, K, L.feat + grp * Dim<FC>::get(kc.F),
                         L.sem + grp * Dim<KC>::get(kc.Ks), hist + grp * 256, res);
    else
      assoc_row<FC, KC, G, (ROWS == kMaxRowsPerBlock)>(P, kc, ek, GL, mp, r, g, mode, ell, K,
                                                       L.feat + grp * Dim<FC>::get(kc.F),
                                                       L.sem + grp * Dim<KC>::get(kc.Ks), res);
  }
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    double v = res.acc[c];
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    res.acc[c] = v;
  }
  if (g == 0) {
    uint32_t cnt = 0;
    if (r < P.n_src) {
      // a capped row is counted by the capped-row kernel, which rewrites it
      cnt = res.overflow ? 0u : res.stored;
      st_global(P.counts + r, cnt);
      if (res.overflow) {
        st_global(P.ovf_stamp + r, (uint32_t)iter + 1u);
        atomicAdd(&P.st->n_overflow, 1u);
      }
    }
#pragma unroll
    for (int c = 0; c < 7; ++c) L.row[grp][c] = res.overflow ? 0.0 : res.acc[c];
    L.cnt[grp] = cnt;
  }
  __syncthreads();
  if (tid < 7) {
    double s = 0.0;
    for (int q = 0; q < RB; ++q) s += L.row[q][tid];
    st_global(P.part_a + (size_t)lb * 8 + tid, s);
  } else if (tid == 7) {
    uint32_t nz = 0, mx = 0;
    for (int q = 0; q < RB; ++q) {
      nz += L.cnt[q];
      mx = max(mx, L.cnt[q]);
    }
    st_global(P.part_cnt + (size_t)lb * 2, nz);
    st_global(P.part_cnt + (size_t)lb * 2 + 1, mx);
  }
}

// Fixed-order tree over red[c][0..W): sums, except the max-row column 8.
// Every thread of the block takes part; red[c][0] holds the result.
template <int W>
__device__ __forceinline__ void tree10(double (*red)[W]) {
  const int tid = threadIdx.x;
  for (int w = W / 2; w > 0; w >>= 1) {
    __syncthreads();
    for (int t = tid; t < 10 * w; t += kBlock) {
      const int c = t / w, i = t - c * w;
      red[c][i] = (c == 8) ? fmax(red[c][i], red[c][i + w]) : red[c][i] + red[c][i + w];
    }
  }
  __syncthreads();
}

// Lanes per row of a problem's next pass (8 or 16).
__device__ __forceinline__ int state_group(const ProblemState* st) {
  return st->group == 4 ? 4 : (st->group == 8 ? 8 : 16);
}

// Active association blocks of a problem at a lane count.
__device__ __forceinline__ int active_blocks(const AssocProblem& P, int group) {
  const int rb = kBlock / group;
  return (P.n_src + rb - 1) / rb;
}

// One launch covers every problem; a problem owns P.n_blocks blocks (its rows
// at the largest lane count) and uses a prefix of them at its current lane
// count.  With xcd_remap the active blocks are renumbered so that each
// XCD holds one contiguous range of rows.
// Register budget: see assoc_kernel.
// FOUR: the four-lanes-per-row path is compiled in (single-problem kernels
// only: it raises the batched variant's register count past its budget; the
// host never selects four lanes for a batched launch).
template <int FC, int KC, bool KNN, bool FOUR>
__device__ __forceinline__ void assoc_pass(const AssocProblem& P, const KConsts& kc) {
  const ProblemState* st = P.st;
  if (st->done) return;
  const int group = state_group(st);
  const int n_act = active_blocks(P, group);
  int lb = (int)blockIdx.x - P.block_begin;
  if (lb >= n_act) return;  // uniform
  if (kc.xcd_remap) lb = xcd_block(lb, n_act);
  __shared__ AssocLds<FC, KC, FOUR ? kMaxRowsPerBlock : kKnnRowsPerBlock> L;
  if constexpr (KNN) {
    __shared__ uint32_t hist[kKnnRowsPerBlock * 256];
    switch (group) {
      case 8: assoc_block<FC, KC, 8, true>(P, kc, lb, L, hist); break;
      default: assoc_block<FC, KC, 16, true>(P, kc, lb, L, hist); break;
    }
  } else {
    switch (group) {
      // two lane counts keep the kernel within its register budget
      case 4:
        if constexpr (FOUR) assoc_block<FC, KC, 4, false>(P, kc, lb, L, nullptr);
        break;
      case 8: assoc_block<FC, KC, 8, false>(P, kc, lb, L, nullptr); break;
      default: assoc_block<FC, KC, 16, false>(P, kc, lb, L, nullptr); break;
    }
  }
}

// ONE: a single problem, its descriptor by value in the kernel arguments (read
// with scalar loads, no dependent load before the first state read).
// Occupancy: a single problem runs at five waves per SIMD (96 VGPRs, a
// 16-byte spill; the 60k street registration 128.3 -> 125.0 ms, six waves
// spill 96 bytes and take 157 ms); several problems at four (128 VGPRs: at
// five the batched variant spills 64 bytes and the 96-start run slows 8%).
template <int FC, int KC, bool ONE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(ONE ? 5 : 4))) void assoc_kernel(
    const AssocProblem* __restrict__ probs, int n_probs, KConsts kc, AssocProblem one) {
  if constexpr (ONE) {
    assoc_pass<FC, KC, false, true>(one, kc);
  } else {
    // a register copy of the problem: the pair stores cannot alias it
    const AssocProblem P = probs[find_problem(probs, n_probs, (int)blockIdx.x)];
    assoc_pass<FC, KC, false, false>(P, kc);
  }
}

// The K-nearest association pass (is_using_kdtree), same launch shape and
// outputs; rows never exceed their cap, so no capped-row work follows.
template <int FC, int KC, bool ONE>
__global__ __launch_bounds__(kBlock) void knn_assoc_kernel(const AssocProblem* __restrict__ probs, int n_probs,
                                                           KConsts kc, AssocProblem one) {
  if constexpr (ONE) {
    assoc_pass<FC, KC, true, false>(one, kc);
  } else {
    const AssocProblem P = probs[find_problem(probs, n_probs, (int)blockIdx.x)];
    assoc_pass<FC, KC, true, false>(P, kc);
  }
}

// ---------------------------------------------------------------------------
// Totals of an association pass, and the rows whose passing pairs exceed the
// cap, which keep their first K passing targets by index (the reference's
// scan over all targets in ascending index, CvoGPU.cu:994-996).  Each block sums a slice of the
// association block partials, adds its capped rows and publishes
// part_ovf[block] = flow6, ip, stored, max row; the last block writes the
// problem's totals (raw flow sums, ip, stored pairs, longest row).
// ---------------------------------------------------------------------------

// The pass totals of a problem by one block: thread t takes association
// blocks t, t + 256, ... (four at a time, for loads in flight), then a fixed
// tree: wave shuffles, then the four waves in order.
__device__ __forceinline__ void pass_totals_one_block(const AssocProblem& P, ProblemState* st) {
  const int tid = threadIdx.x;
  const int n_act = active_blocks(P, state_group(st));
  double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double nz = 0.0, mx = 0.0;
#pragma unroll 4
  for (int b = tid; b < n_act; b += kBlock) {
    const double2* pa = reinterpret_cast<const double2*>(P.part_a + (size_t)b * 8);
    const double2 q0 = pa[0], q1 = pa[1], q2 = pa[2], q3 = pa[3];
    const uint2 cn = reinterpret_cast<const uint2*>(P.part_cnt)[b];
    acc[0] += q0.x; acc[1] += q0.y; acc[2] += q1.x; acc[3] += q1.y;
    acc[4] += q2.x; acc[5] += q2.y; acc[6] += q3.x;
    nz += (double)cn.x;
    mx = fmax(mx, (double)cn.y);
  }
  __shared__ double s_w[9][kBlock / 64];
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    double v = c < 7 ? acc[c] : (c == 7 ? nz : mx);
    for (int o = 32; o > 0; o >>= 1) {
      const double u = __shfl_xor(v, o, 64);
      v = c == 8 ? fmax(v, u) : v + u;
    }
    if (lane == 0) s_w[c][wave] = v;
  }
  __syncthreads();
  if (tid == 0) {
    double t[9];
    for (int c = 0; c < 9; ++c) {
      t[c] = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) t[c] = c == 8 ? fmax(t[c], s_w[c][w]) : t[c] + s_w[c][w];
    }
    for (int c = 0; c < 6; ++c) st->flow[c] = t[c];
    st->ip = t[6];
    st->nnz = (uint64_t)t[7];
    st->max_row = (uint32_t)t[8];
    st->level_assoc = st->level;
  }
}

// A row whose passing pairs exceed the cap keeps the first K passing targets
// by ascending target index (CvoGPU.cu:994-996).  Every passing target lies
// in the distance gate's ball, so one wave finds them there: it counts the
// ball's passing targets and their index range, selects the K smallest
// indices by a most-significant-digit radix select (8 bits a pass, a 256-bin
// LDS histogram), then stores exactly those.  acc: the lane's share of the
// row's flow and ip; returns the stored count.
template <int FC, int KC>
__device__ uint32_t capped_row_ball(const AssocProblem& P, const KConsts& kc, const ExactK& ek,
                                    const GridLevel& GL, const Maps& mp, int r, float ell, uint32_t K,
                                    uint32_t* hist, double acc[7]) {
  const int g = threadIdx.x & 63;
  const int mode = kc.mode;
  const int F = Dim<FC>::get(kc.F);
  const int KS = Dim<KC>::get(kc.Ks);
  float4 x4 = ld_global(P.src_pos + r);
  float x0 = x4.x, x1 = x4.y, x2 = x4.z;
  if (mp.src_xf) {
    float t0, t1, t2;
    affine3(mp.As, mp.bs, x0, x1, x2, t0, t1, t2);
    x0 = t0; x1 = t1; x2 = t2;
  }
  const float l = kc.fixed_ell ? ell : range_ell(ell, x0, x1, x2);
  const float d2t = (float)(-2.0 * (double)l * (double)l * (double)kc.lg);
  const float qg = (float)(1.0 / (2.0 * (double)l * (double)l));
  // with compile-time dimensions the row's and each candidate's features and
  // class vectors are read into registers through global-address loads
  float xfr[FC > 0 ? FC : 1], xsr[KC > 0 ? KC : 1];
  if (FC > 0 && P.src_feat) {
#pragma unroll
    for (int f = 0; f < (FC > 0 ? FC : 1); ++f) xfr[f] = ld_global(P.src_feat + (size_t)r * F + f);
  }
  if (KC > 0 && P.src_sem) {
#pragma unroll
    for (int k = 0; k < (KC > 0 ? KC : 1); ++k) xsr[k] = ld_global(P.src_sem + (size_t)r * KS + k);
  }
  const float* xf = P.src_feat ? (FC > 0 ? xfr : P.src_feat + (size_t)r * kc.F) : nullptr;
  const float* xs = P.src_sem ? (KC > 0 ? xsr : P.src_sem + (size_t)r * kc.Ks) : nullptr;
  float xg[2] = {0.0f, 0.0f};
  if (kc.use_gtype) {
    xg[0] = ld_global(P.src_geo + 2 * (size_t)r);
    xg[1] = ld_global(P.src_geo + 2 * (size_t)r + 1);
  }
  float q0, q1, q2;
  affine3(mp.Rq, mp.Tq, x0, x1, x2, q0, q1, q2);
  const float rq = sqrtf(d2t) * 1.001f + 1e-4f * (1.0f + fmaxf(fabsf(q0), fmaxf(fabsf(q1), fabsf(q2))));
  const CellBox B = cell_box(GL, q0, q1, q2, rq);
  // a passing target's key is its index; ~0 for the others
  auto eval = [&](uint32_t pidx, float4& y4, float& y0, float& y1, float& y2, float& a) -> uint64_t {
    y4 = ld_global(GL.pos + pidx);
    float yfr[FC > 0 ? FC : 1], ysr[KC > 0 ? KC : 1];
    if (FC > 0) {
#pragma unroll
      for (int f = 0; f < (FC > 0 ? FC : 1); ++f) yfr[f] = ld_global(GL.feat + (size_t)pidx * F + f);
    }
    if (KC > 0) {
#pragma unroll
      for (int k = 0; k < (KC > 0 ? KC : 1); ++k) ysr[k] = ld_global(GL.sem + (size_t)pidx * KS + k);
    }
    affine3(mp.A, mp.b, y4.x, y4.y, y4.z, y0, y1, y2);
    const bool pass = pair_kernel<FC, KC>(kc, ek, x0, x1, x2, y0, y1, y2, d2t, qg, l, xf,
                                          FC > 0 ? yfr : (FC != 0 ? GL.feat + (size_t)pidx * F : nullptr), xs,
                                          KC > 0 ? ysr : (KC != 0 ? GL.sem + (size_t)pidx * KS : nullptr), xg,
                                          kc.use_gtype ? GL.geo + 2 * (size_t)pidx : nullptr, a);
    return pass ? (uint64_t)__float_as_uint(y4.w) : ~0ull;
  };
  uint64_t thr;
  int tsh;
  select_smallest<64>(
      [&](auto&& fn) { for_each_in_box<64>(GL, B, g, fn); }, g, K, hist,
      [&](uint32_t pidx) {
        float4 y4;
        float y0, y1, y2, a;
        return eval(pidx, y4, y0, y1, y2, a);
      },
      thr, tsh);
  uint32_t stored = 0;
  for_each_in_box<64>(GL, B, g, [&](bool valid, uint32_t pidx) {
    bool take = false;
    float a = 0.0f, y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
    float4 y4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (valid) {
      const uint64_t key = eval(pidx, y4, y0, y1, y2, a);
      take = selected(key, thr, tsh);
    }
    const uint64_t m = __ballot(take);
    const uint32_t slot = stored + (uint32_t)__popcll(m & ((g == 0) ? 0ull : (~0ull >> (64 - g))));
    if (take && slot < K) {
      store_pair(P, mode, r, K, slot, pidx, y4, a);
      if (mode == MODE_ALIGN) {
        const float c0 = x1 * y2 - x2 * y1, c1 = x2 * y0 - x0 * y2, c2 = x0 * y1 - x1 * y0;
        acc[0] += (double)(c0 * a);
        acc[1] += (double)(c1 * a);
        acc[2] += (double)(c2 * a);
        acc[3] += (double)((y0 - x0) * a);
        acc[4] += (double)((y1 - x1) * a);
        acc[5] += (double)((y2 - x2) * a);
      }
      acc[6] += (double)(kc.ip_ad2 ? sq_dist(x0, x1, x2, y0, y1, y2) * a : a);
    }
    stored += (uint32_t)__popcll(m);
  });
  return min(stored, K);
}

template <int FC, int KC>
__device__ __forceinline__ void overflow_body(const AssocProblem& P, const KConsts& kc) {
  ProblemState* st = P.st;
  if (st->done) return;
  const int tid = threadIdx.x;
  const int mode = kc.mode;
  if (st->n_overflow == 0) {
    // no capped row: block 0 alone sums every association partial (no
    // cross-block publication), the other blocks have nothing to do
    if (blockIdx.x == 0) pass_totals_one_block(P, st);
    return;  // uniform per block
  }
  // (a) this block's slice of the association block partials, in fixed order
  __shared__ double s_tot[10][128];
  {
    const int n_act = active_blocks(P, state_group(st));
    const int per = (n_act + kOvfBlocks - 1) / kOvfBlocks;
    const int b0 = blockIdx.x * per, b1 = min(n_act, b0 + per);
    for (int t = tid; t < 10 * 128; t += kBlock) {
      const int c = t / 128, i = t - c * 128;
      double v = 0.0;
      for (int b = b0 + i; b < b1; b += 128) {
        if (c < 7) v += ld_global(P.part_a + (size_t)b * 8 + c);
        else if (c == 7) v += (double)ld_global(P.part_cnt + (size_t)b * 2);
        else if (c == 8) v = fmax(v, (double)ld_global(P.part_cnt + (size_t)b * 2 + 1));
      }
      s_tot[c][i] = v;
    }
    tree10<128>(s_tot);
  }
  double blk_acc[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) blk_acc[c] = s_tot[c][0];
  uint32_t blk_nnz = (uint32_t)s_tot[7][0], blk_max = (uint32_t)s_tot[8][0];
  __syncthreads();
  // (b) rows whose passing pairs exceed the cap (uniform branch)
  if (st->n_overflow != 0) {
  const float ell = st->ell;
  const uint32_t K = (uint32_t)st->K;
  const uint32_t stamp = (uint32_t)st->iter + 1u;
  Maps mp;
  load_maps(st, mp);
  const ExactK ek = exact_consts(kc);
  const GridLevel G = P.grid->lv[st->level];  // a register copy, as in the association
  const int F = Dim<FC>::get(kc.F);
  const int KS = Dim<KC>::get(kc.Ks);
  const int chunk = (P.n_src + kOvfBlocks - 1) / kOvfBlocks;
  const int r0 = blockIdx.x * chunk, r1 = min(P.n_src, r0 + chunk);
  __shared__ uint32_t s_wave[kBlock / 64];
  __shared__ double s_red[kBlock];
  if (kc.use_geo) {
    // the chunk's capped rows, one wave per row in turn (wave w takes the
    // w-th, (w+4)-th, ... capped row of the chunk)
    __shared__ uint32_t s_hist[kBlock / 64][256];
    __shared__ double s_wacc[kBlock / 64][9];
    const int lane = tid & 63, wave = tid >> 6;
    double wacc[7] = {0, 0, 0, 0, 0, 0, 0};
    uint32_t wnnz = 0, wmax = 0, seen = 0;
    for (int base = r0; base < r1; base += 64) {
      const int rr = base + lane;
      uint64_t m = __ballot(rr < r1 && ld_global(P.ovf_stamp + rr) == stamp);
      while (m) {
        const int b = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        if ((int)(seen % (kBlock / 64)) == wave) {
          const int r = base + b;
          double racc[7] = {0, 0, 0, 0, 0, 0, 0};
          const uint32_t kept = capped_row_ball<FC, KC>(P, kc, ek, G, mp, r, ell, K, s_hist[wave], racc);
          if (lane == 0) st_global(P.counts + r, kept);
#pragma unroll
          for (int c = 0; c < 7; ++c) wacc[c] += racc[c];
          wnnz += kept;
          wmax = max(wmax, kept);
        }
        seen++;
      }
    }
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      double v = wacc[c];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) s_wacc[wave][c] = v;
    }
    if (lane == 0) {
      s_wacc[wave][7] = (double)wnnz;
      s_wacc[wave][8] = (double)wmax;
    }
    __syncthreads();
    // every thread folds the waves in the same order (threads 0-7 publish)
    for (int w = 0; w < kBlock / 64; ++w) {
#pragma unroll
      for (int c = 0; c < 7; ++c) blk_acc[c] += s_wacc[w][c];
      blk_nnz += (uint32_t)s_wacc[w][7];
      blk_max = max(blk_max, (uint32_t)s_wacc[w][8]);
    }
  } else
  for (int r = r0; r < r1; ++r) {
    if (ld_global(P.ovf_stamp + r) != stamp) continue;  // uniform
    float4 x4 = ld_global(P.src_pos + r);
    float x0 = x4.x, x1 = x4.y, x2 = x4.z;
    if (mp.src_xf) {
      float t0, t1, t2;
      affine3(mp.As, mp.bs, x0, x1, x2, t0, t1, t2);
      x0 = t0; x1 = t1; x2 = t2;
    }
    const float l = kc.fixed_ell ? ell : range_ell(ell, x0, x1, x2);
    const float d2t = kc.use_geo ? (float)(-2.0 * (double)l * (double)l * (double)kc.lg) : 1.0f;
    const float qg = (float)(1.0 / (2.0 * (double)l * (double)l));
    const float* xf = P.src_feat ? P.src_feat + (size_t)r * kc.F : nullptr;
    const float* xs = P.src_sem ? P.src_sem + (size_t)r * kc.Ks : nullptr;
    float xg[2] = {0.0f, 0.0f};
    if (kc.use_gtype) {
      xg[0] = ld_global(P.src_geo + 2 * (size_t)r);
      xg[1] = ld_global(P.src_geo + 2 * (size_t)r + 1);
    }
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    uint32_t stored = 0;
    for (int base = 0; base < P.n_tgt && stored < K; base += kBlock) {
      const int j = base + tid;
      bool pass = false;
      float a = 0.0f, y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
      if (j < P.n_tgt) {
        affine3(mp.A, mp.b, ld_global(P.tgt_xyz + 3 * (size_t)j), ld_global(P.tgt_xyz + 3 * (size_t)j + 1), ld_global(P.tgt_xyz + 3 * (size_t)j + 2),
                y0, y1, y2);
        pass = pair_kernel<FC, KC>(kc, ek, x0, x1, x2, y0, y1, y2, d2t, qg, l, xf,
                                   FC != 0 ? P.tgt_feat + (size_t)j * F : nullptr, xs,
                                   KC != 0 ? P.tgt_sem + (size_t)j * KS : nullptr, xg,
                                   kc.use_gtype ? P.tgt_geo + 2 * (size_t)j : nullptr, a);
      }
      // block-wide exclusive prefix of pass flags, in index order
      const uint64_t m = __ballot(pass);
      const int lane = tid & 63, wave = tid >> 6;
      if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
      __syncthreads();
      uint32_t before = 0, tot = 0;
      for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wave) before += s_wave[w];
        tot += s_wave[w];
      }
      before += (uint32_t)__popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
      const uint32_t slot = stored + before;
      if (pass && slot < K) {
        const float4 y4 =
            make_float4(ld_global(P.tgt_xyz + 3 * (size_t)j), ld_global(P.tgt_xyz + 3 * (size_t)j + 1), ld_global(P.tgt_xyz + 3 * (size_t)j + 2), 0.0f);
        store_pair(P, mode, r, K, slot, G.inv_perm[j], y4, a);
        if (mode == MODE_ALIGN) {
          const float c0 = x1 * y2 - x2 * y1, c1 = x2 * y0 - x0 * y2, c2 = x0 * y1 - x1 * y0;
          acc[0] += (double)(c0 * a);
          acc[1] += (double)(c1 * a);
          acc[2] += (double)(c2 * a);
          acc[3] += (double)((y0 - x0) * a);
          acc[4] += (double)((y1 - x1) * a);
          acc[5] += (double)((y2 - x2) * a);
        }
        acc[6] += (double)(kc.ip_ad2 ? sq_dist(x0, x1, x2, y0, y1, y2) * a : a);
      }
      stored += tot;
      __syncthreads();
    }
    const uint32_t kept = min(stored, K);
    if (tid == 0) st_global(P.counts + r, kept);
    blk_nnz += kept;
    blk_max = max(blk_max, kept);
    for (int c = 0; c < 7; ++c) {
      s_red[tid] = acc[c];
      __syncthreads();
      for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) s_red[tid] += s_red[tid + w];
        __syncthreads();
      }
      blk_acc[c] += s_red[0];
      __syncthreads();
    }
  }
  }
  // (c) the block's partial (its slice plus its capped rows); the last block
  // sums all blocks' partials in fixed order into the problem state
  double* out = P.part_ovf + (size_t)blockIdx.x * 10;
  __shared__ int s_last;
  if (tid < 64) {
    if (tid < 7) st_dev(&out[tid], blk_acc[tid]);
    if (tid == 7) {
      st_dev(&out[7], (double)blk_nnz);
      st_dev(&out[8], (double)blk_max);
      st_dev(&out[9], 0.0);
    }
    wait_stores();
    if (tid == 0) s_last = (count_in(&st->blocks_done_ovf) == (uint32_t)(gridDim.x - 1)) ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;  // uniform
  for (int t = tid; t < 10 * 128; t += kBlock) {
    const int c = t / 128, i = t - c * 128;
    s_tot[c][i] = (i < (int)gridDim.x) ? ld_dev(&P.part_ovf[(size_t)i * 10 + c]) : 0.0;
  }
  tree10<128>(s_tot);
  if (tid == 0) {
    for (int c = 0; c < 6; ++c) st->flow[c] = s_tot[c][0];
    st->ip = s_tot[6][0];
    st->nnz = (uint64_t)s_tot[7][0];
    st->max_row = (uint32_t)s_tot[8][0];
    st->level_assoc = st->level;
    st_dev(&st->blocks_done_ovf, 0u);
  }
}

// A single problem's descriptor comes by value in the kernel arguments (no
// dependent load before the first state read); several come from memory.
template <int FC, int KC, bool ONE>
__global__ __launch_bounds__(kBlock) void overflow_kernel(const AssocProblem* __restrict__ probs, KConsts kc,
                                                          AssocProblem one) {
  if constexpr (ONE) overflow_body<FC, KC>(one, kc);
  else {
    const AssocProblem P = probs[blockIdx.y];  // a register copy: stores cannot alias it
    overflow_body<FC, KC>(P, kc);
  }
}

// ---------------------------------------------------------------------------
// Device controller pieces (FP64)
// ---------------------------------------------------------------------------

// Real roots of a s^3 + b s^2 + c s + d by the trigonometric / Cardano closed
// form with two Newton polish steps; complex pairs report their imaginary part
// for the |imag| < 1e-5 rule of compute_step_size (CvoGPU.cu:513-519).
__device__ void cubic_solve(double a, double b, double c, double d, double re[3], double im[3]) {
  const double nan = __longlong_as_double(0x7ff8000000000000ll);
  re[0] = re[1] = re[2] = nan;
  im[0] = im[1] = im[2] = 0.0;
  if (!(a != 0.0)) return;
  const double p2 = b / a, p1 = c / a, p0 = d / a;
  if (!isfinite(p2) || !isfinite(p1) || !isfinite(p0)) return;
  const double sh = p2 / 3.0;
  const double p = p1 - p2 * sh;
  const double q = 2.0 * sh * sh * sh - sh * p1 + p0;
  const double disc = 0.25 * q * q + p * p * p / 27.0;
  auto polish = [&](double s) {
    for (int it = 0; it < 2; ++it) {
      double f = ((s + p2) * s + p1) * s + p0;
      double df = (3.0 * s + 2.0 * p2) * s + p1;
      if (df == 0.0) break;
      double ns = s - f / df;
      if (!isfinite(ns)) break;
      s = ns;
    }
    return s;
  };
  if (disc > 0.0) {
    const double sq = sqrt(disc);
    const double u = cbrt(-0.5 * q + (q > 0 ? -sq : sq));
    const double t = (u != 0.0) ? u - p / (3.0 * u) : 0.0;
    re[0] = polish(t - sh);
    const double ci = sqrt(fmax(0.75 * t * t + p, 0.0));
    re[1] = re[2] = -0.5 * t - sh;
    im[1] = ci;
    im[2] = -ci;
  } else if (p == 0.0) {
    re[0] = re[1] = re[2] = -sh;
  } else {
    const double m = 2.0 * sqrt(-p / 3.0);
    double arg = 3.0 * q / (p * m);
    arg = fmin(1.0, fmax(-1.0, arg));
    const double th = acos(arg) / 3.0;
    for (int k = 0; k < 3; ++k) re[k] = polish(m * cos(th - 2.0943951023931954923 * k) - sh);
  }
}

__device__ double select_step_dev(const KConsts& kc, const double bcde[4]) {
  double re[3], im[3];
  cubic_solve(4.0 * bcde[3], 3.0 * bcde[2], 2.0 * bcde[1], bcde[0], re, im);
  double temp = 1.7976931348623157e308;
  for (int i = 0; i < 3; ++i)
    if (re[i] > 0 && re[i] < temp && fabs(im[i]) < 1e-5) temp = re[i];
  if (temp > (double)kc.max_step) return (double)kc.max_step;
  if (temp < (double)kc.min_step) return (double)kc.min_step;
  return temp;
}

// A_sparsity_indicator_ell_update (CvoGPU.cu:547-665): two FIFO windows with
// float sums; ring buffers of modulus W+1.
__device__ bool indicator_update(ProblemState* st, float ind, int W, float thr) {
  const int cap = W + 1;
  bool dec = false;
  if (st->qs_size < W) {
    st->qs[(st->qs_head + st->qs_size) % cap] = ind;
    st->qs_size++;
    st->ss += ind;
  }
  if (st->qs_size >= W && st->qe_size < W) {
    st->qe[(st->qe_head + st->qe_size) % cap] = ind;
    st->qe_size++;
    st->se += ind;
  }
  if (st->qs_size >= W && st->qe_size >= W) {
    const float ratio = st->se / st->ss;
    if (ratio > 1.0f - thr && ratio < 1.0f + thr) {
      dec = true;
      st->qs_head = st->qs_size = st->qe_head = st->qe_size = 0;
      st->ss = 0.0f;
      st->se = 0.0f;
    } else {
      const float f = st->qe[st->qe_head];
      st->se -= f;
      st->ss += f;
      st->qs[(st->qs_head + st->qs_size) % cap] = f;
      st->qs_size++;
      st->qe_head = (st->qe_head + 1) % cap;
      st->qe_size--;
      st->ss -= st->qs[st->qs_head];
      st->qs_head = (st->qs_head + 1) % cap;
      st->qs_size--;
      st->qe[(st->qe_head + st->qe_size) % cap] = ind;
      st->qe_size++;
      st->se += ind;
    }
  }
  return dec;
}

// Float maps of the next iteration from the FP64 pose: target map [R^T, -R^T T]
// (update_tf, CvoGPU.cu:98-130) and the query map (R, T).
__device__ void set_maps(ProblemState* st) {
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      st->A[i * 3 + j] = (float)st->R[j * 3 + i];
      st->Rq[i * 3 + j] = (float)st->R[i * 3 + j];
    }
    st->b[i] = (float)(-(st->R[0 * 3 + i] * st->T[0] + st->R[1 * 3 + i] * st->T[1] + st->R[2 * 3 + i] * st->T[2]));
    st->Tq[i] = (float)st->T[i];
  }
}

__device__ int choose_level_tab(const float* h, int n_levels, float r_need) {
  int L = 0;
  for (int l = 0; l < n_levels; ++l)
    if (h[l] >= r_need) L = l;
  return L;
}

__device__ float needed_radius(const KConsts& kc, const ProblemState* st, float ell) {
  float tq = fmaxf(fabsf(st->Tq[0]), fmaxf(fabsf(st->Tq[1]), fabsf(st->Tq[2])));
  return kc.rq_coef * ell * (1.0f + kc.max_src_norm / 500.0f) * 1.0001f + 1e-4f * (1.0f + kc.max_src_norm + tq) + 1e-6f;
}

// Lanes per row and grid level of the next association pass at lengthscale
// ell (the level table given as cell sizes or as the grid).
__device__ void set_level(ProblemState* st, const KConsts& kc, float ell, const float* level_h, int n_levels) {
  const bool big = policy_big(kc.pol, ell);
  st->group = big ? kc.pol.group_big : kc.pol.group_small;
  st->level = choose_level_tab(level_h, n_levels,
                               needed_radius(kc, st, ell) * (big ? kc.pol.scale_big : kc.pol.scale_small));
}

__device__ void set_level(ProblemState* st, const KConsts& kc, float ell, const GridDev* grid) {
  const bool big = policy_big(kc.pol, ell);
  st->group = big ? kc.pol.group_big : kc.pol.group_small;
  st->level = choose_level(grid, needed_radius(kc, st, ell) * (big ? kc.pol.scale_big : kc.pol.scale_small));
}

__device__ double apply_step(ProblemState* st, const KConsts& kc, const double xi[6], const float xf[6],
                             double& step);

// One controller step of align_impl (CvoGPU.cu:1342-1433), FP64 pose state.
__device__ __forceinline__ void controller(ProblemState* st, const KConsts& kc, const AssocProblem& P, const float* level_h,
                           int n_levels) {
  cvo_iter_record rec;
  const int k = st->iter;
  double fl[6];
  double n2 = 0.0;
  for (int c = 0; c < 6; ++c) {
    fl[c] = st->flow[c];
    n2 += fl[c] * fl[c];
  }
  const double nrm = sqrt(n2);
  double xi[6];
  for (int c = 0; c < 6; ++c) xi[c] = n2 > 0.0 ? fl[c] / nrm : fl[c];
  float xf[6];
  for (int c = 0; c < 6; ++c) xf[c] = (float)xi[c];
  const double nw = sqrt((double)xf[0] * xf[0] + (double)xf[1] * xf[1] + (double)xf[2] * xf[2]);
  const double nv = sqrt((double)xf[3] * xf[3] + (double)xf[4] * xf[4] + (double)xf[5] * xf[5]);
  rec.ell = st->ell;
  rec.nnz = (int32_t)st->nnz;
  rec.max_row = (int32_t)st->max_row;
  rec.num_neighbors = st->K;
  rec.flow_norm = (float)nrm;
  rec.step = 0.0f;
  rec.dist = 0.0f;
  rec.ip = (float)((double)st->nnz / sqrt((double)kc.n_src_f * (double)kc.n_tgt_f));
  if (nw < (double)kc.eps && nv < (double)kc.eps) {
    st->ret = (nw < 1e-8 && nv < 1e-8) ? -1 : 0;
    st->done = 1;
    st->iters_out = k;
    if (P.trace && k < P.trace_cap) P.trace[k] = rec;
    return;
  }
  double step = select_step_dev(kc, st->bcde);
  const double dist = apply_step(st, kc, xi, xf, step);
  st->step = step;
  st->dist = dist;
  rec.step = (float)step;
  rec.dist = (float)dist;
  const bool dec = indicator_update(st, rec.ip, kc.window, kc.ind_thr);
  if (P.trace && k < P.trace_cap) P.trace[k] = rec;
  set_maps(st);
  if (dist < (double)kc.eps_2) {
    st->done = 1;
    st->iters_out = k;
    return;
  }
  if (k > kc.ell_decay_start && dec) {
    float e = st->ell * kc.ell_decay_rate;
    if (e < kc.ell_min) e = kc.ell_min;
    st->ell = e;
  }
  if (kc.knn > 0) {
    // K nearest: the mean stored row rounded up times is_using_kdtree, at
    // most KDTREE_K_SIZE (CvoGPU.cu:1416-1417); zero stays zero
    const int mean = (int)ceil((double)st->nnz / (double)P.n_src);
    st->K = min(kKnnMaxK, mean * kc.knn);
  } else {
    st->K = min(kc.nn_max, (int)((double)st->max_row * 1.2));
    if (st->K < 1) st->K = 1;
  }
  st->iter = k + 1;
  if (st->iter >= kc.max_iter) {
    st->done = 1;
    st->iters_out = kc.max_iter;
  }
  set_level(st, kc, st->ell, level_h, n_levels);
  st->n_overflow = 0;
}

// Pose update by exp(step xi) (CvoGPU.cu:1362-1379): FP64 by default, or the
// reference's float arithmetic (float step, FP32 Exp_SEK3, pose products in
// double rounded to float).  Returns the step's SE(3) distance; step is
// rounded to float in the reference mode.
__device__ double apply_step(ProblemState* st, const KConsts& kc, const double xi[6], const float xf[6],
                             double& step) {
  double dR[9], dT[3];
  double dist;
  if (kc.mirror_pose) {
    // the reference's float pose arithmetic: float step, FP32 Exp_SEK3, the
    // pose products in double rounded to float (CvoGPU.cu:1289-1341)
    step = (double)(float)step;
    float e[12];
    exp_sek3_f32(xf, (float)step, e);
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) dR[r * 3 + c] = (double)e[r * 4 + c];
      dT[r] = (double)e[r * 4 + 3];
    }
    double Rn[9], Tn[3];
    for (int r = 0; r < 3; ++r) {
      Tn[r] = (double)(float)(st->R[r * 3 + 0] * dT[0] + st->R[r * 3 + 1] * dT[1] + st->R[r * 3 + 2] * dT[2] + st->T[r]);
      for (int c = 0; c < 3; ++c)
        Rn[r * 3 + c] = (double)(float)(st->R[r * 3 + 0] * dR[0 * 3 + c] + st->R[r * 3 + 1] * dR[1 * 3 + c] +
                                        st->R[r * 3 + 2] * dR[2 * 3 + c]);
    }
    for (int i = 0; i < 9; ++i) st->R[i] = Rn[i];
    for (int i = 0; i < 3; ++i) st->T[i] = Tn[i];
    dist = se3_log_norm(dR, dT);
  } else {
    se3_exp(xi, step, dR, dT);
    double Rn[9], Tn[3];
    for (int r = 0; r < 3; ++r) {
      Tn[r] = st->T[r] + st->R[r * 3 + 0] * dT[0] + st->R[r * 3 + 1] * dT[1] + st->R[r * 3 + 2] * dT[2];
      for (int c = 0; c < 3; ++c)
        Rn[r * 3 + c] = st->R[r * 3 + 0] * dR[0 * 3 + c] + st->R[r * 3 + 1] * dR[1 * 3 + c] + st->R[r * 3 + 2] * dR[2 * 3 + c];
    }
    for (int i = 0; i < 9; ++i) st->R[i] = Rn[i];
    for (int i = 0; i < 3; ++i) st->T[i] = Tn[i];
    // ||log(exp(step xi))|| = step ||xi|| exactly while the rotation angle
    // step*|w| stays below pi; the general log otherwise
    const double wn = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
    const double xn = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2] + xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5]);
    dist = (step * wn < 3.0) ? step * xn : se3_log_norm(dR, dT);
  }
  return dist;
}

// Step length of the adaptive-lengthscale loop (AdaptiveCvoGPU.cu:1006-1075):
// the cubic's coefficients rounded to float, its smallest positive real root,
// min_step when there is none, at most 0.8.
__device__ double select_step_adaptive(const KConsts& kc, const double bcde[4]) {
  double re[3], im[3];
  cubic_solve((double)(4.0f * (float)bcde[3]), (double)(3.0f * (float)bcde[2]), (double)(2.0f * (float)bcde[1]),
              (double)(float)bcde[0], re, im);
  double temp = 1.7976931348623157e308;
  for (int i = 0; i < 3; ++i)
    if (re[i] > 0 && re[i] < temp && im[i] == 0.0) temp = re[i];
  double step = temp == 1.7976931348623157e308 ? (double)kc.min_step : (double)(float)temp;
  return step > 0.8 ? 0.8 : step;
}

// One controller step of AdaptiveCvoGPU::align (AdaptiveCvoGPU.cu:1117-1205):
// unnormalised flow (omega / c, v / d), stop when both parts are below eps,
// exp(step xi), stop when the step's distance is below eps_2, then the
// lengthscale moves against its gradient
//   dl = (2/l^3 sum_A -a |y-x|^2 + 1/l^3 sum_Axx a |x'-x|^2 + 1/l^3 sum_Ayy a |y'-y|^2)
//        / (nnz(Axx) + nnz(Ayy) - 2 nnz(A))            (:445-890, unsigned count)
//   l <- l - dl_step dl; past ell_max both shrink to 0.7 ell_max; at least ell_min.
// The source and target self-association totals come from P.xx_st / P.yy_st
// (their passes ran with this one; st->ip holds sum(a |y-x|^2)).
__device__ __forceinline__ void controller_adaptive(ProblemState* st, const KConsts& kc, const AssocProblem& P,
                                                    const float* level_h,
                                    int n_levels) {
  cvo_iter_record rec;
  const int k = st->iter;
  float xf[6];
  double xi[6];
  for (int c = 0; c < 6; ++c) {
    xf[c] = (float)st->flow[c];
    xi[c] = (double)xf[c];
  }
  const double nw = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
  const double nv = sqrt(xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5]);
  rec.ell = st->ell;
  rec.nnz = (int32_t)st->nnz;
  rec.max_row = (int32_t)st->max_row;
  rec.num_neighbors = st->K;
  rec.flow_norm = (float)sqrt(nw * nw + nv * nv);
  rec.step = 0.0f;
  rec.dist = 0.0f;
  rec.ip = (float)((double)st->nnz / sqrt((double)kc.n_src_f * (double)kc.n_tgt_f));
  if (nw < (double)kc.eps && nv < (double)kc.eps) {
    st->ret = (nw < 1e-8 && nv < 1e-8) ? -1 : 0;
    st->done = 1;
    st->iters_out = k;
    if (P.trace && k < P.trace_cap) P.trace[k] = rec;
    return;
  }
  double step = select_step_adaptive(kc, st->bcde);
  const double dist = apply_step(st, kc, xi, xf, step);
  st->step = step;
  st->dist = dist;
  rec.step = (float)step;
  rec.dist = (float)dist;
  if (P.trace && k < P.trace_cap) P.trace[k] = rec;
  set_maps(st);
  if (dist < (double)kc.eps_2) {
    st->done = 1;
    st->iters_out = k;
    return;
  }
  const ProblemState* sx = P.xx_st;
  const ProblemState* sy = P.yy_st;
  const float l3 = st->ell * st->ell * st->ell;
  const float dl_a = (float)((-2.0 * st->ip + sx->ip) / (double)l3);
  const float dl_yy = (float)(sy->ip / (double)l3);
  const uint32_t denom = (uint32_t)sx->nnz + (uint32_t)sy->nnz - 2u * (uint32_t)st->nnz;
  const double dl = (double)(dl_a + dl_yy) / (double)denom;
  st->dl = dl;
  float ell = (float)((double)st->ell - kc.dl_step * dl);
  if (ell >= st->ell_max) {
    ell = (float)((double)st->ell_max * 0.7);
    st->ell_max = (float)((double)st->ell_max * 0.7);
  }
  st->ell = ell < kc.ell_min ? kc.ell_min : ell;
  st->iter = k + 1;
  if (st->iter >= kc.max_iter) {
    st->done = 1;
    st->iters_out = kc.max_iter;
  }
  set_level(st, kc, st->ell, level_h, n_levels);
  st->n_overflow = 0;
}

// The self-association problems follow the registration problem of an
// adaptive-lengthscale run: lengthscale, iteration, done flag; the target
// problem's points move with the target map.  One thread.
__global__ void adaptive_sync_kernel(const AssocProblem* __restrict__ probs, KConsts kc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const ProblemState* s0 = probs[0].st;
  for (int q = 1; q < 3; ++q) {
    const AssocProblem& P = probs[q];
    ProblemState* st = P.st;
    st->ell = s0->ell;
    st->K = s0->K;
    st->iter = s0->iter;
    st->done = s0->done;
    st->n_overflow = 0;
    for (int i = 0; i < 9; ++i) {
      const float e = (i % 4 == 0) ? 1.0f : 0.0f;
      st->A[i] = q == 2 ? s0->A[i] : e;
      st->As[i] = q == 2 ? s0->A[i] : e;
      st->Rq[i] = q == 2 ? s0->Rq[i] : e;
    }
    for (int i = 0; i < 3; ++i) {
      st->b[i] = q == 2 ? s0->b[i] : 0.0f;
      st->bs[i] = q == 2 ? s0->b[i] : 0.0f;
      st->Tq[i] = q == 2 ? s0->Tq[i] : 0.0f;
    }
    st->has_src_xf = q == 2 ? 1 : 0;
    set_level(st, kc, st->ell, P.grid);
  }
}


// ---------------------------------------------------------------------------
// Step terms (compute_step_size_xi + compute_step_size_poly_coeff, CvoGPU.cu:
// 333-462): block partials part_s[block] = B, C, D, E (FP64).
// ---------------------------------------------------------------------------

__device__ __forceinline__ void cross_f(float a0, float a1, float a2, float b0, float b1, float b2, float& o0,
                                        float& o1, float& o2) {
  o0 = a1 * b2 - a2 * b1;
  o1 = a2 * b0 - a0 * b2;
  o2 = a0 * b1 - a1 * b0;
}

// The pass's flow, omega / c and v / d (compute_flow_gpu_no_eigen,
// CvoGPU.cu:2057-2060).
__device__ __forceinline__ void scaled_flow(const ProblemState* st, const KConsts& kc, double tot[6]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    tot[c] = st->flow[c] / (double)kc.c;
    tot[3 + c] = st->flow[3 + c] / (double)kc.d;
  }
}

// The controller step of a problem, run by the step pass's last block: step
// coefficients summed over the step blocks in fixed order, then the scalar
// controller on thread 0 with the state staged in LDS.
__device__ __forceinline__ void controller_tail(const AssocProblem& P, const KConsts& kc, const double tot[6], int nsb) {
  ProblemState* st = P.st;
  const int tid = threadIdx.x;
  __shared__ ProblemState s_st;
  __shared__ float s_h[kMaxLevels];
  __shared__ int s_nl;
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(st);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&s_st);
    for (int i = tid; i < (int)(sizeof(ProblemState) / 4); i += kBlock) dst[i] = src[i];
    if (tid == 0) s_nl = P.grid->n_levels;
    if (tid < kMaxLevels) s_h[tid] = P.grid->lv[tid].h;
  }
  // step partials: each thread sums a strided subset of blocks for all four
  // coefficients, then one fixed-order tree
  __shared__ double s_b[4][kBlock];
  {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = tid; b < nsb; b += kBlock)
#pragma unroll
      for (int c = 0; c < 4; ++c) v[c] += ld_dev(&P.part_s[(size_t)b * 4 + c]);
#pragma unroll
    for (int c = 0; c < 4; ++c) s_b[c][tid] = v[c];
  }
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    for (int t = tid; t < 4 * w; t += kBlock) {
      const int c = t / w, i = t - c * w;
      s_b[c][i] += s_b[c][i + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    for (int c = 0; c < 6; ++c) s_st.flow[c] = tot[c];
    for (int c = 0; c < 4; ++c) s_st.bcde[c] = s_b[c][0];
    if (kc.adaptive) controller_adaptive(&s_st, kc, P, s_h, s_nl);
    else controller(&s_st, kc, P, s_h, s_nl);
  }
  __syncthreads();
  {
    // the pass counters live at the device coherence point: never written
    // back from the staged copy; this pass's counter returns to zero
    constexpr int c0 = (int)(offsetof(ProblemState, blocks_done_assoc) / 4);
    constexpr int c1 = (int)(offsetof(ProblemState, blocks_done_step) / 4) + 1;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&s_st);
    uint32_t* dst = reinterpret_cast<uint32_t*>(st);
    for (int i = tid; i < (int)(sizeof(ProblemState) / 4); i += kBlock)
      if (i < c0 || i >= c1) dst[i] = src[i];
    if (tid == 0) st_dev(&st->blocks_done_step, 0u);
  }
}

// One thread per source row; grid (row blocks, problems).  The last block of
// a problem runs its controller step.
constexpr int kStepLoads = 4;  // pairs whose loads are in flight together per thread
__device__ __forceinline__ void step_pass(const AssocProblem& P, const KConsts& kc) {
  ProblemState* st = P.st;
  if (st->done) return;
  const int tid = threadIdx.x;
  const int r = blockIdx.x * kBlock + tid;
  const int nsb = (P.n_src + kBlock - 1) / kBlock;
  if ((int)blockIdx.x >= nsb) return;  // uniform: blocks beyond this problem's rows
  double tot[6];
  scaled_flow(st, kc, tot);
  const float ell = st->ell;
  const GridLevel G = P.grid->lv[st->level];  // a register copy, as in the association
  Maps mp;
  load_maps(st, mp);
  // normalised flow (compute_flow, CvoGPU.cu:2099-2104), used in float; the
  // adaptive-lengthscale loop steps along the flow as summed
  // (AdaptiveCvoGPU.cu:839-840)
  double n2 = 0.0;
#pragma unroll
  for (int c = 0; c < 6; ++c) n2 += tot[c] * tot[c];
  const double nrm = (n2 > 0.0 && !kc.adaptive) ? sqrt(n2) : 1.0;
  const float w0 = (float)(tot[0] / nrm), w1 = (float)(tot[1] / nrm), w2 = (float)(tot[2] / nrm);
  const float v0 = (float)(tot[3] / nrm), v1 = (float)(tot[4] / nrm), v2 = (float)(tot[5] / nrm);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (r < P.n_src) {
    float4 x4 = ld_global(P.src_pos + r);
    float x0 = x4.x, x1 = x4.y, x2 = x4.z;
    if (mp.src_xf) {
      float t0, t1, t2;
      affine3(mp.As, mp.bs, x0, x1, x2, t0, t1, t2);
      x0 = t0; x1 = t1; x2 = t2;
    }
    float temp_ell = ell;
    if (kc.range_ell_in_step) {
      const float px = sqrtf(x0 * x0 + x1 * x1 + x2 * x2);
      temp_ell = (float)(((double)px / 500.0 + 1.0) * (double)ell);
    }
    const float tc = (float)(1.0 / (2.0 * (double)temp_ell * (double)temp_ell));
    const uint32_t cnt = ld_global(P.counts + r);
    const uint2* row = P.pairs + (size_t)r * P.cap_alloc;
    // one pair's Taylor terms, summed into acc in pair order
    auto step_pair = [&](const float a, const float4 y4) {
      float z0, z1, z2;
      affine3(mp.A, mp.b, y4.x, y4.y, y4.z, z0, z1, z2);
      // xi^k z: xiz = w x z + v, xi^{k+1} z = w x xi^k z
      float a0, a1, a2, b0, b1, b2, c0, c1, c2, d0, d1, d2;
      cross_f(w0, w1, w2, z0, z1, z2, a0, a1, a2);
      a0 += v0; a1 += v1; a2 += v2;
      cross_f(w0, w1, w2, a0, a1, a2, b0, b1, b2);
      cross_f(w0, w1, w2, b0, b1, b2, c0, c1, c2);
      cross_f(w0, w1, w2, c0, c1, c2, d0, d1, d2);
      const float df0 = x0 - z0, df1 = x1 - z1, df2 = x2 - z2;
      const float normxiz2 = a0 * a0 + a1 * a1 + a2 * a2;
      const float xiz_dot_xi2z = -(a0 * b0 + a1 * b1 + a2 * b2);
      const float epsil_const = (b0 * b0 + b1 * b1 + b2 * b2) + 2.0f * (a0 * c0 + a1 * c1 + a2 * c2);
      const float beta = (float)(-2.0 * (double)tc * (double)(a0 * df0 + a1 * df1 + a2 * df2));
      const float gamma = -tc * (normxiz2 + 2.0f * (b0 * df0 + b1 * df1 + b2 * df2));
      const float delta = (float)(2.0 * (double)tc * (double)(xiz_dot_xi2z - (c0 * df0 + c1 * df1 + c2 * df2)));
      const float epsil = -tc * (epsil_const + 2.0f * (d0 * df0 + d1 * df1 + d2 * df2));
      const double ad = (double)a, bd = (double)beta, gd = (double)gamma;
      acc[0] += (double)(a * beta);
      acc[1] += ad * (gd + (double)(beta * beta) / 2.0);
      acc[2] += ad * ((double)(delta + beta * gamma) + (double)(beta * beta * beta) / 6.0);
      acc[3] += ad * ((double)(epsil + beta * delta) + 0.5 * bd * bd * gd + 0.5 * gd * gd +
                      (1.0 / 24.0) * bd * bd * bd * bd);
    };
    // kStepLoads pairs a trip: their slots, then their target gathers, go out
    // before the first pair is summed
    uint32_t s = 0;
    for (; s + kStepLoads <= cnt; s += kStepLoads) {
      uint2 e[kStepLoads];
      float4 y[kStepLoads];
#pragma unroll
      for (int u = 0; u < kStepLoads; ++u) e[u] = ld_global(row + s + u);
#pragma unroll
      for (int u = 0; u < kStepLoads; ++u) y[u] = ld_global(G.pos + e[u].x);
#pragma unroll
      for (int u = 0; u < kStepLoads; ++u) step_pair(__uint_as_float(e[u].y), y[u]);
    }
    for (; s < cnt; ++s) {
      const uint2 e = ld_global(row + s);
      step_pair(__uint_as_float(e.y), ld_global(G.pos + e.x));
    }
  }
  // block partial in fixed order: wave shuffles then LDS over the 4 waves
  __shared__ double s_w[4][kBlock / 64];
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double v = acc[c];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) s_w[c][wave] = v;
  }
  __syncthreads();
  __shared__ int s_last;
  if (tid < 64) {
    if (tid < 4) {
      double v = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) v += s_w[tid][w];
      st_dev(&P.part_s[(size_t)blockIdx.x * 4 + tid], v);
    }
    wait_stores();
    if (tid == 0) s_last = (count_in(&st->blocks_done_step) == (uint32_t)(nsb - 1)) ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;  // uniform
  controller_tail(P, kc, tot, nsb);
}

template <bool ONE>
__global__ __launch_bounds__(kBlock) void step_kernel(const AssocProblem* __restrict__ probs, KConsts kc,
                                                      AssocProblem one) {
  if constexpr (ONE) step_pass(one, kc);
  else {
    const AssocProblem P = probs[blockIdx.y];  // a register copy
    step_pass(P, kc);
  }
}

// Initial state of a problem: pose from a row-major float 4x4, maps, level.
__global__ void init_state_kernel(const AssocProblem* __restrict__ probs, int n_probs, const float* __restrict__ T16,
                                  int stride16, KConsts kc, float ell0, int K0) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_probs) return;
  const AssocProblem& P = probs[p];
  ProblemState* st = P.st;
  const float* T = T16 + (size_t)p * stride16;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) st->R[r * 3 + c] = (double)T[r * 4 + c];
    st->T[r] = (double)T[r * 4 + 3];
  }
  for (int i = 0; i < 9; ++i) st->As[i] = (i % 4 == 0) ? 1.0f : 0.0f;
  for (int i = 0; i < 3; ++i) st->bs[i] = 0.0f;
  st->has_src_xf = 0;
  for (int c = 0; c < 6; ++c) st->flow[c] = 0.0;
  for (int c = 0; c < 4; ++c) st->bcde[c] = 0.0;
  st->ip = 0.0;
  st->step = 0.0;
  st->dist = 0.0;
  st->nnz = 0;
  st->ell = ell0;
  st->K = K0;
  st->iter = 0;
  st->done = 0;
  st->ret = 0;
  st->iters_out = 0;
  st->max_row = 0;
  st->n_overflow = 0;
  st->blocks_done_assoc = st->blocks_done_ovf = st->blocks_done_step = 0;
  st->qs_head = st->qs_size = st->qe_head = st->qe_size = 0;
  st->ss = st->se = 0.0f;
  st->ell_max = kc.ell_max0;
  st->dl = 0.0;
  set_maps(st);
  set_level(st, kc, ell0, P.grid);
}

// Switches every problem to a single scoring pass at lengthscale ell with the
// full row cap, keeping its pose (multi-start scoring by cos angle).
__global__ void rescore_state_kernel(const AssocProblem* __restrict__ probs, int n_probs, KConsts kc, float ell,
                                     int K) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_probs) return;
  const AssocProblem& P = probs[p];
  ProblemState* st = P.st;
  st->ell = ell;
  st->K = K;
  st->done = 0;
  st->n_overflow = 0;
  set_maps(st);
  set_level(st, kc, ell, P.grid);
}

// ---------------------------------------------------------------------------
// Association export: CSR in input row order, target ids ascending per row.
// ---------------------------------------------------------------------------

__global__ void export_counts(const AssocProblem* __restrict__ P, int64_t* __restrict__ counts_orig) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P->n_src) return;
  uint32_t orig = __float_as_uint(P->src_pos[r].w);
  counts_orig[orig] = (int64_t)P->counts[r];
}

// One 64-lane wave per row: gathers (target id, a), bitonic sort by id in LDS.
// One row of the association export, sorted in LDS (bitonic, row <= 1024
// pairs) by target id, or -- for rows with more than sort_threshold pairs when
// sort_threshold > 0 -- by value descending with ties by id, the order the
// reference's multiframe_is_sorting_inner_product gives a row
// (CvoGPU.cu:1080-1095).  Keys: (id, value bits) or (~value bits, id); the
// values are positive, so their bits order like the values.
__global__ __launch_bounds__(64) void export_rows(const AssocProblem* __restrict__ P, const int64_t* __restrict__ row_ptr,
                                                  int32_t* __restrict__ col, float* __restrict__ val,
                                                  int sort_threshold) {
  __shared__ uint64_t s_key[1024];
  const int r = blockIdx.x;
  if (r >= P->n_src) return;
  const GridLevel& G = P->grid->lv[P->st->level_assoc];
  const uint32_t cnt = P->counts[r];
  const uint32_t orig = __float_as_uint(P->src_pos[r].w);
  const bool by_value = sort_threshold > 0 && cnt > (uint32_t)sort_threshold;
  uint32_t n2 = 1;
  while (n2 < cnt) n2 <<= 1;
  // bundle adjustment rows: (target coordinates, a) at the pass's row stride
  // K, the sorted target positions beside them
  const bool ba_rows = P->pairs4 != nullptr;
  const size_t row0 = ba_rows ? (size_t)r * (size_t)P->st->K : (size_t)r * P->cap_alloc;
  for (uint32_t s = threadIdx.x; s < n2; s += 64) {
    uint64_t key = ~0ull;
    if (s < cnt) {
      const uint2 e = ba_rows ? make_uint2(P->pair_ids[row0 + s], __float_as_uint(P->pairs4[row0 + s].w))
                              : P->pairs[row0 + s];
      const uint32_t id = __float_as_uint(G.pos[e.x].w);
      key = by_value ? (((uint64_t)(~e.y)) << 32) | id : (((uint64_t)id) << 32) | e.y;
    }
    s_key[s] = key;
  }
  __syncthreads();
  block_bitonic_sort(s_key, n2);
  const int64_t base = row_ptr[orig];
  for (uint32_t s = threadIdx.x; s < cnt; s += 64) {
    const uint64_t key = s_key[s];
    const uint32_t hi = (uint32_t)(key >> 32), lo = (uint32_t)key;
    col[base + s] = (int32_t)(by_value ? lo : hi);
    val[base + s] = __uint_as_float(by_value ? ~hi : lo);
  }
}

// All-done flag of a set of problems (one block), for the host's chunk loop.
__global__ void all_done_kernel(const AssocProblem* __restrict__ probs, int n_probs, int32_t* __restrict__ out) {
  __shared__ int s_any;
  if (threadIdx.x == 0) s_any = 0;
  __syncthreads();
  for (int q = threadIdx.x; q < n_probs; q += blockDim.x)
    if (!probs[q].st->done) s_any = 1;
  __syncthreads();
  if (threadIdx.x == 0) *out = s_any ? 0 : 1;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

void launch_all_done(const AssocProblem* probs, int n_probs, int32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(all_done_kernel, dim3(1), dim3(256), 0, s, probs, n_probs, out);
}

// A launch, bracketed by the events e0 / e1 at the kernel's own start and end
// when e0 is given (profiling: the events time the kernel alone, as a kernel
// tracer does, not the dispatch around it).
#define CVO_LAUNCH(KERNEL, GRID, BLOCK, STREAM, E0, E1, ...)                                  \
  do {                                                                                      \
    if (E0) hipExtLaunchKernelGGL(KERNEL, GRID, BLOCK, 0, STREAM, E0, E1, 0, __VA_ARGS__); \
    else hipLaunchKernelGGL(KERNEL, GRID, BLOCK, 0, STREAM, __VA_ARGS__);                  \
  } while (0)

template <int FC, int KC>
static void launch_assoc_t(const AssocProblem* probs, int n_probs, int total_blocks, const KConsts& kc,
                           const AssocProblem& one, bool use_one, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  const dim3 g(total_blocks), b(kBlock);
  if (kc.knn > 0) {
    if (use_one) CVO_LAUNCH((knn_assoc_kernel<FC, KC, true>), g, b, s, e0, e1, probs, n_probs, kc, one);
    else CVO_LAUNCH((knn_assoc_kernel<FC, KC, false>), g, b, s, e0, e1, probs, n_probs, kc, one);
  } else {
    if (use_one) CVO_LAUNCH((assoc_kernel<FC, KC, true>), g, b, s, e0, e1, probs, n_probs, kc, one);
    else CVO_LAUNCH((assoc_kernel<FC, KC, false>), g, b, s, e0, e1, probs, n_probs, kc, one);
  }
}

template <int FC, int KC>
static void launch_ovf_t(const AssocProblem* probs, int n_probs, const KConsts& kc, const AssocProblem& one,
                         bool use_one, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  const dim3 g(kOvfBlocks, n_probs), b(kBlock);
  if (use_one) CVO_LAUNCH((overflow_kernel<FC, KC, true>), g, b, s, e0, e1, probs, kc, one);
  else CVO_LAUNCH((overflow_kernel<FC, KC, false>), g, b, s, e0, e1, probs, kc, one);
}

static int fsel(const KConsts& kc) {
  if (!kc.use_col) return 0;
  return kc.F == 5 ? 5 : -1;
}
static int ksel(const KConsts& kc) {
  if (!kc.use_sem) return 0;
  return kc.Ks == 19 ? 19 : -1;
}

#define CVO_DISPATCH_FK(FN, ...)                              \
  do {                                                        \
    int f_ = fsel(kc), k_ = ksel(kc);                         \
    if (f_ == 0 && k_ == 0) FN<0, 0>(__VA_ARGS__);            \
    else if (f_ == 0 && k_ == 19) FN<0, 19>(__VA_ARGS__);     \
    else if (f_ == 0) FN<0, -1>(__VA_ARGS__);                 \
    else if (f_ == 5 && k_ == 0) FN<5, 0>(__VA_ARGS__);       \
    else if (f_ == 5 && k_ == 19) FN<5, 19>(__VA_ARGS__);     \
    else if (f_ == 5) FN<5, -1>(__VA_ARGS__);                 \
    else if (k_ == 0) FN<-1, 0>(__VA_ARGS__);                 \
    else if (k_ == 19) FN<-1, 19>(__VA_ARGS__);               \
    else FN<-1, -1>(__VA_ARGS__);                             \
  } while (0)

void launch_assoc(const AssocProblem* probs, int n_probs, int total_blocks, const KConsts& kc, hipStream_t s,
                  const AssocProblem* host_one, hipEvent_t e0, hipEvent_t e1) {
  AssocProblem one{};
  const bool use_one = n_probs == 1 && host_one;
  if (use_one) one = *host_one;
  CVO_DISPATCH_FK(launch_assoc_t, probs, n_probs, total_blocks, kc, one, use_one, s, e0, e1);
}

void launch_overflow(const AssocProblem* probs, int n_probs, const KConsts& kc, hipStream_t s,
                     const AssocProblem* host_one, hipEvent_t e0, hipEvent_t e1) {
  AssocProblem one{};
  const bool use_one = n_probs == 1 && host_one;
  if (use_one) one = *host_one;
  CVO_DISPATCH_FK(launch_ovf_t, probs, n_probs, kc, one, use_one, s, e0, e1);
}

void launch_step(const AssocProblem* probs, int n_probs, int max_src_rows, const KConsts& kc, hipStream_t s,
                 const AssocProblem* host_one, hipEvent_t e0, hipEvent_t e1) {
  const dim3 g((max_src_rows + kBlock - 1) / kBlock, n_probs), b(kBlock);
  AssocProblem one{};
  if (n_probs == 1 && host_one) {
    one = *host_one;
    CVO_LAUNCH(step_kernel<true>, g, b, s, e0, e1, probs, kc, one);
  } else {
    CVO_LAUNCH(step_kernel<false>, g, b, s, e0, e1, probs, kc, one);
  }
}

void launch_adaptive_sync(const AssocProblem* probs, const KConsts& kc, hipStream_t s) {
  hipLaunchKernelGGL(adaptive_sync_kernel, dim3(1), dim3(64), 0, s, probs, kc);
}

void launch_rescore(const AssocProblem* probs, int n_probs, const KConsts& kc, float ell, int K, hipStream_t s) {
  hipLaunchKernelGGL(rescore_state_kernel, dim3((n_probs + 63) / 64), dim3(64), 0, s, probs, n_probs, kc, ell, K);
}


void launch_init_state(const AssocProblem* probs, int n_probs, const float* T16, int stride16, const KConsts& kc,
                       float ell0, int K0, hipStream_t s) {
  hipLaunchKernelGGL(init_state_kernel, dim3((n_probs + 63) / 64), dim3(64), 0, s, probs, n_probs, T16, stride16, kc,
                     ell0, K0);
}

void launch_grid_count(const float* xyz, int n, const GridLevel& L, uint32_t* key, uint32_t* count, hipStream_t s) {
  hipLaunchKernelGGL(grid_cell_count, dim3((n + 255) / 256), dim3(256), 0, s, xyz, n, L.ox, L.oy, L.oz, L.inv_h, L.nx,
                     L.ny, L.nz, key, count);
}

// The controller's device math on given inputs, for checks against
// published values: the cubic (roots of coef[4i..4i+3], highest power first;
// re / im 3 each), the FP64 SE(3) exponential of dt[i] * xi[6i..] (row-major
// 3x4) with the log norm of its result, and the reference's float exponential
// of the same twist.
__global__ void math_check_kernel(int n_cubic, const double* __restrict__ coef, double* __restrict__ re,
                                  double* __restrict__ im, int n_exp, const double* __restrict__ xi,
                                  const double* __restrict__ dt, double* __restrict__ T64,
                                  double* __restrict__ log_norm, float* __restrict__ T32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_cubic) {
    double r[3], m[3];
    cubic_solve(coef[4 * i], coef[4 * i + 1], coef[4 * i + 2], coef[4 * i + 3], r, m);
    for (int k = 0; k < 3; ++k) {
      re[3 * i + k] = r[k];
      im[3 * i + k] = m[k];
    }
  }
  if (i < n_exp) {
    double R[9], t[3];
    se3_exp(xi + 6 * i, dt[i], R, t);
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) T64[12 * i + 4 * r + c] = R[3 * r + c];
      T64[12 * i + 4 * r + 3] = t[r];
    }
    log_norm[i] = se3_log_norm(R, t);
    float xf[6];
    for (int k = 0; k < 6; ++k) xf[k] = (float)xi[6 * i + k];
    exp_sek3_f32(xf, (float)dt[i], T32 + 12 * i);
  }
}

void launch_math_check(int n_cubic, const double* coef, double* re, double* im, int n_exp, const double* xi,
                       const double* dt, double* T64, double* log_norm, float* T32, hipStream_t s) {
  const int n = n_cubic > n_exp ? n_cubic : n_exp;
  hipLaunchKernelGGL(math_check_kernel, dim3((n + 63) / 64), dim3(64), 0, s, n_cubic, coef, re, im, n_exp, xi, dt,
                     T64, log_norm, T32);
}

void launch_zero(void* p, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 4;
  if (n == 0) return;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(zero_u32, dim3((unsigned)blocks), dim3(256), 0, s, static_cast<uint32_t*>(p), n);
}

void launch_iota(uint32_t* v, int n, hipStream_t s) {
  hipLaunchKernelGGL(iota_u32, dim3((n + 255) / 256), dim3(256), 0, s, v, n);
}

void launch_gather(const uint32_t* order, int n, const float* xyz, const float* feat, int F, const float* sem, int K,
                   const float* geo, float4* pos, float* feat_s, float* sem_s, float* geo_s, uint32_t* inv_perm,
                   hipStream_t s) {
  hipLaunchKernelGGL(gather_sorted, dim3((n + 255) / 256), dim3(256), 0, s, order, n, xyz, feat, F, sem, K, geo, pos,
                     feat_s, sem_s, geo_s, inv_perm);
}

void launch_morton(const float* xyz, int n, const float* bmin, const float* bmax, uint32_t* key, hipStream_t s) {
  float sc[3];
  for (int d = 0; d < 3; ++d) {
    float ext = bmax[d] - bmin[d];
    sc[d] = ext > 0 ? 1023.0f / ext : 0.0f;
  }
  hipLaunchKernelGGL(morton_keys, dim3((n + 255) / 256), dim3(256), 0, s, xyz, n, bmin[0], bmin[1], bmin[2], sc[0],
                     sc[1], sc[2], key);
}

void launch_export(const AssocProblem* P_dev, int n_src, int64_t* counts_orig, const int64_t* row_ptr, int32_t* col,
                   float* val, int phase, int sort_threshold, hipStream_t s) {
  if (phase == 0)
    hipLaunchKernelGGL(export_counts, dim3((n_src + 255) / 256), dim3(256), 0, s, P_dev, counts_orig);
  else
    hipLaunchKernelGGL(export_rows, dim3(n_src), dim3(64), 0, s, P_dev, row_ptr, col, val, sort_threshold);
}

}  // namespace cvo
