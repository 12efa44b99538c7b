// Association with a per-point anisotropic kernel: CvoGPU::compute_association_gpu
// with a 3x3 kernel (src/cvo/CvoGPU.cu:1873-1952, inner_product_non_isotropic_impl;
// fill_in_A_mat_gpu_dense_mat_kernel :719-830, rotate_kernel_and_inverse
// :134-161, mahananobis_distance :164-181).
//
// For a source point x: z = x / |x|, u = (1, 1, -(z0 + z1) / z2), w = z x u,
// U = [u w z] (columns), M = U K^-1 U^T; for a moved target point y',
// d2 = (x - y')^T M (x - y') and k = sigma^2 exp(-d2 / 2) -- no lengthscale and
// no distance gate -- times the colour and semantic factors with their gates
// (geometric type is switched off by the reference for this path).  A row
// keeps the first nearest_neighbors_max passing targets in ascending index.
//
// One wave per source row.  With the geometric factor on, a pair can reach
// sp_thres only when d2 <= 2 ln(amplitude / sp_thres) (the product of the
// enabled factors' amplitudes), an ellipsoid around x when M is positive
// definite: the wave enumerates the target grid cells of its bounding box,
// selects the first K passing targets by index among them (radix select on
// the index) and stores those; the compaction sorts each row by target id.
// A row whose metric is not safely positive definite (see
// ellipsoid_extents), and every row when the geometric factor is off, scans
// all targets in index order instead, 64 at a time, a ballot giving each
// passing lane its slot.  Pairs beyond the d2 bound skip the exponentials.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <math.h>
#include <stdint.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cvo_internal.h"
#include "cvo_device.h"
#include "cvo_ball.h"

namespace cvo {

struct AnisoArgs {
  const float* sx;
  const float* sf;
  const float* ss;
  const float* tx;
  const float* tf;
  const float* ts;
  int n, m, F, Ks, K;
  float A[9], b[3];    // target map y' = A y + b
  float Rq[9], Tq[3];  // its inverse (the pose): grid frame of a mapped point
  float Kinv[9];       // row-major inverse kernel
  float sigma2, c_sigma2, s_sigma2, c_ell2, s_ell2, sp, d2_c_thres, d2_s_thres;
  float d2_cut;  // beyond it the product kernel cannot exceed sp_thres
  int use_geo, use_col, use_sem;
  const GridDev* grid;  // one level over the target; null: every row scans
  uint32_t* counts;     // n
  int32_t* cols;        // n * K
  float* vals;          // n * K
};

// M = U Kinv U^T for the source point x (rotate_kernel_and_inverse).  A point
// at the origin or with z2 = 0 gives a non-finite M, hence no pairs.
__device__ inline void row_metric(const float x[3], const float* Kinv, float M[9]) {
  float z[3] = {x[0], x[1], x[2]};
  const float n2 = fmaf(x[2], x[2], fmaf(x[1], x[1], x[0] * x[0]));
  if (n2 > 0.0f) {
    const float nr = sqrtf(n2);
    z[0] = x[0] / nr;
    z[1] = x[1] / nr;
    z[2] = x[2] / nr;
  }
  const float u[3] = {1.0f, 1.0f, (-1.0f) / z[2] * (z[0] + z[1])};
  const float w[3] = {z[1] * u[2] - z[2] * u[1], z[2] * u[0] - z[0] * u[2], z[0] * u[1] - z[1] * u[0]};
  const float U[9] = {u[0], w[0], z[0], u[1], w[1], z[1], u[2], w[2], z[2]};  // row-major, columns u w z
  float T1[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      T1[r * 3 + c] = fmaf(U[r * 3 + 2], Kinv[2 * 3 + c], fmaf(U[r * 3 + 1], Kinv[1 * 3 + c], U[r * 3 + 0] * Kinv[c]));
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      M[r * 3 + c] = fmaf(T1[r * 3 + 2], U[c * 3 + 2], fmaf(T1[r * 3 + 1], U[c * 3 + 1], T1[r * 3 + 0] * U[c * 3 + 0]));
}

// The product kernel of source row x (metric M, features xf, classes xs)
// against the raw target point (y0r, y1r, y2r) with features yf, classes ys.
__device__ __forceinline__ bool aniso_pair(const AnisoArgs& g, const float x[3], const float M[9], const float* xf,
                                           const float* xs, float y0r, float y1r, float y2r, const float* yf,
                                           const float* ys, float& a) {
  float y[3];
  affine3(g.A, g.b, y0r, y1r, y2r, y[0], y[1], y[2]);
  float k = 1.0f, ck = 1.0f, sk = 1.0f;
  if (g.use_geo) {
    const float d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
    // (d^T M) d
    const float t0 = fmaf(d[2], M[6], fmaf(d[1], M[3], d[0] * M[0]));
    const float t1 = fmaf(d[2], M[7], fmaf(d[1], M[4], d[0] * M[1]));
    const float t2 = fmaf(d[2], M[8], fmaf(d[1], M[5], d[0] * M[2]));
    const float d2 = fmaf(t2, d[2], fmaf(t1, d[1], t0 * d[0]));
    if (d2 > g.d2_cut) return false;
    k = (float)((double)g.sigma2 * exp(-(double)d2 / 2.0));
  }
  if (g.use_col) {
    float d2c = 0.0f;
    for (int f = 0; f < g.F; ++f) {
      const float t = ld_global(xf + f) - ld_global(yf + f);  // both clouds' features in global memory
      d2c = fmaf(t, t, d2c);
    }
    if (!(d2c < g.d2_c_thres)) return false;
    ck = (float)((double)g.c_sigma2 * exp(-(double)d2c / (2.0 * (double)g.c_ell2)));
  }
  if (g.use_sem) {
    float d2s = 0.0f;
    for (int c = 0; c < g.Ks; ++c) {
      const float t = ld_global(xs + c) - ld_global(ys + c);
      d2s = fmaf(t, t, d2s);
    }
    if (!(d2s < g.d2_s_thres)) return false;
    sk = (float)((double)g.s_sigma2 * exp(-(double)d2s / (2.0 * (double)g.s_ell2)));
  }
  a = ck * k * sk * 1.0f;
  return a > g.sp;
}

// Row r against every target in ascending index (the wave's 64 lanes).
__device__ inline void aniso_scan_row(const AnisoArgs& g, int r, int lane, const float x[3], const float M[9]) {
  const float* xf = g.sf + (size_t)r * g.F;
  const float* xs = g.ss + (size_t)r * g.Ks;
  const uint32_t K = (uint32_t)g.K;
  uint32_t stored = 0;
  for (int base = 0; base < g.m && stored < K; base += 64) {
    const int j = base + lane;
    bool pass = false;
    float a = 0.0f;
    if (j < g.m)
      pass = aniso_pair(g, x, M, xf, xs, g.tx[3 * (size_t)j], g.tx[3 * (size_t)j + 1], g.tx[3 * (size_t)j + 2],
                        g.tf + (size_t)j * g.F, g.ts + (size_t)j * g.Ks, a);
    const uint64_t mask = __ballot(pass);
    const uint32_t slot = stored + (uint32_t)__popcll(mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    if (pass && slot < K) {
      g.cols[(size_t)r * K + slot] = j;
      g.vals[(size_t)r * K + slot] = a;
    }
    stored += (uint32_t)__popcll(mask);
  }
  if (lane == 0) g.counts[r] = stored < K ? stored : K;
}

// A row's ellipsoid d^T M d <= c in the grid (raw target) frame: with Rq the
// pose rotation, a grid point's offset from the row's query q = Rq x + Tq is
// delta = -Rq d up to rounding, so delta^T Q delta <= c with Q = Rq M Rq^T.
// The candidates are the cells of its bounding box; within one run of cells
// along x (fixed y, z cells) only x near the ellipsoid's centre line
// x*(dy, dz) = ax dy + az dz (the x minimising the form) can lie inside:
// at most sqrt(c / Qxx) away.
struct RowEllipsoid {
  float q0, q1, q2;  // centre
  float h[3];        // bounding box half-extents
  float ax, az;      // centre line
  float w;           // half-width of a chord along x
  float slack;       // absolute margin for the float mapping and cell bounds
};

// False when M (its symmetric part) is not positive definite, or when the
// float evaluation of d^T M d could fall below the bound for a d outside the
// ellipsoid of c: that evaluation errs by at most ~6u |d|^T |M| |d| <= 6u
// ||M||_F ||M^-1||_F d^T M d (u = 2^-24), so with 8u ||M||_F ||M^-1||_F <=
// 0.01 every d it passes lies within the ellipsoid of cut / 0.99, and
// c = 1.02 cut covers that.
__device__ inline bool row_ellipsoid(const float M[9], float cut, const float Rq[9], float q0, float q1, float q2,
                                     RowEllipsoid& E) {
  const double S[9] = {M[0], 0.5 * ((double)M[1] + (double)M[3]), 0.5 * ((double)M[2] + (double)M[6]),
                       0.0,  M[4],                                  0.5 * ((double)M[5] + (double)M[7]),
                       0.0,  0.0,                                   M[8]};
  const double s00 = S[0], s01 = S[1], s02 = S[2], s11 = S[4], s12 = S[5], s22 = S[8];
  const double c00 = s11 * s22 - s12 * s12, c11 = s00 * s22 - s02 * s02, c22 = s00 * s11 - s01 * s01;
  const double c01 = s02 * s12 - s01 * s22, c02 = s01 * s12 - s02 * s11, c12 = s01 * s02 - s00 * s12;
  const double det = s00 * c00 + s01 * c01 + s02 * c02;
  if (!(s00 > 0.0) || !(c22 > 0.0) || !(det > 0.0) || !isfinite(det)) return false;
  double fm = 0.0;
  for (int i = 0; i < 9; ++i) fm += (double)M[i] * (double)M[i];
  const double fi = (c00 * c00 + c11 * c11 + c22 * c22 + 2.0 * (c01 * c01 + c02 * c02 + c12 * c12)) / (det * det);
  if (!(8.0 * 5.9604644775390625e-8 * sqrt(fm * fi) <= 0.01)) return false;
  // symmetric S and S^-1 in full
  const double Sf[9] = {s00, s01, s02, s01, s11, s12, s02, s12, s22};
  const double If[9] = {c00 / det, c01 / det, c02 / det, c01 / det, c11 / det, c12 / det,
                        c02 / det, c12 / det, c22 / det};
  // Q = Rq S Rq^T and Q^-1 = Rq S^-1 Rq^T
  double Q[9], Qi[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double a = 0.0, b = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          const double rr = (double)Rq[3 * i + k] * (double)Rq[3 * j + l];
          a += rr * Sf[3 * k + l];
          b += rr * If[3 * k + l];
        }
      Q[3 * i + j] = a;
      Qi[3 * i + j] = b;
    }
  if (!(Q[0] > 0.0)) return false;
  const double c = (double)cut * 1.02;
  E.q0 = q0;
  E.q1 = q1;
  E.q2 = q2;
  E.slack = 1e-4f * (1.0f + fmaxf(fabsf(q0), fmaxf(fabsf(q1), fabsf(q2))));
  for (int j = 0; j < 3; ++j) E.h[j] = (float)sqrt(c * fmax(Qi[4 * j], 0.0)) + E.slack;
  E.ax = (float)(-Q[1] / Q[0]);
  E.az = (float)(-Q[2] / Q[0]);
  E.w = (float)sqrt(c / Q[0]) + E.slack;
  return isfinite(E.h[0]) && isfinite(E.h[1]) && isfinite(E.h[2]) && isfinite(E.ax) && isfinite(E.az) &&
         isfinite(E.w);
}

// Calls fn(valid, sorted position) for every target of the box's runs, each
// run clipped to the x window where the ellipsoid can lie for the run's y, z
// cells; the wave's 64 lanes call together.
template <class Fn>
__device__ inline void for_each_in_ellipsoid(const GridLevel& GL, const CellBox& B, const RowEllipsoid& E, int g,
                                             Fn&& fn) {
  for (int rbase = 0; rbase < B.nruns; rbase += 64) {
    const int myrun = rbase + g;
    uint32_t rs = 0, rl = 0;
    if (myrun < B.nruns) {
      const int iz = myrun / B.span1, iy = myrun - iz * B.span1;
      const int cy = B.lo1 + iy, cz = B.lo2 + iz;
      // the run's y and z ranges relative to the centre, with the margin
      const float y0 = GL.oy + (float)cy * GL.h - E.q1 - E.slack, y1 = y0 + GL.h + 2.0f * E.slack;
      const float z0 = GL.oz + (float)cz * GL.h - E.q2 - E.slack, z1 = z0 + GL.h + 2.0f * E.slack;
      const float xmin = E.ax * (E.ax > 0.0f ? y0 : y1) + E.az * (E.az > 0.0f ? z0 : z1);
      const float xmax = E.ax * (E.ax > 0.0f ? y1 : y0) + E.az * (E.az > 0.0f ? z1 : z0);
      const int lo = max(cell_floor((E.q0 + xmin - E.w - GL.ox) * GL.inv_h, GL.nx), B.lo0);
      const int hi = min(cell_floor((E.q0 + xmax + E.w - GL.ox) * GL.inv_h, GL.nx), B.lo0 + B.nx_run - 1);
      if (lo <= hi) {
        const uint32_t c0 = ((uint32_t)cz * (uint32_t)GL.ny + (uint32_t)cy) * (uint32_t)GL.nx + (uint32_t)lo;
        rs = ld_global(GL.cell_start + c0);
        rl = ld_global(GL.cell_start + c0 + (uint32_t)(hi - lo + 1)) - rs;
      }
    }
    uint32_t incl = rl;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t t = __shfl_up(incl, o, 64);
      if (g >= o) incl += t;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    const uint32_t excl = incl - rl;
    for (uint32_t base = 0; base < total; base += 64) {
      const uint32_t idx = base + (uint32_t)g;
      int kk = 0;
#pragma unroll
      for (int s = 32; s > 0; s >>= 1) {
        uint32_t e = __shfl(excl, kk + s, 64);
        if (e <= idx) kk += s;
      }
      const uint32_t run_start = __shfl(rs, kk, 64);
      const uint32_t run_excl = __shfl(excl, kk, 64);
      const bool valid = idx < total;
      fn(valid, valid ? run_start + (idx - run_excl) : 0u);
    }
  }
}

// Row r against the targets of its ellipsoid: the first K passing by index,
// stored in enumeration order.  One pass stores while it counts; a row with
// more than K passing targets then selects the K smallest indices and stores
// those instead.
__device__ inline void aniso_box_row(const AnisoArgs& g, int r, int lane, const float x[3], const float M[9],
                                     const RowEllipsoid& E, uint32_t* hist) {
  const GridLevel GL = g.grid->lv[0];  // a register copy: the pair stores cannot alias it
  const float* xf = g.sf + (size_t)r * g.F;
  const float* xs = g.ss + (size_t)r * g.Ks;
  const uint32_t K = (uint32_t)g.K;
  const CellBox B = cell_box3(GL, E.q0, E.q1, E.q2, E.h[0], E.h[1], E.h[2]);
  auto each = [&](auto&& fn) { for_each_in_ellipsoid(GL, B, E, lane, fn); };
  // a passing target's key is its index; ~0 for the others
  auto eval = [&](uint32_t pidx, float& a) -> uint64_t {
    const float4 y4 = ld_global(GL.pos + pidx);
    const bool pass = aniso_pair(g, x, M, xf, xs, y4.x, y4.y, y4.z, GL.feat ? GL.feat + (size_t)pidx * g.F : nullptr,
                                 GL.sem ? GL.sem + (size_t)pidx * g.Ks : nullptr, a);
    return pass ? (uint64_t)__float_as_uint(y4.w) : ~0ull;
  };
  uint64_t thr = ~0ull;
  int tsh = 0;
  uint32_t stored = 0;
  auto store_pass = [&]() {
    stored = 0;
    each([&](bool valid, uint32_t pidx) {
      float a = 0.0f;
      uint64_t key = ~0ull;
      if (valid) key = eval(pidx, a);
      const bool take = selected(key, thr, tsh);
      const uint64_t mask = __ballot(take);
      const uint32_t slot = stored + (uint32_t)__popcll(mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
      if (take && slot < K) {
        g.cols[(size_t)r * K + slot] = (int32_t)(uint32_t)key;
        g.vals[(size_t)r * K + slot] = a;
      }
      stored += (uint32_t)__popcll(mask);
    });
  };
  store_pass();  // every passing target (thr = ~0, tsh = 0 take all)
  if (stored > K) {
    select_smallest<64>(
        each, lane, K, hist,
        [&](uint32_t pidx) {
          float a;
          return eval(pidx, a);
        },
        thr, tsh);
    store_pass();
  }
  if (lane == 0) g.counts[r] = stored < K ? stored : K;
}

__global__ __launch_bounds__(256) void aniso_assoc_kernel(AnisoArgs g) {
  __shared__ uint32_t s_hist[4][256];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + wave;
  if (r >= g.n) return;  // whole wave; no block barrier below
  const float x[3] = {g.sx[3 * (size_t)r], g.sx[3 * (size_t)r + 1], g.sx[3 * (size_t)r + 2]};
  float M[9];
  row_metric(x, g.Kinv, M);
  RowEllipsoid E;
  float q0 = 0.0f, q1 = 0.0f, q2 = 0.0f;
  if (g.grid) affine3(g.Rq, g.Tq, x[0], x[1], x[2], q0, q1, q2);
  if (g.grid && row_ellipsoid(M, g.d2_cut, g.Rq, q0, q1, q2, E)) aniso_box_row(g, r, lane, x, M, E, s_hist[wave]);
  else aniso_scan_row(g, r, lane, x, M);
}

// CSR compaction: one 64-thread block per row sorts its pairs by target id
// in LDS (rows of at most 1024 pairs) and writes them at row_ptr[r].
__global__ __launch_bounds__(64) void aniso_compact_kernel(const uint32_t* __restrict__ counts,
                                                           const int64_t* __restrict__ row_ptr,
                                                           const int32_t* __restrict__ cols,
                                                           const float* __restrict__ vals, int n, int K,
                                                           int32_t* __restrict__ ocol, float* __restrict__ oval) {
  __shared__ uint64_t s_key[1024];
  const int r = blockIdx.x;
  if (r >= n) return;
  const uint32_t c = counts[r];
  uint32_t n2 = 1;
  while (n2 < c) n2 <<= 1;
  for (uint32_t s = threadIdx.x; s < n2; s += 64)
    s_key[s] = s < c ? ((((uint64_t)(uint32_t)cols[(size_t)r * K + s]) << 32) |
                        (uint64_t)__float_as_uint(vals[(size_t)r * K + s]))
                     : ~0ull;
  __syncthreads();
  block_bitonic_sort(s_key, n2);
  const int64_t base = row_ptr[r];
  for (uint32_t s = threadIdx.x; s < c; s += 64) {
    ocol[base + s] = (int32_t)(uint32_t)(s_key[s] >> 32);
    oval[base + s] = __uint_as_float((uint32_t)s_key[s]);
  }
}

__global__ void widen_counts(const uint32_t* __restrict__ c, int64_t* __restrict__ o, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = (int64_t)c[i];
  if (i == n) o[i] = 0;
}

}  // namespace cvo

using namespace cvo;

namespace {

template <typename T>
T* as(DevBuf& b) {
  return reinterpret_cast<T*>(b.p);
}

bool invert3(const float Kin[9], float out[9]) {
  double k[9];
  for (int i = 0; i < 9; ++i) k[i] = Kin[i];
  const double c00 = k[4] * k[8] - k[5] * k[7], c01 = k[5] * k[6] - k[3] * k[8], c02 = k[3] * k[7] - k[4] * k[6];
  const double det = k[0] * c00 + k[1] * c01 + k[2] * c02;
  if (!(std::isfinite(det)) || det == 0.0) return false;
  const double inv[9] = {c00 / det, (k[2] * k[7] - k[1] * k[8]) / det, (k[1] * k[5] - k[2] * k[4]) / det,
                         c01 / det, (k[0] * k[8] - k[2] * k[6]) / det, (k[2] * k[3] - k[0] * k[5]) / det,
                         c02 / det, (k[1] * k[6] - k[0] * k[7]) / det, (k[0] * k[4] - k[1] * k[3]) / det};
  for (int i = 0; i < 9; ++i) {
    out[i] = (float)inv[i];
    if (!std::isfinite(out[i])) return false;
  }
  return true;
}

}  // namespace

extern "C" cvo_status cvo_compute_association_nonisotropic(cvo_ctx* ctx, const cvo_cloud_view* src,
                                                           const cvo_cloud_view* tgt, const float T[16],
                                                           const float kernel[9], cvo_assoc* out) {
  if (!ctx || !src || !tgt || !T || !kernel || !out) return CVO_ERR_INVALID_ARG;
  std::memset(out, 0, sizeof(*out));
  if (src->n == 0 || tgt->n == 0) return CVO_OK;  // the reference returns without output
  AnisoArgs g{};
  for (int i = 0; i < 9; ++i)
    if (!std::isfinite(kernel[i])) return set_err(ctx, CVO_ERR_INVALID_ARG, "non-finite kernel");
  if (!invert3(kernel, g.Kinv)) return set_err(ctx, CVO_ERR_INVALID_ARG, "singular kernel");
  cvo_dcloud *ds = nullptr, *dt = nullptr;
  cvo_status s = cvo_cloud_upload(ctx, src, &ds);
  if (s == CVO_OK) s = cvo_cloud_upload(ctx, tgt, &dt);
  if (s == CVO_OK) s = validate_pair(ctx, ds, dt);
  auto done = [&](cvo_status r) {
    if (ds) cvo_cloud_free(ctx, ds);
    if (dt) cvo_cloud_free(ctx, dt);
    return r;
  };
  if (s != CVO_OK) return done(s);
  const cvo_params& p = ctx->params;
  hipStream_t st = ctx->stream;
  const int n = ds->n, m = dt->n, K = p.nearest_neighbors_max;
  // the applied target map [R^T, -R^T t] of the pose T (update_tf)
  double R[9], t[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) R[r * 3 + c] = T[r * 4 + c];
    t[r] = T[r * 4 + 3];
  }
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) g.A[i * 3 + j] = (float)R[j * 3 + i];
    g.b[i] = (float)(-(R[0 * 3 + i] * t[0] + R[1 * 3 + i] * t[1] + R[2 * 3 + i] * t[2]));
  }
  // a mapped point y' lies at R y' + t in the target's own frame
  for (int i = 0; i < 9; ++i) g.Rq[i] = (float)R[i];
  for (int i = 0; i < 3; ++i) g.Tq[i] = (float)t[i];
  g.sx = ds->xyz;
  g.sf = ds->feat;
  g.ss = ds->sem;
  g.tx = dt->xyz;
  g.tf = dt->feat;
  g.ts = dt->sem;
  g.n = n;
  g.m = m;
  g.F = ds->F;
  g.Ks = ds->K;
  g.K = K;
  g.use_geo = p.is_using_geometry;
  g.use_col = p.is_using_intensity && ds->F > 0 && ds->feat && dt->feat;
  g.use_sem = p.is_using_semantics && ds->K > 0 && ds->sem && dt->sem;
  g.sigma2 = p.sigma * p.sigma;
  g.c_sigma2 = p.c_sigma * p.c_sigma;
  g.s_sigma2 = p.s_sigma * p.s_sigma;
  g.c_ell2 = p.c_ell * p.c_ell;
  g.s_ell2 = p.s_ell * p.s_ell;
  g.sp = p.sp_thres;
  g.d2_c_thres = g.use_col ? (float)(-2.0 * g.c_ell2 * std::log((double)(p.sp_thres / g.c_sigma2))) : 1.0f;
  g.d2_s_thres = g.use_sem ? (float)(-2.0 * g.s_ell2 * std::log((double)(p.sp_thres / g.s_sigma2))) : 1.0f;
  // a <= sigma^2 exp(-d2/2) c_sigma^2 s_sigma^2 (enabled factors): above this
  // d2 no pair reaches sp_thres; the margin keeps the prune conservative
  double amp = (double)g.sigma2 * (g.use_col ? (double)g.c_sigma2 : 1.0) * (g.use_sem ? (double)g.s_sigma2 : 1.0);
  const double cut = amp > (double)p.sp_thres ? 2.0 * std::log(amp / (double)p.sp_thres) : 0.0;
  g.d2_cut = (float)(cut * (1.0 + 1e-4) + 1e-6);
  // the target grid: with M = U K^-1 U^T and U's columns of lengths >= 1
  // (u, w orthogonal to z and of length >= sqrt 2), an ellipsoid's extent is
  // at most sqrt(cut lambda_max(K)) <= sqrt(cut ||K||_F); cells of half that
  GridHost gh;
  g.grid = nullptr;
  if (g.use_geo) {
    double kf = 0.0;
    for (int i = 0; i < 9; ++i) kf += (double)kernel[i] * (double)kernel[i];
    const double hc = 0.5 * std::sqrt((double)g.d2_cut * std::sqrt(kf));
    if (std::isfinite(hc) && hc > 0.0) {
      s = build_grid(ctx, dt, std::vector<float>{(float)hc}, gh);
      if (s != CVO_OK) return done(s);
      g.grid = gh.dev_ptr;
    }
  }
  DevBuf* B = ctx->b_ba;  // association scratch: pair store, counts, row pointers, CSR
  const size_t nk = (size_t)n * K;
  s = ensure(ctx, B[0], sizeof(uint32_t) * (size_t)n);
  if (s == CVO_OK) s = ensure(ctx, B[1], (sizeof(int32_t) + sizeof(float)) * std::max<size_t>(nk, 1));
  if (s == CVO_OK) s = ensure(ctx, B[2], sizeof(int64_t) * (size_t)(n + 1) * 2);
  if (s != CVO_OK) return done(s);
  g.counts = as<uint32_t>(B[0]);
  g.cols = as<int32_t>(B[1]);
  g.vals = reinterpret_cast<float*>(g.cols + std::max<size_t>(nk, 1));
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (ctx->profiling) {
    s = ensure_events(ctx, 2);
    if (s != CVO_OK) return done(s);
    ev[0] = ctx->ev_pool[0];
    ev[1] = ctx->ev_pool[1];
    const hipError_t er = hipEventRecord(ev[0], st);
    if (er != hipSuccess) return done(hip_check(ctx, er, "event record"));
  }
  hipLaunchKernelGGL(aniso_assoc_kernel, dim3((n + 3) / 4), dim3(256), 0, st, g);
  if (ev[1]) {
    const hipError_t er = hipEventRecord(ev[1], st);
    if (er != hipSuccess) return done(hip_check(ctx, er, "event record"));
  }
  int64_t* cnt = as<int64_t>(B[2]);
  int64_t* rp = cnt + (n + 1);
  hipLaunchKernelGGL(widen_counts, dim3((n + 1 + 255) / 256), dim3(256), 0, st, g.counts, cnt, n);
  size_t tmp = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt, rp, n + 1, st);
  if (e != hipSuccess) return done(hip_check(ctx, e, "scan size"));
  s = ensure(ctx, ctx->b_cub, tmp);
  if (s != CVO_OK) return done(s);
  e = hipcub::DeviceScan::ExclusiveSum(ctx->b_cub.p, tmp, cnt, rp, n + 1, st);
  if (e != hipSuccess) return done(hip_check(ctx, e, "scan"));
  int64_t nnz = 0;
  e = hipMemcpyAsync(&nnz, rp + n, sizeof(int64_t), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return done(hip_check(ctx, e, "nnz"));
  if (ev[1]) {
    // algorithmic bytes: each row's and target's position, features and
    // classes once, the stored pairs, the row counts
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    if (e != hipSuccess) return done(hip_check(ctx, e, "event time"));
    const double per_pt = 12.0 + (g.use_col ? 4.0 * g.F : 0.0) + (g.use_sem ? 4.0 * g.Ks : 0.0);
    ctx->stats.ms[CVO_KSTAT_ANISO] += ms;
    ctx->stats.launches[CVO_KSTAT_ANISO] += 1;
    ctx->stats.bytes[CVO_KSTAT_ANISO] += per_pt * (double)(n + m) + 8.0 * (double)nnz + 4.0 * n;
    ctx->stats.pairs_stored += (double)nnz;
  }
  s = ensure(ctx, B[3], (sizeof(int32_t) + sizeof(float)) * (size_t)std::max<int64_t>(nnz, 1));
  if (s != CVO_OK) return done(s);
  int32_t* ocol = as<int32_t>(B[3]);
  float* oval = reinterpret_cast<float*>(ocol + std::max<int64_t>(nnz, 1));
  hipLaunchKernelGGL(aniso_compact_kernel, dim3(n), dim3(64), 0, st, g.counts, rp, g.cols, g.vals, n, K, ocol, oval);
  e = hipGetLastError();
  if (e != hipSuccess) return done(hip_check(ctx, e, "anisotropic association"));
  out->rows = n;
  out->cols = m;
  out->nnz = nnz;
  out->row_ptr = static_cast<int64_t*>(std::malloc(sizeof(int64_t) * (n + 1)));
  out->col = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
  out->val = static_cast<float*>(std::malloc(sizeof(float) * std::max<int64_t>(nnz, 1)));
  if (!out->row_ptr || !out->col || !out->val) return done(set_err(ctx, CVO_ERR_OOM, "host allocation of association"));
  e = hipMemcpyAsync(out->row_ptr, rp, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && nnz > 0) e = hipMemcpyAsync(out->col, ocol, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && nnz > 0) e = hipMemcpyAsync(out->val, oval, sizeof(float) * nnz, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return done(hip_check(ctx, e, "association download"));
  return done(CVO_OK);
}
