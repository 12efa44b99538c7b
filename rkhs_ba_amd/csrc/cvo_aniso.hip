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
// One wave per source row scans the targets in index order, 64 at a time;
// a ballot gives each passing lane its slot.  Pairs that cannot reach
// sp_thres (d2 beyond 2 ln(amplitude / sp_thres)) skip the exponentials.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <math.h>
#include <stdint.h>

#include <cmath>
#include <cstdlib>
#include <cstring>

#include "cvo_internal.h"
#include "cvo_device.h"

namespace cvo {

struct AnisoArgs {
  const float* sx;
  const float* sf;
  const float* ss;
  const float* tx;
  const float* tf;
  const float* ts;
  int n, m, F, Ks, K;
  float A[9], b[3];  // target map y' = A y + b
  float Kinv[9];     // row-major inverse kernel
  float sigma2, c_sigma2, s_sigma2, c_ell2, s_ell2, sp, d2_c_thres, d2_s_thres;
  float d2_cut;  // beyond it the product kernel cannot exceed sp_thres
  int use_geo, use_col, use_sem;
  uint32_t* counts;  // n
  int32_t* cols;     // n * K
  float* vals;       // n * K
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

__global__ __launch_bounds__(256) void aniso_assoc_kernel(AnisoArgs g) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + wave;
  if (r >= g.n) return;  // whole wave; no block barrier below
  const float x[3] = {g.sx[3 * (size_t)r], g.sx[3 * (size_t)r + 1], g.sx[3 * (size_t)r + 2]};
  float M[9];
  row_metric(x, g.Kinv, M);
  const uint32_t K = (uint32_t)g.K;
  uint32_t stored = 0;
  for (int base = 0; base < g.m && stored < K; base += 64) {
    const int j = base + lane;
    bool pass = false;
    float a = 0.0f;
    if (j < g.m) {
      float y[3];
      affine3(g.A, g.b, g.tx[3 * (size_t)j], g.tx[3 * (size_t)j + 1], g.tx[3 * (size_t)j + 2], y[0], y[1], y[2]);
      float k = 1.0f, ck = 1.0f, sk = 1.0f;
      bool ok = true;
      if (g.use_geo) {
        const float d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
        // (d^T M) d
        const float t0 = fmaf(d[2], M[6], fmaf(d[1], M[3], d[0] * M[0]));
        const float t1 = fmaf(d[2], M[7], fmaf(d[1], M[4], d[0] * M[1]));
        const float t2 = fmaf(d[2], M[8], fmaf(d[1], M[5], d[0] * M[2]));
        const float d2 = fmaf(t2, d[2], fmaf(t1, d[1], t0 * d[0]));
        if (d2 > g.d2_cut) ok = false;
        else k = (float)((double)g.sigma2 * exp(-(double)d2 / 2.0));
      }
      if (ok && g.use_col) {
        float d2c = 0.0f;
        for (int f = 0; f < g.F; ++f) {
          const float t = g.sf[(size_t)r * g.F + f] - g.tf[(size_t)j * g.F + f];
          d2c = fmaf(t, t, d2c);
        }
        if (d2c < g.d2_c_thres) ck = (float)((double)g.c_sigma2 * exp(-(double)d2c / (2.0 * (double)g.c_ell2)));
        else ok = false;
      }
      if (ok && g.use_sem) {
        float d2s = 0.0f;
        for (int c = 0; c < g.Ks; ++c) {
          const float t = g.ss[(size_t)r * g.Ks + c] - g.ts[(size_t)j * g.Ks + c];
          d2s = fmaf(t, t, d2s);
        }
        if (d2s < g.d2_s_thres) sk = (float)((double)g.s_sigma2 * exp(-(double)d2s / (2.0 * (double)g.s_ell2)));
        else ok = false;
      }
      if (ok) {
        a = ck * k * sk * 1.0f;
        pass = a > g.sp;
      }
    }
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

// CSR compaction: one wave per row copies its pairs to row_ptr[r].
__global__ __launch_bounds__(256) void aniso_compact_kernel(const uint32_t* __restrict__ counts,
                                                            const int64_t* __restrict__ row_ptr,
                                                            const int32_t* __restrict__ cols,
                                                            const float* __restrict__ vals, int n, int K,
                                                            int32_t* __restrict__ ocol, float* __restrict__ oval) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + wave;
  if (r >= n) return;
  const uint32_t c = counts[r];
  const int64_t base = row_ptr[r];
  for (uint32_t s = lane; s < c; s += 64) {
    ocol[base + s] = cols[(size_t)r * K + s];
    oval[base + s] = vals[(size_t)r * K + s];
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
  DevBuf* B = ctx->b_ba;  // association scratch: pair store, counts, row pointers, CSR
  const size_t nk = (size_t)n * K;
  s = ensure(ctx, B[0], sizeof(uint32_t) * (size_t)n);
  if (s == CVO_OK) s = ensure(ctx, B[1], (sizeof(int32_t) + sizeof(float)) * std::max<size_t>(nk, 1));
  if (s == CVO_OK) s = ensure(ctx, B[2], sizeof(int64_t) * (size_t)(n + 1) * 2);
  if (s != CVO_OK) return done(s);
  g.counts = as<uint32_t>(B[0]);
  g.cols = as<int32_t>(B[1]);
  g.vals = reinterpret_cast<float*>(g.cols + std::max<size_t>(nk, 1));
  hipLaunchKernelGGL(aniso_assoc_kernel, dim3((n + 3) / 4), dim3(256), 0, st, g);
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
  s = ensure(ctx, B[3], (sizeof(int32_t) + sizeof(float)) * (size_t)std::max<int64_t>(nnz, 1));
  if (s != CVO_OK) return done(s);
  int32_t* ocol = as<int32_t>(B[3]);
  float* oval = reinterpret_cast<float*>(ocol + std::max<int64_t>(nnz, 1));
  hipLaunchKernelGGL(aniso_compact_kernel, dim3((n + 3) / 4), dim3(256), 0, st, g.counts, rp, g.cols, g.vals, n, K,
                     ocol, oval);
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
