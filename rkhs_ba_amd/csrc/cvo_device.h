// Device helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "cvo_internal.h"

namespace cvo {

// Load through a pointer known to address global memory.  Pointers read from
// device structures are generic to the compiler, which then emits flat loads;
// those also count against the LDS wait counter, so a wait for the row's LDS
// data would wait for every gather in flight as well.
template <class T>
__device__ __forceinline__ T ld_global(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
__device__ __forceinline__ float4 ld_global(const float4* p) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  const v4f v = *(const __attribute__((address_space(1))) v4f*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 ld_global(const uint2* p) {
  typedef unsigned v2u __attribute__((ext_vector_type(2)));
  const v2u v = *(const __attribute__((address_space(1))) v2u*)p;
  return make_uint2(v.x, v.y);
}
// Stores through a pointer known to address global memory (a flat store would
// also hold the LDS wait counter until the store is acknowledged).
template <class T>
__device__ __forceinline__ void st_global(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
}
__device__ __forceinline__ void st_global(uint2* p, uint2 v) {
  typedef unsigned v2u __attribute__((ext_vector_type(2)));
  const v2u t = {v.x, v.y};
  *(__attribute__((address_space(1))) v2u*)p = t;
}
__device__ __forceinline__ void st_global(float4* p, float4 v) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  const v4f t = {v.x, v.y, v.z, v.w};
  *(__attribute__((address_space(1))) v4f*)p = t;
}

__device__ __forceinline__ void affine3(const float* A, const float* b, float y0, float y1, float y2, float& o0,
                                        float& o1, float& o2) {
  o0 = fmaf(A[2], y2, fmaf(A[1], y1, fmaf(A[0], y0, b[0])));
  o1 = fmaf(A[5], y2, fmaf(A[4], y1, fmaf(A[3], y0, b[1])));
  o2 = fmaf(A[8], y2, fmaf(A[7], y1, fmaf(A[6], y0, b[2])));
}

// compute_range_ell (CvoGPU.cu:90-94); distance with the canonical FMA order.
__device__ __forceinline__ float range_ell(float ell, float x0, float x1, float x2) {
  float dist = sqrtf(fmaf(x2, x2, fmaf(x1, x1, x0 * x0)));
  return (float)(((double)dist / 500.0 + 1.0) * (double)ell);
}

// exp(s [w; v]) for a unit twist, FP64, exact small-angle series.
__device__ inline void se3_exp(const double xi[6], double s, double R[9], double t[3]) {
  const double ph[3] = {xi[0] * s, xi[1] * s, xi[2] * s};
  const double u[3] = {xi[3] * s, xi[4] * s, xi[5] * s};
  const double th2 = ph[0] * ph[0] + ph[1] * ph[1] + ph[2] * ph[2];
  const double th = sqrt(th2);
  double ca, cb, cc;
  if (th < 1e-3) {
    ca = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
    cb = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
    cc = 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0;
  } else {
    double sn, cs;
    sincos(th, &sn, &cs);
    ca = sn / th;
    cb = (1.0 - cs) / th2;
    cc = (th - sn) / (th2 * th);
  }
  const double W[9] = {0, -ph[2], ph[1], ph[2], 0, -ph[0], -ph[1], ph[0], 0};
  double V[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double w2 = W[r * 3 + 0] * W[0 * 3 + c] + W[r * 3 + 1] * W[1 * 3 + c] + W[r * 3 + 2] * W[2 * 3 + c];
      double I = (r == c) ? 1.0 : 0.0;
      R[r * 3 + c] = I + ca * W[r * 3 + c] + cb * w2;
      V[r * 3 + c] = I + cb * W[r * 3 + c] + cc * w2;
    }
  for (int r = 0; r < 3; ++r) t[r] = V[r * 3 + 0] * u[0] + V[r * 3 + 1] * u[1] + V[r * 3 + 2] * u[2];
}

// ||log(T)|| of a rigid transform, from the rotation's unit quaternion
// (the Sophus SE3d log norm of dist_se3_cpu, src/cvo/CvoGPU.cpp:28-33).
__device__ inline double se3_log_norm(const double R[9], const double t[3]) {
  double qw, qx, qy, qz;
  const double tr = R[0] + R[4] + R[8];
  if (tr > 0) {
    double s = sqrt(tr + 1.0) * 2.0;
    qw = 0.25 * s; qx = (R[7] - R[5]) / s; qy = (R[2] - R[6]) / s; qz = (R[3] - R[1]) / s;
  } else if (R[0] > R[4] && R[0] > R[8]) {
    double s = sqrt(1.0 + R[0] - R[4] - R[8]) * 2.0;
    qw = (R[7] - R[5]) / s; qx = 0.25 * s; qy = (R[1] + R[3]) / s; qz = (R[2] + R[6]) / s;
  } else if (R[4] > R[8]) {
    double s = sqrt(1.0 + R[4] - R[0] - R[8]) * 2.0;
    qw = (R[2] - R[6]) / s; qx = (R[1] + R[3]) / s; qy = 0.25 * s; qz = (R[5] + R[7]) / s;
  } else {
    double s = sqrt(1.0 + R[8] - R[0] - R[4]) * 2.0;
    qw = (R[3] - R[1]) / s; qx = (R[2] + R[6]) / s; qy = (R[5] + R[7]) / s; qz = 0.25 * s;
  }
  const double qn = sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
  qw /= qn; qx /= qn; qy /= qn; qz /= qn;
  const double n2 = qx * qx + qy * qy + qz * qz, n = sqrt(n2);
  double f;
  if (n2 < 1e-20) f = 2.0 / qw - 2.0 * n2 / (qw * qw * qw);
  else if (fabs(qw) < 1e-10) f = (qw > 0 ? M_PI : -M_PI) / n;
  else f = 2.0 * atan(n / qw) / n;
  const double w[3] = {f * qx, f * qy, f * qz};
  const double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  const double k2 = th < 1e-6 ? 1.0 / 12.0 : (1.0 - th * cos(0.5 * th) / (2.0 * sin(0.5 * th))) / (th * th);
  double s2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  for (int r = 0; r < 3; ++r) {
    double acc = 0.0;
    for (int c = 0; c < 3; ++c) {
      double w2 = W[r * 3 + 0] * W[0 * 3 + c] + W[r * 3 + 1] * W[1 * 3 + c] + W[r * 3 + 2] * W[2 * 3 + c];
      double vi = ((r == c) ? 1.0 : 0.0) - 0.5 * W[r * 3 + c] + k2 * w2;
      acc += vi * t[c];
    }
    s2 += acc * acc;
  }
  return sqrt(s2);
}

// Exp_SEK3 of the reference in FP32 (LieGroup.cpp:244-275), including its
// small-angle branch (J = I below 1e-6, so a pure translation step moves v
// unscaled): out = [R | J v] row-major 3x4.  Used only when the library is
// asked to reproduce the reference's float pose arithmetic.
__device__ inline void exp_sek3_f32(const float xi[6], float dt, float out[12]) {
  const float w0 = xi[0], w1 = xi[1], w2 = xi[2];
  const float theta = sqrtf(w0 * w0 + w1 * w1 + w2 * w2);
  float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, J[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (!(theta < 1e-6f)) {
    const float A[9] = {0, -w2, w1, w2, 0, -w0, -w1, w0, 0};
    float A2[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) A2[r * 3 + c] = A[r * 3 + 0] * A[0 * 3 + c] + A[r * 3 + 1] * A[1 * 3 + c] + A[r * 3 + 2] * A[2 * 3 + c];
    const float theta2 = theta * theta;
    const float st = sinf(dt * theta), ct = cosf(dt * theta);
    const float omc = (1 - ct) / theta2;
    const float k3 = (dt * theta - st) / (theta2 * theta);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        const float I = r == c ? 1.0f : 0.0f;
        R[r * 3 + c] = I + (st / theta) * A[r * 3 + c] + omc * A2[r * 3 + c];
        J[r * 3 + c] = dt * I + omc * A[r * 3 + c] + k3 * A2[r * 3 + c];
      }
  }
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) out[r * 4 + c] = R[r * 3 + c];
    out[r * 4 + 3] = J[r * 3 + 0] * xi[3] + J[r * 3 + 1] * xi[4] + J[r * 3 + 2] * xi[5];
  }
}

__device__ inline int choose_level(const GridDev* grid, float r_need) {
  int L = 0;
  for (int l = 0; l < grid->n_levels; ++l)
    if (grid->lv[l].h >= r_need) L = l;
  return L;
}

// Ascending bitonic sort of key[0, n2) in LDS by the threads of one block
// (n2 a power of two; pad with ~0).  Every thread of the block calls it.
__device__ inline void block_bitonic_sort(uint64_t* key, uint32_t n2) {
  for (uint32_t k = 2; k <= n2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const bool up = (i & k) == 0;
          const uint64_t a = key[i], b = key[ixj];
          if ((a > b) == up) {
            key[i] = b;
            key[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

}  // namespace cvo
