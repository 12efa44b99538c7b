// CPU restatement (oracle) of the reference CVO two-view path and RKHS-BA window.
//
// TEST INFRASTRUCTURE ONLY: the checker of the HIP library.  Loaded by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the product.
//
// What pins each function (independent of this file's own formulas):
//   orc_assoc            brute force (brute=1) vs grid (brute=0) set equality;
//                        tests/golden vectors of a few hand-sized pairs computed
//                        with numpy from the kernel definition; gate/threshold
//                        closed forms.
//   orc_flow             numpy FP64 closed form on tiny inputs (golden).
//   orc_step_coeffs      numpy polynomial composition of the 4th-order path
//                        expansion (golden) -- see tests/golden/make_golden.py.
//   orc_cubic_roots      numpy.roots (companion-matrix eigenvalues) golden.
//   orc_exp_*, orc_log_* scipy.linalg.expm / logm golden.
//   orc_indicator_sequence  scripted sequences with hand-derived decisions.
//   orc_align            convergence to the generating pose of seeded scenes.
//   BA functions         finite differences, scipy expm, numpy dense assembly.
//
// Build: oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off).

#include "cvo_oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <complex>
#include <cstdint>
#include <cstring>
#include <deque>
#include <limits>
#include <vector>

namespace {

constexpr int kKnnK = 512;  // KDTREE_K_SIZE (thirdparty/cugicp/cukdtree/cukdtree.h:12)

// ---------------------------------------------------------------------------
// Canonical float helpers shared (by definition, not by code) with the HIP path.
// ---------------------------------------------------------------------------

struct Pose64 {
  double R[9];  // row-major
  double T[3];
};

Pose64 pose_from16(const float T[16]) {
  Pose64 P;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) P.R[r * 3 + c] = T[r * 4 + c];
    P.T[r] = T[r * 4 + 3];
  }
  return P;
}

// update_tf (CvoGPU.cu:98-130): the target is mapped by [R^T, -R^T T].
// Canonical: A = float(R^T), b = float(-(R^T T)) evaluated in double.
void applied_transform(const Pose64& P, float A[9], float b[3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i * 3 + j] = static_cast<float>(P.R[j * 3 + i]);
  for (int i = 0; i < 3; ++i)
    b[i] = static_cast<float>(-(P.R[0 * 3 + i] * P.T[0] + P.R[1 * 3 + i] * P.T[1] + P.R[2 * 3 + i] * P.T[2]));
}

// transform_point_R_T (CvoGPU_impl.cu:32-83) with the canonical FMA order.
inline void xform_f(const float A[9], const float b[3], const float* y, float* o) {
  for (int i = 0; i < 3; ++i)
    o[i] = std::fmaf(A[i * 3 + 2], y[2], std::fmaf(A[i * 3 + 1], y[1], std::fmaf(A[i * 3 + 0], y[0], b[i])));
}

// squared_dist(p_b, p_a) (gpu_utils.cuh:72-78): diff = b - a.
inline float sqdist3_f(const float* b, const float* a) {
  float dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
  return std::fmaf(dz, dz, std::fmaf(dy, dy, dx * dx));
}

// squared_dist<float>(a, b, dim) (gpu_utils.cuh:33-41): sequential accumulation.
inline float sqdist_n_f(const float* a, const float* b, int n) {
  float r = 0.0f;
  for (int i = 0; i < n; ++i) {
    float t = a[i] - b[i];
    r = std::fmaf(t, t, r);
  }
  return r;
}

inline float dot_n_f(const float* a, const float* b, int n) {
  float r = 0.0f;
  for (int i = 0; i < n; ++i) r = std::fmaf(a[i], b[i], r);
  return r;
}

struct Consts {
  float sigma2, c2, c_sigma2, s_ell, s_sigma2;
  float lg, lc, ls;  // logf(sp/..)
  float d2_c_thres, d2_s_thres;
  int geo, col, sem, gtype;
  float sp;
};

Consts make_consts(const cvo_params* p) {
  Consts k;
  k.sigma2 = p->sigma * p->sigma;
  k.c2 = p->c_ell * p->c_ell;
  k.c_sigma2 = p->c_sigma * p->c_sigma;
  k.s_ell = p->s_ell;
  k.s_sigma2 = p->s_sigma * p->s_sigma;
  k.lg = logf(p->sp_thres / k.sigma2);
  k.lc = logf(p->sp_thres / k.c_sigma2);
  k.ls = logf(p->sp_thres / k.s_sigma2);
  k.geo = p->is_using_geometry;
  k.col = p->is_using_intensity;
  k.sem = p->is_using_semantics;
  k.gtype = p->is_using_geometric_type;
  k.sp = p->sp_thres;
  // fill_in_A_mat_gpu thresholds (CvoGPU.cu:981-987)
  k.d2_c_thres = k.col ? static_cast<float>(-2.0 * (double)k.c2 * (double)k.lc) : 1.0f;
  k.d2_s_thres = k.sem ? static_cast<float>(-2.0 * (double)k.s_ell * (double)k.s_ell * (double)k.ls) : 1.0f;
  return k;
}

// compute_range_ell (CvoGPU.cu:90-94), applied unconditionally at :978-979.
inline float range_ell(float ell, const float* x) {
  float dist = sqrtf(std::fmaf(x[2], x[2], std::fmaf(x[1], x[1], x[0] * x[0])));
  return static_cast<float>(((double)dist / 500.0 + 1.0) * (double)ell);
}

inline float d2_geo_thres(const Consts& k, float l) {
  return k.geo ? static_cast<float>(-2.0 * (double)l * (double)l * (double)k.lg) : 1.0f;
}

struct CloudRef {
  int n, F, K;
  const float *xyz, *feat, *sem, *geo;
};

CloudRef ref(const cvo_cloud_view* v) {
  CloudRef c{v->n, v->feat_dim, v->num_classes, v->xyz, v->feat, v->sem, v->geo_type};
  return c;
}

// Mirror evaluation of one pair (CvoGPU.cu:997-1076).  Returns false when a gate
// rejects; a = product kernel before the sp_thres test.  Gate margins (relative
// distance to each threshold) are reported for the near-cutoff exemption.
struct PairEval {
  bool ok;
  int gate_stage;  // 0 geo-type, 1 geo, 2 color, 3 sem, 4 product
  float a;
  double margin;
};

// gmin: geometric-type similarity floor, 0.01 (fill_in_A_mat_gpu, CvoGPU.cu:1015)
// or 0.001 (fill_in_A_mat_cukdtree, :898).
inline PairEval eval_pair_mirror(const Consts& k, const CloudRef& S, const CloudRef& Tg, int i, int j,
                                 const float* xi, const float* yp, float l, float d2_thres, float gmin = 0.01f) {
  PairEval r{false, 0, 0.0f, 1e30};
  float geo_sim = 1.0f, kk = 1.0f, ck = 1.0f, sk = 1.0f;
  if (k.gtype) {
    const float* ga = S.geo + 2 * (size_t)i;
    const float* gb = Tg.geo + 2 * (size_t)j;
    float n2a = dot_n_f(ga, ga, 2), n2b = dot_n_f(gb, gb, 2), dab = dot_n_f(ga, gb, 2);
    geo_sim = dab * dab / (n2a * n2b);
    r.margin = std::min(r.margin, std::fabs((double)geo_sim - (double)gmin) / (double)gmin);
    if (geo_sim < gmin) { r.gate_stage = 0; return r; }
  }
  if (k.geo) {
    float d2 = sqdist3_f(yp, xi);
    r.margin = std::min(r.margin, std::fabs((double)d2 - (double)d2_thres) / std::max(1e-30, (double)d2_thres));
    if (!(d2 < d2_thres)) { r.gate_stage = 1; return r; }
    kk = static_cast<float>((double)k.sigma2 * std::exp((double)(-d2) / (2.0 * (double)l * (double)l)));
  }
  if (k.col) {
    float d2c = sqdist_n_f(S.feat + (size_t)i * S.F, Tg.feat + (size_t)j * Tg.F, S.F);
    r.margin = std::min(r.margin, std::fabs((double)d2c - (double)k.d2_c_thres) / std::max(1e-30, (double)k.d2_c_thres));
    if (!(d2c < k.d2_c_thres)) { r.gate_stage = 2; return r; }
    ck = static_cast<float>((double)k.c_sigma2 * std::exp((double)(-d2c) / (2.0 * (double)k.c2)));
  }
  if (k.sem) {
    float d2s = sqdist_n_f(S.sem + (size_t)i * S.K, Tg.sem + (size_t)j * Tg.K, S.K);
    r.margin = std::min(r.margin, std::fabs((double)d2s - (double)k.d2_s_thres) / std::max(1e-30, (double)k.d2_s_thres));
    if (!(d2s < k.d2_s_thres)) { r.gate_stage = 3; return r; }
    sk = static_cast<float>((double)k.s_sigma2 * std::exp((double)(-d2s) / (2.0 * (double)k.s_ell * (double)k.s_ell)));
  }
  r.a = ck * kk * sk * geo_sim;
  r.gate_stage = 4;
  r.ok = true;
  return r;
}

// FP64 evaluation of one pair (same algorithm in double).
inline PairEval eval_pair_fp64(const Consts& k, const cvo_params* p, const CloudRef& S, const CloudRef& Tg,
                               int i, int j, const double* xi, const double* yp, double l, double gmin = 0.01) {
  PairEval r{false, 0, 0.0f, 1e30};
  double geo_sim = 1, kk = 1, ck = 1, sk = 1;
  if (k.gtype) {
    const float* ga = S.geo + 2 * (size_t)i;
    const float* gb = Tg.geo + 2 * (size_t)j;
    double n2a = (double)ga[0] * ga[0] + (double)ga[1] * ga[1];
    double n2b = (double)gb[0] * gb[0] + (double)gb[1] * gb[1];
    double dab = (double)ga[0] * gb[0] + (double)ga[1] * gb[1];
    geo_sim = dab * dab / (n2a * n2b);
    if (geo_sim < gmin) return r;
  }
  double sigma2 = (double)p->sigma * p->sigma;
  if (k.geo) {
    double dx = yp[0] - xi[0], dy = yp[1] - xi[1], dz = yp[2] - xi[2];
    double d2 = dx * dx + dy * dy + dz * dz;
    double thr = -2.0 * l * l * std::log((double)p->sp_thres / sigma2);
    if (!(d2 < thr)) return r;
    kk = sigma2 * std::exp(-d2 / (2.0 * l * l));
  }
  if (k.col) {
    double c2 = (double)p->c_ell * p->c_ell, cs2 = (double)p->c_sigma * p->c_sigma;
    double d2c = 0;
    for (int f = 0; f < S.F; ++f) {
      double t = (double)S.feat[(size_t)i * S.F + f] - Tg.feat[(size_t)j * Tg.F + f];
      d2c += t * t;
    }
    if (!(d2c < -2.0 * c2 * std::log((double)p->sp_thres / cs2))) return r;
    ck = cs2 * std::exp(-d2c / (2.0 * c2));
  }
  if (k.sem) {
    double s2 = (double)p->s_ell * p->s_ell, ss2 = (double)p->s_sigma * p->s_sigma;
    double d2s = 0;
    for (int f = 0; f < S.K; ++f) {
      double t = (double)S.sem[(size_t)i * S.K + f] - Tg.sem[(size_t)j * Tg.K + f];
      d2s += t * t;
    }
    if (!(d2s < -2.0 * s2 * std::log((double)p->sp_thres / ss2))) return r;
    sk = ss2 * std::exp(-d2s / (2.0 * s2));
  }
  double a = ck * kk * sk * geo_sim;
  r.a = static_cast<float>(a);
  r.margin = a;  // exact value carried in margin for the fp64 path
  r.ok = true;
  r.gate_stage = 4;
  return r;
}

// Uniform grid over the transformed target (the oracle's own radius search; it
// returns candidate supersets, the exact gates decide).
struct Grid {
  double h = 1.0;
  std::vector<std::pair<uint64_t, int>> keyed;  // (cell key, index) sorted
  static uint64_t key(int64_t cx, int64_t cy, int64_t cz) {
    const int64_t off = 1 << 20;
    return ((uint64_t)(cx + off) << 42) | ((uint64_t)(cy + off) << 21) | (uint64_t)(cz + off);
  }
  void build(const float* pts, int m, double cell) {
    h = cell;
    keyed.resize(m);
    for (int j = 0; j < m; ++j) {
      int64_t cx = (int64_t)std::floor(pts[3 * j] / h), cy = (int64_t)std::floor(pts[3 * j + 1] / h),
              cz = (int64_t)std::floor(pts[3 * j + 2] / h);
      keyed[j] = {key(cx, cy, cz), j};
    }
    std::sort(keyed.begin(), keyed.end());
  }
  template <class F>
  void query(const float* x, double r, F&& f) const {
    int64_t lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = (int64_t)std::floor((x[d] - r) / h);
      hi[d] = (int64_t)std::floor((x[d] + r) / h);
    }
    for (int64_t cx = lo[0]; cx <= hi[0]; ++cx)
      for (int64_t cy = lo[1]; cy <= hi[1]; ++cy)
        for (int64_t cz = lo[2]; cz <= hi[2]; ++cz) {
          uint64_t kk = key(cx, cy, cz);
          auto it = std::lower_bound(keyed.begin(), keyed.end(), std::make_pair(kk, -1));
          for (; it != keyed.end() && it->first == kk; ++it) f(it->second);
        }
  }
};

struct AssocOut {
  std::vector<int32_t> counts, cols;
  std::vector<float> vals;
  orc_counters cnt{};
};

// fill_in_A_mat_gpu semantics (CvoGPU.cu:951-1099).  yprime: the transformed
// target in float (mirror) -- for precision FP64 the double transform is used.
void assoc_core(const cvo_params* p, const CloudRef& S, const CloudRef& Tg, const Pose64& P, float ell,
                int K, int precision, int brute, AssocOut& out) {
  const Consts k = make_consts(p);
  const int N = S.n, M = Tg.n;
  out.counts.assign(N, 0);
  out.cols.assign((size_t)N * K, -1);
  out.vals.assign((size_t)N * K, 0.0f);
  std::vector<float> yp((size_t)M * 3);
  std::vector<double> yp64;
  float A[9], b[3];
  applied_transform(P, A, b);
  for (int j = 0; j < M; ++j) xform_f(A, b, Tg.xyz + 3 * (size_t)j, &yp[3 * (size_t)j]);
  if (precision == ORC_FP64) {
    yp64.resize((size_t)M * 3);
    for (int j = 0; j < M; ++j) {
      const float* y = Tg.xyz + 3 * (size_t)j;
      for (int i = 0; i < 3; ++i) {
        double s = -(P.R[0 * 3 + i] * P.T[0] + P.R[1 * 3 + i] * P.T[1] + P.R[2 * 3 + i] * P.T[2]);
        for (int c = 0; c < 3; ++c) s += P.R[c * 3 + i] * (double)y[c];
        yp64[3 * (size_t)j + i] = s;
      }
    }
  }
  bool use_grid = !brute && k.geo;
  Grid grid;
  double rmax = 0;
  if (use_grid) {
    for (int i = 0; i < N; ++i) {
      float l = range_ell(ell, S.xyz + 3 * (size_t)i);
      rmax = std::max(rmax, std::sqrt(std::max(0.0, (double)d2_geo_thres(k, l))));
    }
    rmax = rmax * 1.001 + 1e-6;
    // the grid holds the float transformed points for both precisions (it only
    // proposes candidates; the exact gates decide)
    grid.build(yp.data(), M, std::max(rmax, 1e-6));
  }
  int64_t c_geo = 0, c_col = 0, c_sem = 0, c_pass = 0, c_store = 0;
  int32_t mrow = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : c_geo, c_col, c_sem, c_pass, c_store) reduction(max : mrow)
  for (int i = 0; i < N; ++i) {
    const float* xi = S.xyz + 3 * (size_t)i;
    float l = range_ell(ell, xi);
    float d2t = d2_geo_thres(k, l);
    double l64 = ((double)std::sqrt((double)xi[0] * xi[0] + (double)xi[1] * xi[1] + (double)xi[2] * xi[2]) / 500.0 + 1.0) * (double)ell;
    double xi64[3] = {xi[0], xi[1], xi[2]};
    std::vector<int> cand;
    if (use_grid) {
      double r = std::sqrt(std::max(0.0, (double)d2t)) * 1.001 + 1e-5 * (1.0 + std::fabs(xi[0]) + std::fabs(xi[1]) + std::fabs(xi[2]));
      grid.query(xi, r, [&](int j) { cand.push_back(j); });
      std::sort(cand.begin(), cand.end());
    }
    int stored = 0;
    auto visit = [&](int j) {
      PairEval e;
      if (precision == ORC_FP64) e = eval_pair_fp64(k, p, S, Tg, i, j, xi64, &yp64[3 * (size_t)j], l64);
      else e = eval_pair_mirror(k, S, Tg, i, j, xi, &yp[3 * (size_t)j], l, d2t);
      if (e.gate_stage >= 2) c_geo++;
      if (e.gate_stage >= 3) c_col++;
      if (e.gate_stage >= 4) c_sem++;
      bool pass = e.ok && (precision == ORC_FP64 ? (double)e.margin > (double)p->sp_thres : e.a > p->sp_thres);
      if (pass) {
        c_pass++;
        if (stored < K) {
          out.cols[(size_t)i * K + stored] = j;
          out.vals[(size_t)i * K + stored] = e.a;
          stored++;
        }
      }
    };
    if (use_grid) {
      for (int j : cand) visit(j);
    } else {
      for (int j = 0; j < M; ++j) visit(j);
    }
    out.counts[i] = stored;
    c_store += stored;
    mrow = std::max(mrow, stored);
  }
  // gate stages are counted per candidate; stage>=2 means the geometric gate passed
  // (or geometry is off).  Without geometry every target is a geo "pass".
  out.cnt.geo = c_geo;
  out.cnt.color = c_col;
  out.cnt.sem = c_sem;
  out.cnt.passed = c_pass;
  out.cnt.stored = c_store;
  out.cnt.max_row = mrow;
}

// K-nearest association of CvoGPU::align with is_using_kdtree (se_kernel_kdtree
// CvoGPU.cu:1104-1154, fill_in_A_mat_cukdtree :832-949, NearestKSearch
// cukdtree.cuh:365-376): a row takes its K nearest targets, nearest first
// (ties by ascending target index), and stores those that pass the gates at
// the fixed lengthscale ell (no range scaling, :863) with the geometric-type
// floor 0.001 (:898).  Nearness is the squared distance the gate uses (float
// for the mirror and policy precisions, double for FP64).  The reference's
// tree is approximate (its backtracking never revisits the root's far side,
// cukdtree.cuh:326) and queries it with T x instead of T^-1 x (CvoGPU_impl.cu:
// 459 with the tree over the untransformed target, CvoState.cu:98); this
// restatement is the exact K nearest in the frame the gates use.
// brute: scan every target per row instead of the grid.
void knn_assoc_core(const cvo_params* p, const CloudRef& S, const CloudRef& Tg, const Pose64& P, float ell, int K,
                    int precision, int brute, AssocOut& out) {
  const Consts k = make_consts(p);
  const int N = S.n, M = Tg.n;
  const int Kc = std::max(K, 0);
  out.counts.assign(N, 0);
  out.cols.assign((size_t)N * std::max(Kc, 1), -1);
  out.vals.assign((size_t)N * std::max(Kc, 1), 0.0f);
  std::vector<float> yp((size_t)M * 3);
  std::vector<double> yp64;
  float A[9], b[3];
  applied_transform(P, A, b);
  for (int j = 0; j < M; ++j) xform_f(A, b, Tg.xyz + 3 * (size_t)j, &yp[3 * (size_t)j]);
  const bool fp64 = precision == ORC_FP64;
  if (fp64) {
    yp64.resize((size_t)M * 3);
    for (int j = 0; j < M; ++j) {
      const float* y = Tg.xyz + 3 * (size_t)j;
      for (int i = 0; i < 3; ++i) {
        double s = -(P.R[0 * 3 + i] * P.T[0] + P.R[1 * 3 + i] * P.T[1] + P.R[2 * 3 + i] * P.T[2]);
        for (int c = 0; c < 3; ++c) s += P.R[c * 3 + i] * (double)y[c];
        yp64[3 * (size_t)j + i] = s;
      }
    }
  }
  const float l = ell;
  const float d2t = d2_geo_thres(k, l);
  const double l64 = (double)ell;
  const double thr64 = -2.0 * l64 * l64 * std::log((double)p->sp_thres / ((double)p->sigma * p->sigma));
  const double rball = std::sqrt(std::max(0.0, std::max((double)d2t, thr64))) * 1.001 + 1e-6;
  Grid grid;
  if (!brute) grid.build(yp.data(), M, std::max(rball, 1e-6));
  int64_t c_pass = 0, c_store = 0;
  int32_t mrow = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : c_pass, c_store) reduction(max : mrow)
  for (int i = 0; i < N; ++i) {
    const float* xi = S.xyz + 3 * (size_t)i;
    const double xi64[3] = {xi[0], xi[1], xi[2]};
    // (distance, index) of the ball targets
    std::vector<std::pair<double, int>> ball;
    auto consider = [&](int j) {
      if (fp64) {
        const double* y = &yp64[3 * (size_t)j];
        const double dx = y[0] - xi64[0], dy = y[1] - xi64[1], dz = y[2] - xi64[2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < thr64) ball.push_back({d2, j});
      } else {
        const float d2 = sqdist3_f(&yp[3 * (size_t)j], xi);
        if (d2 < d2t) ball.push_back({(double)d2, j});
      }
    };
    if (brute) {
      for (int j = 0; j < M; ++j) consider(j);
    } else {
      const double r = rball + 1e-5 * (1.0 + std::fabs(xi[0]) + std::fabs(xi[1]) + std::fabs(xi[2]));
      grid.query(xi, r, consider);
    }
    const size_t take = std::min(ball.size(), (size_t)Kc);
    std::partial_sort(ball.begin(), ball.begin() + take, ball.end());
    int stored = 0;
    for (size_t q = 0; q < take; ++q) {
      const int j = ball[q].second;
      PairEval e;
      if (fp64) e = eval_pair_fp64(k, p, S, Tg, i, j, xi64, &yp64[3 * (size_t)j], l64, 0.001);
      else e = eval_pair_mirror(k, S, Tg, i, j, xi, &yp[3 * (size_t)j], l, d2t, 0.001f);
      const bool pass = e.ok && (fp64 ? (double)e.margin > (double)p->sp_thres : e.a > p->sp_thres);
      if (!pass) continue;
      c_pass++;
      out.cols[(size_t)i * Kc + stored] = j;
      out.vals[(size_t)i * Kc + stored] = e.a;
      stored++;
    }
    out.counts[i] = stored;
    c_store += stored;
    mrow = std::max(mrow, stored);
  }
  out.cnt.passed = c_pass;
  out.cnt.stored = c_store;
  out.cnt.max_row = mrow;
}

// The row cap of the next K-nearest pass: the mean stored row rounded up
// times is_using_kdtree, at most KDTREE_K_SIZE (CvoGPU.cu:1416-1417).
int knn_next_k(const cvo_params* p, int64_t nnz, int rows) {
  const int mean = (int)std::ceil((double)nnz / (double)rows);
  return std::min(kKnnK, mean * p->is_using_kdtree);
}

// ---------------------------------------------------------------------------
// Flow (CvoGPU.cu:2001-2120) and step coefficients (:333-462).
// ---------------------------------------------------------------------------

void flow_core(const cvo_params* p, int N, const float* X, const float* Y, int K, const int32_t* counts,
               const int32_t* cols, const float* vals, int precision, double out6[6], double unit6[6]) {
  double sum[6] = {0, 0, 0, 0, 0, 0};
  if (precision == ORC_POLICY) {
    // the library's flow: float products a * (x x y) and a * (y - x), summed
    // in FP64; omega / c and v / d on the totals; FP64 normalisation
    std::vector<double> rows((size_t)N * 6);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      const float* x = X + 3 * (size_t)i;
      double acc[6] = {0, 0, 0, 0, 0, 0};
      for (int s = 0; s < counts[i]; ++s) {
        const int j = cols[(size_t)i * K + s];
        const float a = vals[(size_t)i * K + s];
        const float* y = Y + 3 * (size_t)j;
        const float c0 = x[1] * y[2] - x[2] * y[1], c1 = x[2] * y[0] - x[0] * y[2], c2 = x[0] * y[1] - x[1] * y[0];
        acc[0] += (double)(c0 * a);
        acc[1] += (double)(c1 * a);
        acc[2] += (double)(c2 * a);
        acc[3] += (double)((y[0] - x[0]) * a);
        acc[4] += (double)((y[1] - x[1]) * a);
        acc[5] += (double)((y[2] - x[2]) * a);
      }
      for (int d = 0; d < 6; ++d) rows[(size_t)i * 6 + d] = acc[d];
    }
    for (int i = 0; i < N; ++i)
      for (int d = 0; d < 6; ++d) sum[d] += rows[(size_t)i * 6 + d];
    for (int d = 0; d < 3; ++d) {
      sum[d] /= (double)p->c;
      sum[3 + d] /= (double)p->d;
    }
    double z = 0.0;
    for (int d = 0; d < 6; ++d) z += sum[d] * sum[d];
    const double s = z > 0.0 ? std::sqrt(z) : 1.0;
    for (int d = 0; d < 6; ++d) {
      out6[d] = sum[d];
      unit6[d] = z > 0.0 ? sum[d] / s : sum[d];
    }
  } else if (precision == ORC_MIRROR) {
    std::vector<double> rows((size_t)N * 6);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      const float* x = X + 3 * (size_t)i;
      float om[3] = {0, 0, 0}, v[3] = {0, 0, 0};
      for (int s = 0; s < counts[i]; ++s) {
        int j = cols[(size_t)i * K + s];
        float a = vals[(size_t)i * K + s];
        const float* y = Y + 3 * (size_t)j;
        float cr[3] = {x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]};
        for (int d = 0; d < 3; ++d) {
          om[d] = om[d] + cr[d] * a;
          v[d] = v[d] + (y[d] - x[d]) * a;
        }
      }
      for (int d = 0; d < 3; ++d) {
        rows[(size_t)i * 6 + d] = (double)(om[d] / p->c);
        rows[(size_t)i * 6 + 3 + d] = (double)(v[d] / p->d);
      }
    }
    for (int i = 0; i < N; ++i)
      for (int d = 0; d < 6; ++d) sum[d] += rows[(size_t)i * 6 + d];
    float f[6];
    for (int d = 0; d < 6; ++d) f[d] = static_cast<float>(sum[d]);
    float z = 0.0f;
    for (int d = 0; d < 6; ++d) z += f[d] * f[d];
    if (z > 0.0f) {
      float s = sqrtf(z);
      for (int d = 0; d < 6; ++d) f[d] = f[d] / s;
    }
    for (int d = 0; d < 6; ++d) {
      out6[d] = sum[d];
      unit6[d] = f[d];
    }
  } else {
    std::vector<double> rows((size_t)N * 6);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      const float* xf = X + 3 * (size_t)i;
      double x[3] = {xf[0], xf[1], xf[2]};
      double om[3] = {0, 0, 0}, v[3] = {0, 0, 0};
      for (int s = 0; s < counts[i]; ++s) {
        int j = cols[(size_t)i * K + s];
        double a = vals[(size_t)i * K + s];
        const float* yf = Y + 3 * (size_t)j;
        double y[3] = {yf[0], yf[1], yf[2]};
        double cr[3] = {x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]};
        for (int d = 0; d < 3; ++d) {
          om[d] += cr[d] * a;
          v[d] += (y[d] - x[d]) * a;
        }
      }
      for (int d = 0; d < 3; ++d) {
        rows[(size_t)i * 6 + d] = om[d] / (double)p->c;
        rows[(size_t)i * 6 + 3 + d] = v[d] / (double)p->d;
      }
    }
    for (int i = 0; i < N; ++i)
      for (int d = 0; d < 6; ++d) sum[d] += rows[(size_t)i * 6 + d];
    double z = 0;
    for (int d = 0; d < 6; ++d) z += sum[d] * sum[d];
    double s = z > 0 ? std::sqrt(z) : 1.0;
    for (int d = 0; d < 6; ++d) {
      out6[d] = sum[d];
      unit6[d] = z > 0 ? sum[d] / s : sum[d];
    }
  }
}

template <class T>
inline void cross3(const T* a, const T* b, T* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
template <class T>
inline T dot3(const T* a, const T* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

void step_core(const cvo_params* p, float ell, int N, const float* X, const float* Y, int K, const int32_t* counts,
               const int32_t* cols, const float* vals, const double flow6[6], int precision, double bcde[4]) {
  double tot[4] = {0, 0, 0, 0};
  std::vector<double> rows((size_t)N * 4);
  if (precision == ORC_POLICY) {
    // the library's step terms: xi^k z by repeated cross products in float
    const float w0 = (float)flow6[0], w1 = (float)flow6[1], w2 = (float)flow6[2];
    const float v0 = (float)flow6[3], v1 = (float)flow6[4], v2 = (float)flow6[5];
    auto cr = [&](float a0, float a1, float a2, float& o0, float& o1, float& o2) {
      o0 = w1 * a2 - w2 * a1;
      o1 = w2 * a0 - w0 * a2;
      o2 = w0 * a1 - w1 * a0;
    };
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      const float* x = X + 3 * (size_t)i;
      float temp_ell = ell;
      if (p->is_using_range_ell) {
        const float px = sqrtf(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        temp_ell = (float)(((double)px / 500.0 + 1.0) * (double)ell);
      }
      const float tc = (float)(1.0 / (2.0 * (double)temp_ell * (double)temp_ell));
      double acc[4] = {0, 0, 0, 0};
      for (int s = 0; s < counts[i]; ++s) {
        const int j = cols[(size_t)i * K + s];
        const float a = vals[(size_t)i * K + s];
        const float* z = Y + 3 * (size_t)j;
        float a0, a1, a2, b0, b1, b2, c0, c1, c2, d0, d1, d2;
        cr(z[0], z[1], z[2], a0, a1, a2);
        a0 += v0; a1 += v1; a2 += v2;
        cr(a0, a1, a2, b0, b1, b2);
        cr(b0, b1, b2, c0, c1, c2);
        cr(c0, c1, c2, d0, d1, d2);
        const float df0 = x[0] - z[0], df1 = x[1] - z[1], df2 = x[2] - z[2];
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
      }
      for (int c = 0; c < 4; ++c) rows[(size_t)i * 4 + c] = acc[c];
    }
  } else if (precision == ORC_MIRROR) {
    float w[3] = {(float)flow6[0], (float)flow6[1], (float)flow6[2]};
    float v[3] = {(float)flow6[3], (float)flow6[4], (float)flow6[5]};
    // Omega = skew(w) and its powers (float), compute_step_size_xi (:333-378)
    float W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    auto mm = [](const float* A, const float* B, float* C) {
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) C[r * 3 + c] = A[r * 3 + 0] * B[0 * 3 + c] + A[r * 3 + 1] * B[1 * 3 + c] + A[r * 3 + 2] * B[2 * 3 + c];
    };
    auto mv = [](const float* A, const float* x, float* o) {
      for (int r = 0; r < 3; ++r) o[r] = A[r * 3 + 0] * x[0] + A[r * 3 + 1] * x[1] + A[r * 3 + 2] * x[2];
    };
    float W2[9], W3[9], W4[9];
    mm(W, W, W2);
    mm(W2, W, W3);
    mm(W3, W, W4);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      const float* x = X + 3 * (size_t)i;
      float px_norm = sqrtf(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
      float temp_ell = ell;
      if (p->is_using_range_ell) temp_ell = static_cast<float>(((double)px_norm / 500.0 + 1.0) * (double)ell);
      float temp_coef = static_cast<float>(1.0 / (2.0 * (double)temp_ell * (double)temp_ell));
      double B = 0, C = 0, D = 0, E = 0;
      for (int s = 0; s < counts[i]; ++s) {
        int j = cols[(size_t)i * K + s];
        float Aij = vals[(size_t)i * K + s];
        const float* z = Y + 3 * (size_t)j;
        float xiz[3], t1[3], t2[3], xi2z[3], xi3z[3], xi4z[3];
        cross3(w, z, xiz);
        for (int d = 0; d < 3; ++d) xiz[d] += v[d];
        mv(W2, z, t1); mv(W, v, t2);
        for (int d = 0; d < 3; ++d) xi2z[d] = t1[d] + t2[d];
        mv(W3, z, t1); mv(W2, v, t2);
        for (int d = 0; d < 3; ++d) xi3z[d] = t1[d] + t2[d];
        mv(W4, z, t1); mv(W3, v, t2);
        for (int d = 0; d < 3; ++d) xi4z[d] = t1[d] + t2[d];
        float normxiz2 = dot3(xiz, xiz);
        float xiz_dot_xi2z = -dot3(xiz, xi2z);
        float epsil_const = dot3(xi2z, xi2z) + 2.0f * dot3(xiz, xi3z);
        float diff[3] = {x[0] - z[0], x[1] - z[1], x[2] - z[2]};
        float two_xi2z[3] = {2.0f * xi2z[0], 2.0f * xi2z[1], 2.0f * xi2z[2]};
        float m_xi3z[3] = {-xi3z[0], -xi3z[1], -xi3z[2]};
        float two_xi4z[3] = {2.0f * xi4z[0], 2.0f * xi4z[1], 2.0f * xi4z[2]};
        float beta = static_cast<float>(-2.0 * (double)temp_coef * (double)dot3(xiz, diff));
        float gamma = -temp_coef * (normxiz2 + dot3(two_xi2z, diff));
        float delta = static_cast<float>(2.0 * (double)temp_coef * (double)(xiz_dot_xi2z + dot3(m_xi3z, diff)));
        float epsil = -temp_coef * (epsil_const + dot3(two_xi4z, diff));
        double bi = (double)(Aij * beta);
        double ci = (double)Aij * ((double)gamma + (double)(beta * beta) / 2.0);
        double di = (double)Aij * ((double)(delta + beta * gamma) + (double)(beta * beta * beta) / 6.0);
        double ei = (double)Aij * ((double)(epsil + beta * delta) + 0.5 * (double)beta * (double)beta * (double)gamma +
                                   0.5 * (double)gamma * (double)gamma +
                                   (1 / 24.0) * (double)beta * (double)beta * (double)beta * (double)beta);
        B += bi; C += ci; D += di; E += ei;
      }
      rows[(size_t)i * 4 + 0] = B; rows[(size_t)i * 4 + 1] = C;
      rows[(size_t)i * 4 + 2] = D; rows[(size_t)i * 4 + 3] = E;
    }
  } else {
    double w[3] = {flow6[0], flow6[1], flow6[2]};
    double v[3] = {flow6[3], flow6[4], flow6[5]};
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      const float* xf = X + 3 * (size_t)i;
      double x[3] = {xf[0], xf[1], xf[2]};
      double temp_ell = ell;
      if (p->is_using_range_ell) temp_ell = (std::sqrt(dot3(x, x)) / 500.0 + 1.0) * (double)ell;
      double c = 1.0 / (2.0 * temp_ell * temp_ell);
      double B = 0, C = 0, D = 0, E = 0;
      for (int s = 0; s < counts[i]; ++s) {
        int j = cols[(size_t)i * K + s];
        double a = vals[(size_t)i * K + s];
        const float* zf = Y + 3 * (size_t)j;
        double z[3] = {zf[0], zf[1], zf[2]};
        // xi^k z = w x xi^{k-1} z (homogeneous twist powers), xi z = w x z + v
        double x1[3], x2[3], x3[3], x4[3];
        cross3(w, z, x1);
        for (int d = 0; d < 3; ++d) x1[d] += v[d];
        cross3(w, x1, x2);
        cross3(w, x2, x3);
        cross3(w, x3, x4);
        double diff[3] = {x[0] - z[0], x[1] - z[1], x[2] - z[2]};
        double beta = -2.0 * c * dot3(x1, diff);
        double gamma = -c * (dot3(x1, x1) + 2.0 * dot3(x2, diff));
        double delta = 2.0 * c * (-dot3(x1, x2) - dot3(x3, diff));
        double epsil = -c * (dot3(x2, x2) + 2.0 * dot3(x1, x3) + 2.0 * dot3(x4, diff));
        B += a * beta;
        C += a * (gamma + beta * beta / 2.0);
        D += a * (delta + beta * gamma + beta * beta * beta / 6.0);
        E += a * (epsil + beta * delta + 0.5 * beta * beta * gamma + 0.5 * gamma * gamma + beta * beta * beta * beta / 24.0);
      }
      rows[(size_t)i * 4 + 0] = B; rows[(size_t)i * 4 + 1] = C;
      rows[(size_t)i * 4 + 2] = D; rows[(size_t)i * 4 + 3] = E;
    }
  }
  for (int i = 0; i < N; ++i)
    for (int d = 0; d < 4; ++d) tot[d] += rows[(size_t)i * 4 + d];
  for (int d = 0; d < 4; ++d) bcde[d] = tot[d];
}

// Roots of a cubic by the Aberth-Ehrlich simultaneous iteration in long double
// (a different method from the device's closed form, so the two cross-check).
void cubic_roots(const double c[4], double re[3], double im[3]) {
  typedef std::complex<long double> C;
  long double a0 = c[0];
  if (a0 == 0.0L || !std::isfinite((double)a0)) {
    for (int i = 0; i < 3; ++i) { re[i] = std::numeric_limits<double>::quiet_NaN(); im[i] = 0; }
    return;
  }
  C q[4] = {C(1.0L), C((long double)c[1] / a0), C((long double)c[2] / a0), C((long double)c[3] / a0)};
  for (int i = 1; i < 4; ++i) {
    if (!std::isfinite((double)q[i].real())) {
      for (int k = 0; k < 3; ++k) { re[k] = std::numeric_limits<double>::quiet_NaN(); im[k] = 0; }
      return;
    }
  }
  auto P = [&](C z) { return ((q[0] * z + q[1]) * z + q[2]) * z + q[3]; };
  auto dP = [&](C z) { return (3.0L * q[0] * z + 2.0L * q[1]) * z + q[2]; };
  long double bound = 1.0L;
  for (int i = 1; i < 4; ++i) bound = std::max(bound, 1.0L + std::abs(q[i]));
  C z[3];
  for (int k = 0; k < 3; ++k) z[k] = std::polar(bound * 0.5L, (long double)(0.4 + 2.0943951023931954923 * k));
  for (int it = 0; it < 500; ++it) {
    long double change = 0;
    for (int k = 0; k < 3; ++k) {
      C pk = P(z[k]), dk = dP(z[k]);
      if (std::abs(pk) == 0.0L) continue;
      C ratio = pk / dk;
      C s = 0;
      for (int m = 0; m < 3; ++m)
        if (m != k) s += 1.0L / (z[k] - z[m]);
      C w = ratio / (1.0L - ratio * s);
      if (!std::isfinite((double)std::abs(w))) w = ratio;
      z[k] -= w;
      change = std::max(change, std::abs(w) / std::max(1.0L, std::abs(z[k])));
    }
    if (change < 1e-17L) break;
  }
  // Polish with Newton on each root; conjugate pairs of a real polynomial
  for (int k = 0; k < 3; ++k) {
    for (int it = 0; it < 3; ++it) {
      C d = dP(z[k]);
      if (std::abs(d) == 0.0L) break;
      z[k] -= P(z[k]) / d;
    }
    re[k] = (double)z[k].real();
    im[k] = (double)z[k].imag();
    // a real root reached from a complex start keeps an imaginary residue
    // far below rounding (1e-98 and the like): it is real
    if (std::fabs(im[k]) <= 1e-14 * std::max(1.0, std::fabs(re[k]))) im[k] = 0.0;
  }
}

// compute_step_size root selection (CvoGPU.cu:504-538).  Note the reference's
// clamp runs on temp_step itself, so "no usable root" (temp_step = DBL_MAX) ends
// at max_step.
double select_step(const cvo_params* p, const double bcde[4]) {
  double coef[4] = {4.0 * bcde[3], 3.0 * bcde[2], 2.0 * bcde[1], bcde[0]};
  double re[3], im[3];
  cubic_roots(coef, re, im);
  double temp = std::numeric_limits<double>::max();
  for (int i = 0; i < 3; ++i)
    if (re[i] > 0 && re[i] < temp && std::fabs(im[i]) < 1e-5) temp = re[i];
  double step;
  if (temp > (double)p->max_step) step = p->max_step;
  else if (temp < (double)p->min_step) step = p->min_step;
  else step = temp;
  return step;
}

// Exp_SEK3 float form (LieGroup.cpp:244-275) including its theta<1e-6 branch.
void exp_sek3_f32(const float xi[6], float dt, float out[12]) {
  float w[3] = {xi[0], xi[1], xi[2]};
  float theta = sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, J[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (!(theta < 1e-6f)) {
    float A[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    float A2[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) A2[r * 3 + c] = A[r * 3 + 0] * A[0 * 3 + c] + A[r * 3 + 1] * A[1 * 3 + c] + A[r * 3 + 2] * A[2 * 3 + c];
    float theta2 = theta * theta;
    float st = sinf(dt * theta), ct = cosf(dt * theta);
    float omc = (1 - ct) / theta2;
    float k3 = (dt * theta - st) / (theta2 * theta);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        float I = r == c ? 1.0f : 0.0f;
        R[r * 3 + c] = I + (st / theta) * A[r * 3 + c] + omc * A2[r * 3 + c];
        J[r * 3 + c] = dt * I + omc * A[r * 3 + c] + k3 * A2[r * 3 + c];
      }
  }
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) out[r * 4 + c] = R[r * 3 + c];
    out[r * 4 + 3] = J[r * 3 + 0] * xi[3] + J[r * 3 + 1] * xi[4] + J[r * 3 + 2] * xi[5];
  }
}

// exp(dt * [w; v]) in FP64 with exact small-angle series.
void exp_se3_f64(const double xi[6], double dt, double out[12]) {
  double ph[3] = {xi[0] * dt, xi[1] * dt, xi[2] * dt};
  double u[3] = {xi[3] * dt, xi[4] * dt, xi[5] * dt};
  double th2 = ph[0] * ph[0] + ph[1] * ph[1] + ph[2] * ph[2];
  double th = std::sqrt(th2);
  double A, B, Cc;  // sin/th, (1-cos)/th^2, (th - sin)/th^3
  if (th < 1e-3) {
    A = 1 - th2 / 6 + th2 * th2 / 120;
    B = 0.5 - th2 / 24 + th2 * th2 / 720;
    Cc = 1.0 / 6 - th2 / 120 + th2 * th2 / 5040;
  } else {
    A = std::sin(th) / th;
    B = (1 - std::cos(th)) / th2;
    Cc = (th - std::sin(th)) / (th2 * th);
  }
  double W[9] = {0, -ph[2], ph[1], ph[2], 0, -ph[0], -ph[1], ph[0], 0};
  double W2[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) W2[r * 3 + c] = W[r * 3 + 0] * W[0 * 3 + c] + W[r * 3 + 1] * W[1 * 3 + c] + W[r * 3 + 2] * W[2 * 3 + c];
  double V[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double I = r == c ? 1.0 : 0.0;
      out[r * 4 + c] = I + A * W[r * 3 + c] + B * W2[r * 3 + c];
      V[r * 3 + c] = I + B * W[r * 3 + c] + Cc * W2[r * 3 + c];
    }
  for (int r = 0; r < 3; ++r) out[r * 4 + 3] = V[r * 3 + 0] * u[0] + V[r * 3 + 1] * u[1] + V[r * 3 + 2] * u[2];
}

// SE(3) log (tangent [upsilon; omega]) via the rotation's unit quaternion, the
// same construction as Sophus::SE3d::log used by dist_se3_cpu.
void log_se3(const double T[12], double out[6]) {
  double R[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) R[r * 3 + c] = T[r * 4 + c];
  // Shepperd's method
  double qw, qx, qy, qz;
  double tr = R[0] + R[4] + R[8];
  if (tr > 0) {
    double s = std::sqrt(tr + 1.0) * 2;
    qw = 0.25 * s; qx = (R[7] - R[5]) / s; qy = (R[2] - R[6]) / s; qz = (R[3] - R[1]) / s;
  } else if (R[0] > R[4] && R[0] > R[8]) {
    double s = std::sqrt(1.0 + R[0] - R[4] - R[8]) * 2;
    qw = (R[7] - R[5]) / s; qx = 0.25 * s; qy = (R[1] + R[3]) / s; qz = (R[2] + R[6]) / s;
  } else if (R[4] > R[8]) {
    double s = std::sqrt(1.0 + R[4] - R[0] - R[8]) * 2;
    qw = (R[2] - R[6]) / s; qx = (R[1] + R[3]) / s; qy = 0.25 * s; qz = (R[5] + R[7]) / s;
  } else {
    double s = std::sqrt(1.0 + R[8] - R[0] - R[4]) * 2;
    qw = (R[3] - R[1]) / s; qx = (R[2] + R[6]) / s; qy = (R[5] + R[7]) / s; qz = 0.25 * s;
  }
  double qn = std::sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
  qw /= qn; qx /= qn; qy /= qn; qz /= qn;
  double n2 = qx * qx + qy * qy + qz * qz, n = std::sqrt(n2);
  double f;  // omega = f * q_vec
  if (n2 < 1e-20) {
    f = 2.0 / qw - 2.0 * n2 / (qw * qw * qw);
  } else if (std::fabs(qw) < 1e-10) {
    f = (qw > 0 ? M_PI : -M_PI) / n;
  } else {
    f = 2.0 * std::atan(n / qw) / n;
  }
  double w[3] = {f * qx, f * qy, f * qz};
  double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  double W2[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) W2[r * 3 + c] = W[r * 3 + 0] * W[0 * 3 + c] + W[r * 3 + 1] * W[1 * 3 + c] + W[r * 3 + 2] * W[2 * 3 + c];
  double k2;
  if (th < 1e-6) k2 = 1.0 / 12.0;
  else k2 = (1.0 - th * std::cos(0.5 * th) / (2.0 * std::sin(0.5 * th))) / (th * th);
  double t[3] = {T[3], T[7], T[11]};
  for (int r = 0; r < 3; ++r) {
    double s = 0;
    for (int c = 0; c < 3; ++c) {
      double Vi = (r == c ? 1.0 : 0.0) - 0.5 * W[r * 3 + c] + k2 * W2[r * 3 + c];
      s += Vi * t[c];
    }
    out[r] = s;
  }
  out[3] = w[0]; out[4] = w[1]; out[5] = w[2];
}

double log_se3_norm(const double T[12]) {
  double v[6];
  log_se3(T, v);
  double s = 0;
  for (int i = 0; i < 6; ++i) s += v[i] * v[i];
  return std::sqrt(s);
}

// A_sparsity_indicator_ell_update (CvoGPU.cu:547-665) with float sums.
template <class T>
struct Indicator {
  std::deque<T> qs, qe;
  T ss = 0, se = 0;
  bool update(T ind, int W, T thr) {
    bool dec = false;
    if ((int)qs.size() < W) { qs.push_back(ind); ss += ind; }
    if ((int)qs.size() >= W && (int)qe.size() < W) { qe.push_back(ind); se += ind; }
    if ((int)qs.size() >= W && (int)qe.size() >= W) {
      T ratio = se / ss;
      if (ratio > (T)1 - thr && ratio < (T)1 + thr) {
        dec = true;
        qs.clear(); qe.clear();
        ss = 0; se = 0;
      } else {
        T f = qe.front();
        se -= f;
        ss += f;
        qs.push_back(f);
        qe.pop_front();
        ss -= qs.front();
        qs.pop_front();
        qe.push_back(ind);
        se += ind;
      }
    }
    return dec;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" int orc_assoc(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                         const float T[16], float ell, int32_t K, int32_t precision, int32_t brute,
                         int32_t* counts, int32_t* cols, float* vals, orc_counters* cnt) {
  if (!p || !src || !tgt || !T || K <= 0) return 1;
  AssocOut o;
  assoc_core(p, ref(src), ref(tgt), pose_from16(T), ell, K, precision, brute, o);
  if (counts) std::memcpy(counts, o.counts.data(), sizeof(int32_t) * o.counts.size());
  if (cols) std::memcpy(cols, o.cols.data(), sizeof(int32_t) * o.cols.size());
  if (vals) std::memcpy(vals, o.vals.data(), sizeof(float) * o.vals.size());
  if (cnt) *cnt = o.cnt;
  return 0;
}

extern "C" int orc_assoc_knn(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                             const float T[16], float ell, int32_t K, int32_t precision, int32_t brute,
                             int32_t* counts, int32_t* cols, float* vals, orc_counters* cnt) {
  if (!p || !src || !tgt || !T || K <= 0) return 1;
  AssocOut o;
  knn_assoc_core(p, ref(src), ref(tgt), pose_from16(T), ell, K, precision, brute, o);
  if (counts) std::memcpy(counts, o.counts.data(), sizeof(int32_t) * o.counts.size());
  if (cols) std::memcpy(cols, o.cols.data(), sizeof(int32_t) * o.cols.size());
  if (vals) std::memcpy(vals, o.vals.data(), sizeof(float) * o.vals.size());
  if (cnt) *cnt = o.cnt;
  return 0;
}

extern "C" int orc_pair_value(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                              const float T[16], float ell, int32_t i, int32_t j, double* a_out,
                              double* gate_margin) {
  const Consts k = make_consts(p);
  CloudRef S = ref(src), Tg = ref(tgt);
  float A[9], b[3], yp[3];
  applied_transform(pose_from16(T), A, b);
  xform_f(A, b, Tg.xyz + 3 * (size_t)j, yp);
  const float* xi = S.xyz + 3 * (size_t)i;
  float l = range_ell(ell, xi);
  PairEval e = eval_pair_mirror(k, S, Tg, i, j, xi, yp, l, d2_geo_thres(k, l));
  // also the exact (double) product value, for the sp_thres margin
  double a_exact = -1.0;
  if (e.ok) a_exact = (double)e.a;
  if (a_out) *a_out = a_exact;
  if (gate_margin) *gate_margin = e.margin;
  return e.ok ? 0 : 2;
}

// orc_pair_value over n (i, j) pairs: a (-1 when a gate rejects) and the
// smallest relative gate margin of each.
extern "C" int orc_pair_values(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                               const float T[16], float ell, int64_t n, const int32_t* ii, const int32_t* jj,
                               double* a_out, double* margin_out) {
  const Consts k = make_consts(p);
  CloudRef S = ref(src), Tg = ref(tgt);
  float A[9], b[3];
  applied_transform(pose_from16(T), A, b);
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < n; ++q) {
    float yp[3];
    xform_f(A, b, Tg.xyz + 3 * (size_t)jj[q], yp);
    const float* xi = S.xyz + 3 * (size_t)ii[q];
    const float l = range_ell(ell, xi);
    const PairEval e = eval_pair_mirror(k, S, Tg, ii[q], jj[q], xi, yp, l, d2_geo_thres(k, l));
    a_out[q] = e.ok ? (double)e.a : -1.0;
    margin_out[q] = e.margin;
  }
  return 0;
}

// CvoGPU::inner_product_cpu (src/cvo/CvoGPU.cpp:204-231) with
// se_kernel_init_ell_cpu (:114-186): the target moved into the source frame by
// T^-1 (float, canonical FMA order), a radius search of squared radius
// d2_thres = -2 l^2 log(sp / sigma^2) at the fixed lengthscale l = ell (no
// range scaling, no row cap), and for every target inside it
// a = ck k sk -- k = float(sigma^2 exp(-d2 / 2l^2)) with geometry,
// ck = float(c_sigma^2 exp(-|f_x - f_y|^2 / 2 c_ell^2)) with intensity,
// sk = float(s_sigma^2 exp(-|s_x - s_y|^2 / 2 s_ell^2)) with semantics, no
// colour / semantic / geometric-type gate -- counted when a > sp_thres.  The
// reference sums the float sparse matrix; here the sum is FP64.  Searching
// the grid or scanning every target (brute) gives the same set.  pairs_out
// (optional): the number of pairs counted.
extern "C" double orc_inner_product_cpu(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                        const float T[16], float ell, int32_t brute, int64_t* pairs_out) {
  const Consts k = make_consts(p);
  CloudRef S = ref(src), Tg = ref(tgt);
  const int N = S.n, M = Tg.n;
  if (N == 0 || M == 0) {
    if (pairs_out) *pairs_out = 0;
    return 0.0;
  }
  float A[9], b[3];
  applied_transform(pose_from16(T), A, b);
  std::vector<float> yp((size_t)M * 3);
  for (int j = 0; j < M; ++j) xform_f(A, b, Tg.xyz + 3 * (size_t)j, &yp[3 * (size_t)j]);
  const float l = ell;
  const float d2t = static_cast<float>(-2.0 * (double)l * (double)l * (double)k.lg);
  Grid grid;
  const double r = std::sqrt(std::max(0.0, (double)d2t)) * 1.001 + 1e-6;
  if (!brute) grid.build(yp.data(), M, std::max(r, 1e-6));
  double total = 0.0;
  int64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total, cnt)
  for (int i = 0; i < N; ++i) {
    const float* xi = S.xyz + 3 * (size_t)i;
    auto visit = [&](int j) {
      const float d2 = sqdist3_f(&yp[3 * (size_t)j], xi);
      if (!(d2 < d2t)) return;
      float kk = 1.0f, ck = 1.0f, sk = 1.0f;
      if (k.geo) kk = static_cast<float>((double)k.sigma2 * std::exp((double)(-d2) / (2.0 * (double)l * (double)l)));
      if (k.col) {
        const float d2c = sqdist_n_f(S.feat + (size_t)i * S.F, Tg.feat + (size_t)j * Tg.F, S.F);
        ck = static_cast<float>((double)k.c_sigma2 * std::exp((double)(-d2c) / (2.0 * (double)k.c2)));
      }
      if (k.sem) {
        const float d2s = sqdist_n_f(S.sem + (size_t)i * S.K, Tg.sem + (size_t)j * Tg.K, S.K);
        sk = static_cast<float>((double)k.s_sigma2 *
                                std::exp((double)(-d2s) / (2.0 * (double)k.s_ell * (double)k.s_ell)));
      }
      const float a = ck * kk * sk;
      if (a > k.sp) {
        total += (double)a;
        cnt++;
      }
    };
    if (brute) {
      for (int j = 0; j < M; ++j) visit(j);
    } else {
      grid.query(xi, r, visit);
    }
  }
  if (pairs_out) *pairs_out = cnt;
  return total;
}

extern "C" int orc_transform_target(const float T[16], int32_t m, const float* xyz, float* out) {
  float A[9], b[3];
  applied_transform(pose_from16(T), A, b);
  for (int j = 0; j < m; ++j) xform_f(A, b, xyz + 3 * (size_t)j, out + 3 * (size_t)j);
  return 0;
}

extern "C" int orc_flow(const cvo_params* p, int32_t n, const float* src_xyz, const float* yprime, int32_t K,
                        const int32_t* counts, const int32_t* cols, const float* vals, int32_t precision,
                        double out6[6], double unit6[6]) {
  flow_core(p, n, src_xyz, yprime, K, counts, cols, vals, precision, out6, unit6);
  return 0;
}

extern "C" int orc_step_coeffs(const cvo_params* p, float ell, int32_t n, const float* src_xyz, const float* yprime,
                               int32_t K, const int32_t* counts, const int32_t* cols, const float* vals,
                               const double flow6[6], int32_t precision, double bcde[4]) {
  step_core(p, ell, n, src_xyz, yprime, K, counts, cols, vals, flow6, precision, bcde);
  return 0;
}

extern "C" int orc_cubic_roots(const double c[4], double re[3], double im[3]) {
  cubic_roots(c, re, im);
  return 0;
}

extern "C" double orc_select_step(const cvo_params* p, const double bcde[4]) { return select_step(p, bcde); }

extern "C" int orc_exp_sek3_f32(const float xi[6], float dt, float out12[12]) {
  exp_sek3_f32(xi, dt, out12);
  return 0;
}

extern "C" int orc_exp_se3_f64(const double xi[6], double dt, double out12[12]) {
  exp_se3_f64(xi, dt, out12);
  return 0;
}

extern "C" double orc_log_se3_norm(const double T12[12]) { return log_se3_norm(T12); }

extern "C" int orc_log_se3(const double T12[12], double out6[6]) {
  log_se3(T12, out6);
  return 0;
}

extern "C" int orc_indicator_sequence(int32_t window, float threshold, int32_t n, const float* ind, int32_t* dec) {
  Indicator<float> st;
  for (int i = 0; i < n; ++i) dec[i] = st.update(ind[i], window, threshold) ? 1 : 0;
  return 0;
}

extern "C" int orc_align(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                         const float T_init[16], int32_t precision, int32_t brute, float T_out[16],
                         int32_t* iters, int32_t* ret_code, cvo_iter_record* trace, int32_t trace_cap) {
  if (!p || !src || !tgt || !T_init || !T_out) return 1;
  CloudRef S = ref(src), Tg = ref(tgt);
  const int N = S.n, M = Tg.n;
  // pose state: float (mirror) or double
  float Rf[9], Tf[3];
  double Rd[9], Td[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) { Rf[r * 3 + c] = T_init[r * 4 + c]; Rd[r * 3 + c] = T_init[r * 4 + c]; }
    Tf[r] = T_init[r * 4 + 3];
    Td[r] = T_init[r * 4 + 3];
  }
  float ell = p->ell_init;
  double ell64 = p->ell_init;
  // ORC_POLICY: mirror kernels, FP64 pose; ORC_MIRROR: float pose as well
  const int kern = precision == ORC_FP64 ? ORC_FP64 : ORC_MIRROR;  // association
  const int fkern = precision;                                         // flow and step terms
  const bool float_pose = precision == ORC_MIRROR;
  const bool float_ell = precision != ORC_FP64;
  const bool knn = p->is_using_kdtree != 0;
  int K = knn ? kKnnK : p->nearest_neighbors_max;  // CvoGPU.cu:1287
  Indicator<float> ind32;
  Indicator<double> ind64;
  int ret = 0, k = 0;
  std::vector<float> yp((size_t)M * 3);
  for (; k < p->MAX_ITER; ++k) {
    Pose64 P;
    for (int i = 0; i < 9; ++i) P.R[i] = float_pose ? (double)Rf[i] : Rd[i];
    for (int i = 0; i < 3; ++i) P.T[i] = float_pose ? (double)Tf[i] : Td[i];
    float A[9], b[3];
    applied_transform(P, A, b);
    for (int j = 0; j < M; ++j) xform_f(A, b, Tg.xyz + 3 * (size_t)j, &yp[3 * (size_t)j]);
    AssocOut o;
    float ell_used = float_ell ? ell : (float)ell64;
    int stride = K;
    if (knn) {
      knn_assoc_core(p, S, Tg, P, ell_used, K, kern, brute, o);
      stride = std::max(K, 1);
    } else {
      assoc_core(p, S, Tg, P, ell_used, K, kern, brute, o);
    }
    double sum6[6], unit6[6];
    flow_core(p, N, S.xyz, yp.data(), stride, o.counts.data(), o.cols.data(), o.vals.data(), fkern, sum6, unit6);
    double bcde[4];
    step_core(p, ell_used, N, S.xyz, yp.data(), stride, o.counts.data(), o.cols.data(), o.vals.data(), unit6, fkern,
              bcde);
    double step = select_step(p, bcde);
    if (precision == ORC_MIRROR) step = (double)(float)step;  // cvo_state->step is float
    // stop test on the normalised flow (CvoGPU.cu:1354-1360)
    double u6[6];
    for (int d = 0; d < 6; ++d) u6[d] = precision == ORC_POLICY ? (double)(float)unit6[d] : unit6[d];
    double nw = std::sqrt(u6[0] * u6[0] + u6[1] * u6[1] + u6[2] * u6[2]);
    double nv = std::sqrt(u6[3] * u6[3] + u6[4] * u6[4] + u6[5] * u6[5]);
    cvo_iter_record rec{};
    rec.ell = ell_used;
    rec.step = (float)step;
    rec.nnz = (int32_t)o.cnt.stored;
    rec.max_row = o.cnt.max_row;
    rec.num_neighbors = K;
    rec.flow_norm = (float)std::sqrt(sum6[0] * sum6[0] + sum6[1] * sum6[1] + sum6[2] * sum6[2] + sum6[3] * sum6[3] +
                                     sum6[4] * sum6[4] + sum6[5] * sum6[5]);
    if (nw < p->eps && nv < p->eps) {
      if (nw < 1e-8 && nv < 1e-8) ret = -1;
      if (trace && k < trace_cap) trace[k] = rec;
      break;
    }
    double dRT[12];
    if (float_pose) {
      float xi[6];
      for (int d = 0; d < 6; ++d) xi[d] = (float)unit6[d];
      float dt[12];
      exp_sek3_f32(xi, (float)step, dt);
      double dR[9], dT[3];
      for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) dR[r * 3 + c] = dt[r * 4 + c];
        dT[r] = dt[r * 4 + 3];
      }
      float Tn[3], Rn[9];
      for (int r = 0; r < 3; ++r)
        Tn[r] = (float)((double)Rf[r * 3 + 0] * dT[0] + (double)Rf[r * 3 + 1] * dT[1] + (double)Rf[r * 3 + 2] * dT[2] + (double)Tf[r]);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          Rn[r * 3 + c] = (float)((double)Rf[r * 3 + 0] * dR[0 * 3 + c] + (double)Rf[r * 3 + 1] * dR[1 * 3 + c] + (double)Rf[r * 3 + 2] * dR[2 * 3 + c]);
      std::memcpy(Rf, Rn, sizeof(Rf));
      std::memcpy(Tf, Tn, sizeof(Tf));
      for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) dRT[r * 4 + c] = dR[r * 3 + c];
        dRT[r * 4 + 3] = dT[r];
      }
    } else {
      exp_se3_f64(unit6, step, dRT);
      double Rn[9], Tn[3];
      for (int r = 0; r < 3; ++r) {
        Tn[r] = Td[r];
        for (int c = 0; c < 3; ++c) Tn[r] += Rd[r * 3 + c] * dRT[c * 4 + 3];
        for (int c = 0; c < 3; ++c) {
          double s = 0;
          for (int m = 0; m < 3; ++m) s += Rd[r * 3 + m] * dRT[m * 4 + c];
          Rn[r * 3 + c] = s;
        }
      }
      std::memcpy(Rd, Rn, sizeof(Rd));
      std::memcpy(Td, Tn, sizeof(Td));
    }
    double dist = log_se3_norm(dRT);
    rec.dist = (float)dist;
    float ip = (float)((double)o.cnt.stored / std::sqrt((double)N * (double)M));
    rec.ip = ip;
    bool dec = float_ell ? ind32.update(ip, p->indicator_window_size, p->indicator_stable_threshold)
                                       : ind64.update((double)ip, p->indicator_window_size, (double)p->indicator_stable_threshold);
    if (trace && k < trace_cap) trace[k] = rec;
    if (dist < p->eps_2) break;
    if (k > p->ell_decay_start && dec) {
      ell = ell * p->ell_decay_rate;
      if (ell < p->ell_min) ell = p->ell_min;
      ell64 = ell64 * (double)p->ell_decay_rate;
      if (ell64 < (double)p->ell_min) ell64 = p->ell_min;
    }
    if (knn) K = knn_next_k(p, o.cnt.stored, N);
    else K = std::min(p->nearest_neighbors_max, (int)(o.cnt.max_row * 1.2));
  }
  if (iters) *iters = k;
  if (ret_code) *ret_code = ret;
  // final update_tf: transform = [R^T, -R^T T]
  std::memset(T_out, 0, sizeof(float) * 16);
  T_out[15] = 1.0f;
  if (float_pose) {
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) T_out[r * 4 + c] = Rf[c * 3 + r];
    for (int r = 0; r < 3; ++r)
      T_out[r * 4 + 3] = -(T_out[r * 4 + 0] * Tf[0] + T_out[r * 4 + 1] * Tf[1] + T_out[r * 4 + 2] * Tf[2]);
  } else {
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) T_out[r * 4 + c] = (float)Rd[c * 3 + r];
      T_out[r * 4 + 3] = (float)(-(Rd[0 * 3 + r] * Td[0] + Rd[1 * 3 + r] * Td[1] + Rd[2 * 3 + r] * Td[2]));
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Adaptive-lengthscale CVO (AdaptiveCvoGPU::align, AdaptiveCvoGPU.cu:1077-1248)
// ---------------------------------------------------------------------------

namespace {

// fill_in_A_mat_gpu of AdaptiveCvoGPU.cu:263-408: rows of X against Y in
// ascending target index, the first K that pass: distance gate at the fixed
// lengthscale, colour gate always, class factor without a gate (when
// is_using_semantics), no geometric type; each factor an FP64 exp rounded to
// float.  Also the per-row sums of a |y - x|^2 (float row sums for the
// mirror, FP64 of the float products otherwise).
struct AcvoRows {
  std::vector<int32_t> counts, cols;
  std::vector<float> vals;
  std::vector<double> ad2;  // per row
  int64_t nnz = 0;
};

void acvo_assoc(const cvo_params* p, int N, const float* X, const float* XF, const float* XS, int M, const float* Y,
                const float* YF, const float* YS, int F, int C, float ell, int K, int precision, AcvoRows& o) {
  const float s2 = p->sigma * p->sigma;
  const float l = ell;
  const float d2_thres = (float)(-2.0 * (double)l * (double)l * (double)logf(p->sp_thres / s2));
  const float d2_c_thres = (float)(-2.0 * (double)p->c_ell * (double)p->c_ell *
                                   (double)logf(p->sp_thres / p->c_sigma / p->c_sigma));
  const bool sem = p->is_using_semantics && C > 0 && XS && YS;
  o.counts.assign(N, 0);
  o.cols.assign((size_t)N * K, -1);
  o.vals.assign((size_t)N * K, 0.0f);
  o.ad2.assign(N, 0.0);
  int64_t nnz = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nnz)
  for (int i = 0; i < N; ++i) {
    const float* x = X + 3 * (size_t)i;
    int n = 0;
    float rowf = 0.0f;
    double rowd = 0.0;
    for (int j = 0; j < M && n < K; ++j) {
      const float* y = Y + 3 * (size_t)j;
      const float d2 = sqdist3_f(y, x);
      if (!(d2 < d2_thres)) continue;
      const float d2c = sqdist_n_f(XF + (size_t)i * F, YF + (size_t)j * F, F);
      if (!(d2c < d2_c_thres)) continue;
      const float k = (float)((double)s2 * std::exp(-(double)d2 / (2.0 * (double)l * (double)l)));
      const float ck = (float)((double)(p->c_sigma * p->c_sigma) *
                               std::exp(-(double)d2c / (2.0 * (double)p->c_ell * (double)p->c_ell)));
      float sk = 1.0f;
      if (sem) {
        const float d2s = sqdist_n_f(XS + (size_t)i * C, YS + (size_t)j * C, C);
        sk = (float)((double)(p->s_sigma * p->s_sigma) *
                     std::exp(-(double)d2s / (2.0 * (double)p->s_ell * (double)p->s_ell)));
      }
      const float a = ck * k * sk;
      if (!(a > p->sp_thres)) continue;
      o.cols[(size_t)i * K + n] = j;
      o.vals[(size_t)i * K + n] = a;
      n++;
      rowf += d2 * a;
      rowd += (double)(d2 * a);
    }
    o.counts[i] = n;
    o.ad2[i] = precision == ORC_MIRROR ? (double)rowf : rowd;
    nnz += n;
  }
  o.nnz = nnz;
}

// Step of the adaptive loop (AdaptiveCvoGPU.cu:1039-1074): the cubic
// 4E s^3 + 3D s^2 + 2C s + B with float coefficients, its smallest positive
// real root, min_step when there is none, at most 0.8.
float adaptive_select_step(const cvo_params* p, const double bcde[4]) {
  const double cf[4] = {(double)(4.0f * (float)bcde[3]), (double)(3.0f * (float)bcde[2]),
                        (double)(2.0f * (float)bcde[1]), (double)(float)bcde[0]};
  double re[3], im[3];
  cubic_roots(cf, re, im);
  double temp = std::numeric_limits<double>::max();
  for (int i = 0; i < 3; ++i)
    if (std::isfinite(re[i]) && re[i] > 0 && re[i] < temp && im[i] == 0.0) temp = re[i];
  float step = temp == std::numeric_limits<double>::max() ? p->min_step : (float)temp;
  return step > 0.8f ? 0.8f : step;
}

}  // namespace

extern "C" float orc_select_step_adaptive(const cvo_params* p, const double bcde[4]) {
  return adaptive_select_step(p, bcde);
}

extern "C" int orc_assoc_adaptive(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                  const float T[16], float ell, int32_t K, int32_t precision, int32_t* counts,
                                  int32_t* cols, float* vals, double* ad2) {
  if (!p || !src || !tgt || !T || K <= 0 || !src->feat || !tgt->feat) return 1;
  const int M = tgt->n;
  std::vector<float> yp((size_t)M * 3);
  float Am[9], b[3];
  applied_transform(pose_from16(T), Am, b);
  for (int j = 0; j < M; ++j) xform_f(Am, b, tgt->xyz + 3 * (size_t)j, &yp[3 * (size_t)j]);
  AcvoRows o;
  const int C = p->is_using_semantics ? src->num_classes : 0;
  acvo_assoc(p, src->n, src->xyz, src->feat, src->sem, M, yp.data(), tgt->feat, tgt->sem, src->feat_dim, C, ell, K,
             precision, o);
  if (counts) std::memcpy(counts, o.counts.data(), sizeof(int32_t) * o.counts.size());
  if (cols) std::memcpy(cols, o.cols.data(), sizeof(int32_t) * o.cols.size());
  if (vals) std::memcpy(vals, o.vals.data(), sizeof(float) * o.vals.size());
  if (ad2) std::memcpy(ad2, o.ad2.data(), sizeof(double) * o.ad2.size());
  return 0;
}

extern "C" double orc_inner_product_adaptive(const cvo_params* p, const cvo_cloud_view* src,
                                             const cvo_cloud_view* tgt, const float T_s2t[16]) {
  if (!p || !src || !tgt || !T_s2t || !src->feat || !tgt->feat) return -1.0;
  if (src->n == 0 || tgt->n == 0) return 0.0;
  // the target moved by T itself: the applied map of its inverse (float, as
  // the library forms it)
  float Ti[16];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) Ti[r * 4 + c] = T_s2t[c * 4 + r];
    Ti[r * 4 + 3] = (float)(-((double)T_s2t[0 * 4 + r] * T_s2t[3] + (double)T_s2t[1 * 4 + r] * T_s2t[7] +
                              (double)T_s2t[2 * 4 + r] * T_s2t[11]));
  }
  Ti[12] = Ti[13] = Ti[14] = 0.0f;
  Ti[15] = 1.0f;
  const int M = tgt->n;
  std::vector<float> yp((size_t)M * 3);
  float Am[9], b[3];
  applied_transform(pose_from16(Ti), Am, b);
  for (int j = 0; j < M; ++j) xform_f(Am, b, tgt->xyz + 3 * (size_t)j, &yp[3 * (size_t)j]);
  AcvoRows o;
  const int C = p->is_using_semantics ? src->num_classes : 0;
  acvo_assoc(p, src->n, src->xyz, src->feat, src->sem, M, yp.data(), tgt->feat, tgt->sem, src->feat_dim, C,
             p->ell_init, M, ORC_MIRROR, o);
  double sum = 0.0;
  for (int i = 0; i < src->n; ++i)
    for (int q = 0; q < o.counts[i]; ++q) sum += (double)o.vals[(size_t)i * M + q];
  return (double)(float)(sum / (double)src->n * 1e6 / (double)M);
}

extern "C" int orc_align_adaptive(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                  const float T_init[16], int32_t precision, float T_out[16], int32_t* iters,
                                  int32_t* ret_code, cvo_iter_record* trace, int32_t trace_cap) {
  if (!p || !src || !tgt || !T_init || !T_out || !src->feat || !tgt->feat) return 1;
  if (precision != ORC_MIRROR && precision != ORC_POLICY) return 1;
  const int N = src->n, M = tgt->n, F = src->feat_dim;
  const int C = p->is_using_semantics ? src->num_classes : 0;
  const int K = std::min(kKnnK, p->nearest_neighbors_max);
  const bool float_pose = precision == ORC_MIRROR;
  cvo_params ps = *p;
  ps.is_using_range_ell = 0;  // the step's 1/(2 l^2) uses the lengthscale as is
  float Rf[9], Tf[3];
  double Rd[9], Td[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) Rf[r * 3 + c] = Rd[r * 3 + c] = T_init[r * 4 + c];
    Tf[r] = Td[r] = T_init[r * 4 + 3];
  }
  float ell = p->ell_init, ell_max = p->ell_max;
  int ret = 0, k = 0;
  std::vector<float> yp((size_t)M * 3);
  AcvoRows A, Axx, Ayy;
  for (; k < p->MAX_ITER; ++k) {
    Pose64 P;
    for (int i = 0; i < 9; ++i) P.R[i] = float_pose ? (double)Rf[i] : Rd[i];
    for (int i = 0; i < 3; ++i) P.T[i] = float_pose ? (double)Tf[i] : Td[i];
    float Am[9], b[3];
    applied_transform(P, Am, b);
    for (int j = 0; j < M; ++j) xform_f(Am, b, tgt->xyz + 3 * (size_t)j, &yp[3 * (size_t)j]);
    acvo_assoc(p, N, src->xyz, src->feat, src->sem, M, yp.data(), tgt->feat, tgt->sem, F, C, ell, K, precision, A);
    acvo_assoc(p, N, src->xyz, src->feat, src->sem, N, src->xyz, src->feat, src->sem, F, C, ell, K, precision, Axx);
    acvo_assoc(p, M, yp.data(), tgt->feat, tgt->sem, M, yp.data(), tgt->feat, tgt->sem, F, C, ell, K, precision, Ayy);
    double flow6[6], unit6[6];
    flow_core(p, N, src->xyz, yp.data(), K, A.counts.data(), A.cols.data(), A.vals.data(), precision, flow6, unit6);
    float xf[6];
    for (int d = 0; d < 6; ++d) xf[d] = (float)flow6[d];
    double bcde[4];
    step_core(&ps, ell, N, src->xyz, yp.data(), K, A.counts.data(), A.cols.data(), A.vals.data(), flow6, precision,
              bcde);
    const double nw = std::sqrt((double)xf[0] * xf[0] + (double)xf[1] * xf[1] + (double)xf[2] * xf[2]);
    const double nv = std::sqrt((double)xf[3] * xf[3] + (double)xf[4] * xf[4] + (double)xf[5] * xf[5]);
    cvo_iter_record rec{};
    rec.ell = ell;
    rec.nnz = (int32_t)A.nnz;
    rec.num_neighbors = K;
    rec.flow_norm = (float)std::sqrt(nw * nw + nv * nv);
    if (nw < p->eps && nv < p->eps) {
      if (nw < 1e-8 && nv < 1e-8) ret = -1;
      if (trace && k < trace_cap) trace[k] = rec;
      break;
    }
    // step: coefficients in float, smallest positive real root, min_step if
    // none, at most 0.8 (AdaptiveCvoGPU.cu:1039-1074)
    const float step = adaptive_select_step(p, bcde);
    rec.step = step;
    double dRT[12];
    if (float_pose) {
      float dt[12];
      exp_sek3_f32(xf, step, dt);
      float Tn[3], Rn[9];
      for (int r = 0; r < 3; ++r)
        Tn[r] = (float)((double)Rf[r * 3 + 0] * dt[3] + (double)Rf[r * 3 + 1] * dt[7] + (double)Rf[r * 3 + 2] * dt[11] +
                        (double)Tf[r]);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          Rn[r * 3 + c] = (float)((double)Rf[r * 3 + 0] * dt[c] + (double)Rf[r * 3 + 1] * dt[4 + c] +
                                  (double)Rf[r * 3 + 2] * dt[8 + c]);
      std::memcpy(Rf, Rn, sizeof(Rf));
      std::memcpy(Tf, Tn, sizeof(Tf));
      for (int i = 0; i < 12; ++i) dRT[i] = dt[i];
    } else {
      double xi[6];
      for (int d = 0; d < 6; ++d) xi[d] = (double)xf[d];
      exp_se3_f64(xi, (double)step, dRT);
      double Rn[9], Tn[3];
      for (int r = 0; r < 3; ++r) {
        Tn[r] = Td[r];
        for (int c = 0; c < 3; ++c) Tn[r] += Rd[r * 3 + c] * dRT[c * 4 + 3];
        for (int c = 0; c < 3; ++c) {
          double sacc = 0;
          for (int m = 0; m < 3; ++m) sacc += Rd[r * 3 + m] * dRT[m * 4 + c];
          Rn[r * 3 + c] = sacc;
        }
      }
      std::memcpy(Rd, Rn, sizeof(Rd));
      std::memcpy(Td, Tn, sizeof(Td));
    }
    const double dist = log_se3_norm(dRT);
    rec.dist = (float)dist;
    rec.ip = (float)((double)A.nnz / std::sqrt((double)N * (double)M));
    if (trace && k < trace_cap) trace[k] = rec;
    if (dist < p->eps_2) break;
    // the lengthscale gradient (AdaptiveCvoGPU.cu:548-720, 839-890)
    const float l3 = ell * ell * ell;
    double sa = 0, sxx = 0, syy = 0;
    for (int i = 0; i < N; ++i) sa += A.ad2[i];
    for (int i = 0; i < N; ++i) sxx += Axx.ad2[i];
    for (int i = 0; i < M; ++i) syy += Ayy.ad2[i];
    float dl_a, dl_yy;
    if (precision == ORC_MIRROR) {
      // per row: 2/l^3 * (-sum_A) then + 1/l^3 * sum_Axx, in float; rows summed in FP64
      double ta = 0, ty = 0;
      for (int i = 0; i < N; ++i)
        ta += (double)(2.0f / l3 * (float)(-A.ad2[i])) + (double)(1.0f / l3 * (float)Axx.ad2[i]);
      for (int i = 0; i < M; ++i) ty += (double)(1.0f / l3 * (float)Ayy.ad2[i]);
      dl_a = (float)ta;
      dl_yy = (float)ty;
    } else {
      dl_a = (float)((-2.0 * sa + sxx) / (double)l3);
      dl_yy = (float)(syy / (double)l3);
    }
    const uint32_t denom = (uint32_t)Axx.nnz + (uint32_t)Ayy.nnz - 2u * (uint32_t)A.nnz;
    const double dl = (double)(dl_a + dl_yy) / (double)denom;
    float e = (float)((double)ell - p->dl_step * dl);
    if (e >= ell_max) {
      e = (float)((double)ell_max * 0.7);
      ell_max = (float)((double)ell_max * 0.7);
    }
    ell = e < p->ell_min ? p->ell_min : e;
  }
  if (iters) *iters = k;
  if (ret_code) *ret_code = ret;
  std::memset(T_out, 0, sizeof(float) * 16);
  T_out[15] = 1.0f;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) T_out[r * 4 + c] = float_pose ? Rf[c * 3 + r] : (float)Rd[c * 3 + r];
    T_out[r * 4 + 3] = float_pose ? -(Rf[0 * 3 + r] * Tf[0] + Rf[1 * 3 + r] * Tf[1] + Rf[2 * 3 + r] * Tf[2])
                                  : (float)(-(Rd[0 * 3 + r] * Td[0] + Rd[1 * 3 + r] * Td[1] + Rd[2 * 3 + r] * Td[2]));
  }
  return 0;
}

extern "C" double orc_inner_product(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                    const float T[16], float ell, int32_t precision) {
  AssocOut o;
  int K = p->nearest_neighbors_max;
  assoc_core(p, ref(src), ref(tgt), pose_from16(T), ell, K, precision, 0, o);
  double s = 0;
  for (int i = 0; i < src->n; ++i)
    for (int q = 0; q < o.counts[i]; ++q) s += (double)o.vals[(size_t)i * K + q];
  return s;
}

extern "C" double orc_time_iterations(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                      const float T[16], float ell, int32_t K, int32_t n_iter, int32_t threads) {
  int old = omp_get_max_threads();
  if (threads > 0) omp_set_num_threads(threads);
  CloudRef S = ref(src), Tg = ref(tgt);
  Pose64 P = pose_from16(T);
  std::vector<float> yp((size_t)Tg.n * 3);
  auto t0 = std::chrono::steady_clock::now();
  for (int it = 0; it < n_iter; ++it) {
    float A[9], b[3];
    applied_transform(P, A, b);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Tg.n; ++j) xform_f(A, b, Tg.xyz + 3 * (size_t)j, &yp[3 * (size_t)j]);
    AssocOut o;
    assoc_core(p, S, Tg, P, ell, K, ORC_MIRROR, 0, o);
    double sum6[6], unit6[6], bcde[4];
    flow_core(p, S.n, S.xyz, yp.data(), K, o.counts.data(), o.cols.data(), o.vals.data(), ORC_MIRROR, sum6, unit6);
    step_core(p, ell, S.n, S.xyz, yp.data(), K, o.counts.data(), o.cols.data(), o.vals.data(), unit6, ORC_MIRROR, bcde);
    volatile double sink = select_step(p, bcde);
    (void)sink;
  }
  auto t1 = std::chrono::steady_clock::now();
  omp_set_num_threads(old);
  return std::chrono::duration<double>(t1 - t0).count() / std::max(1, n_iter);
}

// ===========================================================================
// IRLS bundle adjustment (src/cvo/IRLS.cpp:148-367, IRLS_State_GPU.cu:144-336,
// IRLS_State_GPU.cpp:10-146, IRLS_Cost_CPU.hpp:266-371,
// local_parameterization_se3.hpp:21-99).  The Ceres solve is replaced by the
// documented Levenberg-Marquardt schedule shared with the device solver:
//   D = clamp(diag(H), 1e-6, 1e32); solve (H + D/mu) d = -g;
//   rho = (F - F_new) / (-(g.d + d.H.d/2)); accept if rho > 1e-3:
//     mu /= max(1/3, 1 - (2 rho - 1)^3), v = 2;  else mu /= v, v *= 2;
//   mu starts at 1e4; stop on max iterations, |dF|/F < 1e-5, ||g||inf < 1e-5,
//   ||d|| < 1e-5 (||x|| + 1e-5).
// ===========================================================================

namespace {

struct Mat34 {
  double m[12];
  double& operator[](int i) { return m[i]; }
  const double& operator[](int i) const { return m[i]; }
  operator double*() { return m; }
  operator const double*() const { return m; }
};

// T * Exp([rho; phi]) with the translation part first (Exp_SE3(.., false),
// LieGroup.cpp:168-192), exact small-angle series.
void se3_plus(const double T[12], const double d[6], double out[12]) {
  double xi[6] = {d[3], d[4], d[5], d[0], d[1], d[2]};  // [w; v]
  double E[12];
  exp_se3_f64(xi, 1.0, E);
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c)
      out[r * 4 + c] = T[r * 4 + 0] * E[0 * 4 + c] + T[r * 4 + 1] * E[1 * 4 + c] + T[r * 4 + 2] * E[2 * 4 + c];
    out[r * 4 + 3] = T[r * 4 + 0] * E[3] + T[r * 4 + 1] * E[7] + T[r * 4 + 2] * E[11] + T[r * 4 + 3];
  }
}

// LocalParameterizationSE3::ComputeJacobian (local_parameterization_se3.hpp:49-83)
void plus_jacobian(const double T[12], double J[72]) {
  for (int i = 0; i < 72; ++i) J[i] = 0.0;
  const double d1[3] = {T[0], T[4], T[8]}, d2[3] = {T[1], T[5], T[9]}, d3[3] = {T[2], T[6], T[10]};
  auto S = [&](int r, int c, double v) { J[r * 6 + c] = v; };
  S(0, 4, -d3[0]); S(0, 5, d2[0]); S(4, 4, -d3[1]); S(4, 5, d2[1]); S(8, 4, -d3[2]); S(8, 5, d2[2]);
  S(1, 3, d3[0]); S(1, 5, -d1[0]); S(5, 3, d3[1]); S(5, 5, -d1[1]); S(9, 3, d3[2]); S(9, 5, -d1[2]);
  S(2, 3, -d2[0]); S(2, 4, d1[0]); S(6, 3, -d2[1]); S(6, 4, d1[1]); S(10, 3, -d2[2]); S(10, 4, d1[2]);
  for (int c = 0; c < 3; ++c) {
    S(3, c, T[0 * 4 + c]);
    S(7, c, T[1 * 4 + c]);
    S(11, c, T[2 * 4 + c]);
  }
}

// PairwiseAnalyticalDiffFunctor::Evaluate (IRLS_Cost_CPU.hpp:305-357)
void pair_residual(const double T1[12], const double T2[12], const double p1[3], const double p2[3], double a,
                   double* r, double J1[12], double J2[12]) {
  double P1[3], P2[3], diff[3];
  for (int i = 0; i < 3; ++i) {
    P1[i] = T1[i * 4 + 0] * p1[0] + T1[i * 4 + 1] * p1[1] + T1[i * 4 + 2] * p1[2] + T1[i * 4 + 3];
    P2[i] = T2[i * 4 + 0] * p2[0] + T2[i * 4 + 1] * p2[1] + T2[i * 4 + 2] * p2[2] + T2[i * 4 + 3];
    diff[i] = P1[i] - P2[i];
  }
  *r = a * (diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
  const double h1[4] = {p1[0], p1[1], p1[2], 1.0}, h2[4] = {p2[0], p2[1], p2[2], 1.0};
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 4; ++k) {
      J1[i * 4 + k] = diff[i] * h1[k];   // diff^T DT1
      J2[i * 4 + k] = -diff[i] * h2[k];  // -diff^T DT2
    }
}

struct Pair {
  int e;
  float p1[3], p2[3];
  double a;
  int i1 = -1, i2 = -1;  // point indices in the edge's two frames
};

// Local (6-dof) Jacobians of one pair: J_k(12) * ComputeJacobian(T_k)(12x6).
void local_jacobians(const double T1[12], const double T2[12], const Pair& q, double* r, double j1[6], double j2[6]) {
  double p1[3] = {q.p1[0], q.p1[1], q.p1[2]}, p2[3] = {q.p2[0], q.p2[1], q.p2[2]};
  double J1[12], J2[12], P1j[72], P2j[72];
  pair_residual(T1, T2, p1, p2, (double)q.a, r, J1, J2);
  plus_jacobian(T1, P1j);
  plus_jacobian(T2, P2j);
  for (int c = 0; c < 6; ++c) {
    double s1 = 0, s2 = 0;
    for (int k = 0; k < 12; ++k) {
      s1 += J1[k] * P1j[k * 6 + c];
      s2 += J2[k] * P2j[k * 6 + c];
    }
    j1[c] = s1;
    j2[c] = s2;
  }
}

// Dense normal equations over free frames: H (6F x 6F), g (6F), cost = 1/2 sum r^2.
void normal_equations(int n_frames, const std::vector<Mat34>& poses, const std::vector<int>& free_idx, int nfree,
                      const std::vector<std::pair<int, int>>& edge_frames, const std::vector<char>& edge_on,
                      const std::vector<Pair>& pairs, std::vector<double>& H, std::vector<double>& g, double* cost,
                      int threads) {
  const int n = 6 * nfree;
  H.assign((size_t)n * n, 0.0);
  g.assign(n, 0.0);
  double c = 0.0;
  int nt = std::max(1, threads);
  std::vector<std::vector<double>> Hs(nt, std::vector<double>((size_t)n * n, 0.0)), gs(nt, std::vector<double>(n, 0.0));
  std::vector<double> cs(nt, 0.0);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (size_t k = 0; k < pairs.size(); ++k) {
    const Pair& q = pairs[k];
    if (!edge_on[q.e]) continue;
    int t = omp_get_thread_num();
    int f1 = edge_frames[q.e].first, f2 = edge_frames[q.e].second;
    double r, j1[6], j2[6];
    local_jacobians(poses[f1], poses[f2], q, &r, j1, j2);
    cs[t] += 0.5 * r * r;
    int o[2] = {free_idx[f1], free_idx[f2]};
    const double* js[2] = {j1, j2};
    for (int u = 0; u < 2; ++u) {
      if (o[u] < 0) continue;
      for (int a = 0; a < 6; ++a) {
        gs[t][6 * o[u] + a] += js[u][a] * r;
        for (int w = 0; w < 2; ++w) {
          if (o[w] < 0) continue;
          for (int b = 0; b < 6; ++b) Hs[t][(size_t)(6 * o[u] + a) * n + 6 * o[w] + b] += js[u][a] * js[w][b];
        }
      }
    }
  }
  for (int t = 0; t < nt; ++t) {
    for (size_t i = 0; i < H.size(); ++i) H[i] += Hs[t][i];
    for (int i = 0; i < n; ++i) g[i] += gs[t][i];
    c += cs[t];
  }
  *cost = c;
}

bool cholesky_solve(std::vector<double> A, int n, const std::vector<double>& b, std::vector<double>& x) {
  for (int j = 0; j < n; ++j) {
    double s = A[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) s -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
    if (!(s > 0)) return false;
    double d = std::sqrt(s);
    A[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      double t = A[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) t -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
      A[(size_t)i * n + j] = t / d;
    }
  }
  x = b;
  for (int i = 0; i < n; ++i) {
    double t = x[i];
    for (int k = 0; k < i; ++k) t -= A[(size_t)i * n + k] * x[k];
    x[i] = t / A[(size_t)i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double t = x[i];
    for (int k = i + 1; k < n; ++k) t -= A[(size_t)k * n + i] * x[k];
    x[i] = t / A[(size_t)i * n + i];
  }
  return true;
}

// change_of_all_poses (IRLS.cpp:99-108): sum ||log(T_old^-1 T_new)||
double pose_change(const std::vector<Mat34>& a, const std::vector<Mat34>& b) {
  double s = 0;
  for (size_t f = 0; f < a.size(); ++f) {
    double D[12];
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c)
        D[r * 4 + c] = a[f][0 * 4 + r] * b[f][0 * 4 + c] + a[f][1 * 4 + r] * b[f][1 * 4 + c] + a[f][2 * 4 + r] * b[f][2 * 4 + c];
      D[r * 4 + 3] = a[f][0 * 4 + r] * (b[f][3] - a[f][3]) + a[f][1 * 4 + r] * (b[f][7] - a[f][7]) +
                     a[f][2 * 4 + r] * (b[f][11] - a[f][11]);
    }
    s += log_se3_norm(D);
  }
  return s;
}

// World points of a frame in float (CvoFrameGPU::transform_pointcloud, float
// pose, canonical FMA order).
std::vector<float> world_points(const cvo_cloud_view& v, const double T[12]) {
  float A[9], b[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) A[r * 3 + c] = (float)T[r * 4 + c];
    b[r] = (float)T[r * 4 + 3];
  }
  std::vector<float> out((size_t)v.n * 3);
  for (int j = 0; j < v.n; ++j) xform_f(A, b, v.xyz + 3 * (size_t)j, &out[3 * (size_t)j]);
  return out;
}

}  // namespace

extern "C" int orc_ba_pair_residual(const double T1[12], const double T2[12], const double p1[3], const double p2[3],
                                    double a, double* r, double J1[12], double J2[12]) {
  pair_residual(T1, T2, p1, p2, a, r, J1, J2);
  return 0;
}

extern "C" int orc_se3_plus(const double T[12], const double delta[6], double out[12]) {
  se3_plus(T, delta, out);
  return 0;
}

extern "C" int orc_se3_plus_jacobian(const double T[12], double J[72]) {
  plus_jacobian(T, J);
  return 0;
}

extern "C" int orc_ba_normal_equations(int32_t n_frames, const double* poses, int32_t n_pairs,
                                       const int32_t* pair_frames, const float* p1, const float* p2, const float* a,
                                       double* H, double* g, double* cost) {
  std::vector<Mat34> P(n_frames);
  for (int f = 0; f < n_frames; ++f)
    for (int i = 0; i < 12; ++i) P[f][i] = poses[12 * f + i];
  std::vector<int> free_idx(n_frames);
  for (int f = 0; f < n_frames; ++f) free_idx[f] = f;
  // one "edge" per distinct (f1, f2) is not needed here: every pair carries its frames
  std::vector<std::pair<int, int>> ef(n_pairs);
  std::vector<char> on(n_pairs, 1);
  std::vector<Pair> pairs(n_pairs);
  for (int k = 0; k < n_pairs; ++k) {
    ef[k] = {pair_frames[2 * k], pair_frames[2 * k + 1]};
    pairs[k].e = k;
    for (int i = 0; i < 3; ++i) {
      pairs[k].p1[i] = p1[3 * k + i];
      pairs[k].p2[i] = p2[3 * k + i];
    }
    pairs[k].a = a[k];
  }
  std::vector<double> Hv, gv;
  normal_equations(n_frames, P, free_idx, n_frames, ef, on, pairs, Hv, gv, cost, 1);
  std::memcpy(H, Hv.data(), sizeof(double) * Hv.size());
  std::memcpy(g, gv.data(), sizeof(double) * gv.size());
  return 0;
}

namespace {

// BinaryStateCPU::update_inner_product + se_kernel (IRLS_State_CPU.cpp:146-309):
// frame 2 mapped into frame 1 by pose1^-1 pose2 (float), every mapped frame-2
// point radius-searched over the frame-1 points at the edge's fixed lengthscale
// (no range-ell, no row cap).  Gates: geometric type dot / (|a| |b|) >= 0.01,
// distance, colour, semantics; kept when ck k sk geo > sp_thres, with weight
// ck k (the semantic and geometric-type factors are not in the weight,
// IRLS_State_CPU.cpp:261).  Pairs (i in frame 1, idx in frame 2) in sensor
// coordinates.
void cpu_state_pairs(const cvo_params* p, const cvo_cloud_view& c1, const cvo_cloud_view& c2, const double X1[12],
                     const double X2[12], float ell, int e, std::vector<Pair>& out, int64_t* nnz) {
  // f1_to_f2 = pose1^-1 pose2
  double R1[9], t1[3], R2[9], t2[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) {
      R1[r * 3 + c] = X1[r * 4 + c];
      R2[r * 3 + c] = X2[r * 4 + c];
    }
    t1[r] = X1[r * 4 + 3];
    t2[r] = X2[r * 4 + 3];
  }
  float A[9], b[3];
  for (int r = 0; r < 3; ++r) {
    double bb = 0;
    for (int c = 0; c < 3; ++c) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += R1[k * 3 + r] * R2[k * 3 + c];
      A[r * 3 + c] = (float)s;
      bb += R1[c * 3 + r] * (t2[c] - t1[c]);
    }
    b[r] = (float)bb;
  }
  std::vector<float> y((size_t)c2.n * 3);
  for (int j = 0; j < c2.n; ++j) {
    const float* q = c2.xyz + 3 * (size_t)j;
    for (int r = 0; r < 3; ++r) y[3 * (size_t)j + r] = A[r * 3 + 0] * q[0] + A[r * 3 + 1] * q[1] + A[r * 3 + 2] * q[2] + b[r];
  }
  const float sigma2 = p->sigma * p->sigma, c_ell2 = p->c_ell * p->c_ell, s_ell2 = p->s_ell * p->s_ell;
  const float c_sigma2 = p->c_sigma * p->c_sigma, s_sigma2 = p->s_sigma * p->s_sigma;
  float d2_thresh = 1, d2_c_thresh = 1, d2_s_thresh = 1;
  if (p->is_using_geometry) d2_thresh = (float)(-2.0 * ell * ell * std::log((double)(p->sp_thres / sigma2)));
  if (p->is_using_intensity) d2_c_thresh = (float)(-2.0 * c_ell2 * std::log((double)(p->sp_thres / c_sigma2)));
  if (p->is_using_semantics) d2_s_thresh = (float)(-2.0 * s_ell2 * std::log((double)(p->sp_thres / s_sigma2)));
  Grid grid;
  const double r = std::sqrt((double)d2_thresh);
  grid.build(c1.xyz, c1.n, std::max(r, 1e-3));
  std::vector<std::vector<Pair>> per(c2.n);
#pragma omp parallel for schedule(dynamic, 64)
  for (int idx = 0; idx < c2.n; ++idx) {
    const float* q = &y[3 * (size_t)idx];
    std::vector<std::pair<int, float>> m;
    grid.query(q, r, [&](int i) {
      const float* x = c1.xyz + 3 * (size_t)i;
      float d2 = 0;
      for (int d = 0; d < 3; ++d) d2 += (x[d] - q[d]) * (x[d] - q[d]);
      if (d2 < d2_thresh) m.push_back({i, d2});
    });
    std::sort(m.begin(), m.end());
    for (auto& mm : m) {
      const int i = mm.first;
      const float d2 = mm.second;
      float sk = 1, ck = 1, k = 1, geo_sim = 1;
      if (p->is_using_geometric_type && c1.geo_type && c2.geo_type) {
        const float* ga = c1.geo_type + 2 * (size_t)i;
        const float* gb = c2.geo_type + 2 * (size_t)idx;
        geo_sim = (ga[0] * gb[0] + ga[1] * gb[1]) / std::sqrt(ga[0] * ga[0] + ga[1] * ga[1]) /
                  std::sqrt(gb[0] * gb[0] + gb[1] * gb[1]);
        if (geo_sim < 0.01f) continue;
      }
      if (p->is_using_geometry) k = (float)(sigma2 * std::exp(-d2 / (2.0 * ell * ell)));
      if (p->is_using_intensity) {
        const float d2c = sqdist_n_f(c1.feat + (size_t)i * c1.feat_dim, c2.feat + (size_t)idx * c2.feat_dim, c1.feat_dim);
        if (!(d2c < d2_c_thresh)) continue;
        ck = (float)(c_sigma2 * std::exp(-d2c / (2.0 * c_ell2)));
      }
      if (p->is_using_semantics) {
        const float d2s =
            sqdist_n_f(c1.sem + (size_t)i * c1.num_classes, c2.sem + (size_t)idx * c2.num_classes, c1.num_classes);
        if (!(d2s < d2_s_thresh)) continue;
        sk = (float)(s_sigma2 * std::exp(-d2s / (2.0 * s_ell2)));
      }
      const float a = ck * k * sk * geo_sim;
      if (a > p->sp_thres) {
        Pair pr;
        pr.e = e;
        for (int d = 0; d < 3; ++d) {
          pr.p1[d] = c1.xyz[3 * (size_t)i + d];
          pr.p2[d] = c2.xyz[3 * (size_t)idx + d];
        }
        pr.a = (double)ck * k;
        pr.i1 = i;
        pr.i2 = idx;
        per[idx].push_back(pr);
      }
    }
  }
  int64_t cnt = 0;
  for (auto& v : per) {
    cnt += (int64_t)v.size();
    out.insert(out.end(), v.begin(), v.end());
  }
  *nnz = cnt;
}

int align_multiframe_impl(const cvo_params* p, int32_t n_frames, const cvo_cloud_view* frames, double* poses,
                          const uint8_t* hold_const, int32_t n_edges, const int32_t* edges, int32_t threads,
                          int32_t* iters_out, double* per_iter_seconds, int32_t max_records, bool cpu_states) {
  if (!p || n_frames <= 0 || !frames || !poses || n_edges < 0) return 1;
  int old = omp_get_max_threads();
  if (threads > 0) omp_set_num_threads(threads);
  std::vector<Mat34> X(n_frames);
  for (int f = 0; f < n_frames; ++f)
    for (int i = 0; i < 12; ++i) X[f][i] = poses[12 * f + i];
  std::vector<int> free_idx(n_frames, -1);
  int nfree = 0;
  for (int f = 0; f < n_frames; ++f)
    if (!hold_const || !hold_const[f]) free_idx[f] = nfree++;
  std::vector<std::pair<int, int>> ef(n_edges);
  for (int e = 0; e < n_edges; ++e) ef[e] = {edges[2 * e], edges[2 * e + 1]};
  // BinaryStateGPU state (IRLS_State_GPU.cu:18-69)
  std::vector<double> ell(n_edges, p->multiframe_ell_init);
  const double ell_max = (double)p->multiframe_ell_init * 1.2, ell_min = p->multiframe_ell_min;
  std::vector<int> K(n_edges, p->multiframe_num_neighbors), last_max(n_edges, 0), iters_per_ell(n_edges, 0);
  std::vector<int64_t> nnz_last(n_edges, 0), nnz(n_edges, 0);
  Pose64 I64;
  for (int i = 0; i < 9; ++i) I64.R[i] = (i % 4 == 0) ? 1.0 : 0.0;
  for (int i = 0; i < 3; ++i) I64.T[i] = 0.0;
  int it = 0;
  const int n = 6 * nfree;
  while (true) {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<Mat34> X_old = X;
    // world points (float) and per-edge association
    std::vector<std::vector<float>> W(n_frames);
    for (int f = 0; f < n_frames; ++f) W[f] = world_points(frames[f], X[f]);
    std::vector<Pair> pairs;
    int counter_nonzero = 0;
    std::vector<char> on(n_edges, 0);
    for (int e = 0; e < n_edges; ++e) {
      const int f1 = ef[e].first, f2 = ef[e].second;
      nnz_last[e] = nnz[e];
      if (cpu_states) {
        // BinaryStateCPU: every stored pair becomes a residual
        // (IRLS_State_CPU.cpp:48-81, no nonzero threshold per edge)
        cpu_state_pairs(p, frames[f1], frames[f2], X[f1].m, X[f2].m, (float)ell[e], e, pairs, &nnz[e]);
        if (nnz[e] > p->multiframe_min_nonzeros) counter_nonzero++;
        on[e] = nnz[e] > 0;
        continue;
      }
      if (last_max[e] > 0) K[e] = std::min(p->multiframe_num_neighbors, (int)(last_max[e] * 1.1));
      CloudRef S = ref(&frames[f1]), Tg = ref(&frames[f2]);
      S.xyz = W[f1].data();
      Tg.xyz = W[f2].data();
      AssocOut o;
      // BinaryStateGPU::update_inner_product: K nearest with is_using_kdtree
      // (IRLS_State_GPU.cu:171-215), the ball association otherwise
      if (p->is_using_kdtree) knn_assoc_core(p, S, Tg, I64, (float)ell[e], K[e], ORC_MIRROR, 0, o);
      else assoc_core(p, S, Tg, I64, (float)ell[e], K[e], ORC_MIRROR, 0, o);
      nnz[e] = o.cnt.stored;
      last_max[e] = o.cnt.max_row;
      if (std::getenv("ORC_BA_DEBUG"))
        std::fprintf(stderr, "it %d edge %d (%d,%d) ell %.6f K %d nnz %lld max_row %d\n", it, e, f1, f2, ell[e], K[e],
                     (long long)nnz[e], last_max[e]);
      if (nnz[e] > p->multiframe_min_nonzeros) counter_nonzero++;
      on[e] = (nnz[e] > 0 && nnz[e] >= p->multiframe_min_nonzeros);
      if (!on[e]) continue;
      for (int i = 0; i < S.n; ++i)
        for (int s = 0; s < o.counts[i]; ++s) {
          Pair q;
          q.e = e;
          int j = o.cols[(size_t)i * K[e] + s];
          for (int d = 0; d < 3; ++d) {
            q.p1[d] = frames[f1].xyz[3 * (size_t)i + d];
            q.p2[d] = frames[f2].xyz[3 * (size_t)j + d];
          }
          q.a = o.vals[(size_t)i * K[e] + s];
          pairs.push_back(q);
        }
    }
    if (counter_nonzero == 0) break;  // IRLS.cpp:249-250
    // Levenberg-Marquardt over the fixed pairs
    if (n > 0) {
      std::vector<double> H, g, Hc, gc;
      double F = 0, Fc = 0;
      normal_equations(n_frames, X, free_idx, nfree, ef, on, pairs, H, g, &F, threads);
      double mu = 1e4, v = 2.0;
      for (int lm = 0; lm < p->multiframe_iterations_per_solve; ++lm) {
        double gmax = 0;
        for (int i = 0; i < n; ++i) gmax = std::max(gmax, std::fabs(g[i]));
        if (gmax < 1e-5) break;
        std::vector<double> A = H, d;
        for (int i = 0; i < n; ++i) {
          double D = std::min(std::max(H[(size_t)i * n + i], 1e-6), 1e32);
          A[(size_t)i * n + i] += D / mu;
        }
        std::vector<double> mg(n);
        for (int i = 0; i < n; ++i) mg[i] = -g[i];
        if (!cholesky_solve(A, n, mg, d)) {
          mu /= v;
          v *= 2;
          continue;
        }
        double dn = 0, xn = 0;
        for (int i = 0; i < n; ++i) dn += d[i] * d[i];
        for (int f = 0; f < n_frames; ++f)
          for (int i = 0; i < 12; ++i) xn += X[f][i] * X[f][i];
        if (std::sqrt(dn) < 1e-5 * (std::sqrt(xn) + 1e-5)) break;
        std::vector<Mat34> Xc = X;
        for (int f = 0; f < n_frames; ++f)
          if (free_idx[f] >= 0) se3_plus(X[f], &d[6 * free_idx[f]], Xc[f]);
        normal_equations(n_frames, Xc, free_idx, nfree, ef, on, pairs, Hc, gc, &Fc, threads);
        double gd = 0, dHd = 0;
        for (int i = 0; i < n; ++i) {
          gd += g[i] * d[i];
          double s = 0;
          for (int j = 0; j < n; ++j) s += H[(size_t)i * n + j] * d[j];
          dHd += d[i] * s;
        }
        const double pred = -(gd + 0.5 * dHd);
        const double rho = pred > 0 ? (F - Fc) / pred : -1.0;
        if (std::getenv("ORC_BA_DEBUG"))
          std::fprintf(stderr, "it %d lm %d F %.9e Fc %.9e pred %.6e rho %.6e mu %.3e\n", it, lm, F, Fc, pred, rho, mu);
        if (rho > 1e-3) {
          const double rel = std::fabs(F - Fc) / std::max(F, 1e-300);
          X = Xc;
          H = Hc;
          g = gc;
          F = Fc;
          mu = mu / std::max(1.0 / 3.0, 1.0 - std::pow(2.0 * rho - 1.0, 3));
          mu = std::min(mu, 1e16);
          v = 2.0;
          if (rel < 1e-5) break;
        } else {
          mu /= v;
          v *= 2.0;
        }
      }
    }
    // per-edge lengthscale (BinaryStateGPU::update_ell, IRLS_State_GPU.cpp:117-146;
    // BinaryStateCPU::update_ell, IRLS_State_CPU.cpp:287-290: decay while above
    // the minimum, every outer iteration)
    bool all_min = true;
    for (int e = 0; e < n_edges; ++e) {
      if (cpu_states) {
        if (ell[e] > (double)p->multiframe_ell_min) ell[e] = ell[e] * (double)p->multiframe_ell_decay_rate;
        all_min = all_min && std::fabs((float)ell[e] - p->multiframe_ell_min) < 1e-4;
        continue;
      }
      iters_per_ell[e]++;
      if (nnz_last[e] > 0 && nnz[e] <= nnz_last[e] && iters_per_ell[e] / p->multiframe_iter_per_ell > 0) {
        ell[e] = ell[e] * (double)p->multiframe_ell_decay_rate;
        iters_per_ell[e] = 0;
      }
      if (ell[e] < ell_min) ell[e] = ell_min;
      else if (ell[e] > ell_max) ell[e] = ell_max;
      all_min = all_min && std::fabs((float)ell[e] - p->multiframe_ell_min) < 1e-4;
    }
    const double change = pose_change(X_old, X);
    auto t1 = std::chrono::steady_clock::now();
    if (per_iter_seconds && it < max_records) per_iter_seconds[it] = std::chrono::duration<double>(t1 - t0).count();
    it++;
    if (all_min && change < 1e-5) break;
    if (it >= p->multiframe_max_iters) break;
  }
  for (int f = 0; f < n_frames; ++f)
    for (int i = 0; i < 12; ++i) poses[12 * f + i] = X[f][i];
  if (iters_out) *iters_out = it;
  omp_set_num_threads(old);
  return 0;
}

}  // namespace

extern "C" int orc_align_multiframe(const cvo_params* p, int32_t n_frames, const cvo_cloud_view* frames,
                                    double* poses, const uint8_t* hold_const, int32_t n_edges, const int32_t* edges,
                                    int32_t threads, int32_t* iters_out, double* per_iter_seconds, int32_t max_records) {
  return align_multiframe_impl(p, n_frames, frames, poses, hold_const, n_edges, edges, threads, iters_out,
                               per_iter_seconds, max_records, false);
}

extern "C" int orc_align_multiframe_cpu_states(const cvo_params* p, int32_t n_frames, const cvo_cloud_view* frames,
                                               double* poses, const uint8_t* hold_const, int32_t n_edges,
                                               const int32_t* edges, int32_t threads, int32_t* iters_out,
                                               double* per_iter_seconds, int32_t max_records) {
  return align_multiframe_impl(p, n_frames, frames, poses, hold_const, n_edges, edges, threads, iters_out,
                               per_iter_seconds, max_records, true);
}

extern "C" int orc_cpu_state_pairs(const cvo_params* p, const cvo_cloud_view* c1, const cvo_cloud_view* c2,
                                   const double X1[12], const double X2[12], float ell, int32_t cap, int32_t* rows,
                                   int32_t* cols, double* weights, int64_t* nnz) {
  if (!p || !c1 || !c2 || !X1 || !X2 || !nnz || cap < 0) return 1;
  std::vector<Pair> pairs;
  cpu_state_pairs(p, *c1, *c2, X1, X2, ell, 0, pairs, nnz);
  const int n = (int)std::min<int64_t>(*nnz, cap);
  for (int q = 0; q < n; ++q) {
    rows[q] = pairs[q].i1;
    cols[q] = pairs[q].i2;
    weights[q] = pairs[q].a;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// The reference's CPU two-view CVO, cvo::cvo::align (src/cvo/Cvo.cpp:885-1089),
// restated as the CPU baseline of the small configuration.  Differences from
// the GPU path it restates: a neighbour search per iteration over the moved
// target (the reference rebuilds a kd-tree, Cvo.cpp:368-369; here a grid gives
// the same radius sets), fixed lengthscale (no range-ell) and no row cap; the
// flow is not normalised (:692-693); float step coefficients with the exact
// imag() == 0 root test, min_step when no root and a 0.8 cap (:701-830); the
// pose update with the dynamic float Exp_SEK3 (LieGroup.cpp:214-241); the stop
// distance is the Frobenius norm of the matrix logarithm (:262-272); and the
// indicator is A.sum() with decrease and increase branches (:1287-1381).
// ---------------------------------------------------------------------------

namespace {

struct LegacyRow {
  std::vector<int> j;
  std::vector<float> a;
};

// se_kernel (Cvo.cpp:349-456) at lengthscale l over the moved target yp.
void legacy_se_kernel(const orc_legacy_params* lp, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                      const std::vector<float>& yp, float l, std::vector<LegacyRow>& rows) {
  const int N = src->n, M = tgt->n, F = src->feat_dim, KS = src->num_classes;
  const float s2 = lp->sigma * lp->sigma;
  const float d2_thres = (float)(-2.0 * l * l * std::log((double)(lp->sp_thres / s2)));
  const float d2_c_thres =
      (float)(-2.0 * lp->c_ell * lp->c_ell * std::log((double)(lp->sp_thres / lp->c_sigma / lp->c_sigma)));
  const double r = std::sqrt((double)std::max(d2_thres, 0.0f));
  Grid g;
  g.build(yp.data(), M, std::max(r, 1e-3));
  rows.assign(N, LegacyRow());
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < N; ++i) {
    const float* x = src->xyz + 3 * (size_t)i;
    std::vector<std::pair<int, float>> m;
    g.query(x, r, [&](int j) {
      const float* y = &yp[3 * (size_t)j];
      float d2 = 0;
      for (int d = 0; d < 3; ++d) d2 += (x[d] - y[d]) * (x[d] - y[d]);
      if (d2 < d2_thres) m.push_back({j, d2});
    });
    std::sort(m.begin(), m.end());  // the sparse row's column order
    LegacyRow& R = rows[i];
    for (auto& mm : m) {
      const int j = mm.first;
      const float d2 = mm.second;
      const float d2_color = F > 0 ? sqdist_n_f(src->feat + (size_t)i * F, tgt->feat + (size_t)j * F, F) : 0.0f;
      if (!(d2_color < d2_c_thres)) continue;
      const float k = (float)(s2 * std::exp(-(double)d2 / (2.0 * l * l)));
      const float ck = (float)(lp->c_sigma * lp->c_sigma * std::exp(-(double)d2_color / (2.0 * lp->c_ell * lp->c_ell)));
      float sk = 1.0f;
      if (lp->use_semantics && KS > 0) {
        const float d2s = sqdist_n_f(src->sem + (size_t)i * KS, tgt->sem + (size_t)j * KS, KS);
        sk = (float)(lp->s_sigma * lp->s_sigma * std::exp(-(double)d2s / (2.0 * lp->s_ell * lp->s_ell)));
      }
      const float a = ck * k * sk * 1.0f;
      if (a > lp->sp_thres) {
        R.j.push_back(j);
        R.a.push_back(a);
      }
    }
  }
}

inline void cross_f3(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

inline void matvec3(const float* A, const float* x, float* o) {
  for (int r = 0; r < 3; ++r) o[r] = A[r * 3 + 0] * x[0] + A[r * 3 + 1] * x[1] + A[r * 3 + 2] * x[2];
}

inline void matmul3(const float* A, const float* B, float* o) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) o[r * 3 + c] = A[r * 3 + 0] * B[0 * 3 + c] + A[r * 3 + 1] * B[1 * 3 + c] + A[r * 3 + 2] * B[2 * 3 + c];
}

// compute_flow (Cvo.cpp:557-698): per row (1/c) A_i [x_i x y_j] and
// (1/d) A_i [y_j - x_i] in float, summed over rows in double; not normalised.
void legacy_flow(const orc_legacy_params* lp, const cvo_cloud_view* src, const std::vector<float>& yp,
                 const std::vector<LegacyRow>& rows, float omega[3], float v[3]) {
  const int N = src->n;
  const float ic = 1.0f / lp->c, id = 1.0f / lp->d;
  std::vector<double> po((size_t)N * 3), pv((size_t)N * 3);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) {
    const float* x = src->xyz + 3 * (size_t)i;
    float o[3] = {0, 0, 0}, w[3] = {0, 0, 0};
    const LegacyRow& R = rows[i];
    for (size_t q = 0; q < R.j.size(); ++q) {
      const float* y = &yp[3 * (size_t)R.j[q]];
      const float s = ic * R.a[q], t = id * R.a[q];
      float cr[3];
      cross_f3(x, y, cr);
      for (int d = 0; d < 3; ++d) {
        o[d] += s * cr[d];
        w[d] += t * (y[d] - x[d]);
      }
    }
    for (int d = 0; d < 3; ++d) {
      po[3 * (size_t)i + d] = (double)o[d];
      pv[3 * (size_t)i + d] = (double)w[d];
    }
  }
  double so[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
  for (int i = 0; i < N; ++i)
    for (int d = 0; d < 3; ++d) {
      so[d] += po[3 * (size_t)i + d];
      sv[d] += pv[3 * (size_t)i + d];
    }
  for (int d = 0; d < 3; ++d) {
    omega[d] = (float)so[d];
    v[d] = (float)sv[d];
  }
}

// compute_step_size (Cvo.cpp:701-830).
float legacy_step(const orc_legacy_params* lp, const cvo_cloud_view* src, const std::vector<float>& yp, int M,
                  const std::vector<LegacyRow>& rows, const float omega[3], const float v[3], float ell,
                  double bcde_out[4]) {
  const int N = src->n;
  const float W[9] = {0, -omega[2], omega[1], omega[2], 0, -omega[0], -omega[1], omega[0], 0};
  float W2[9], W3[9], W4[9];
  matmul3(W, W, W2);
  matmul3(W2, W, W3);
  matmul3(W3, W, W4);
  float Wv[3], W2v[3], W3v[3];
  matvec3(W, v, Wv);
  matvec3(W2, v, W2v);
  matvec3(W3, v, W3v);
  std::vector<float> xiz((size_t)M * 3), xi2z((size_t)M * 3), xi3z((size_t)M * 3), xi4z((size_t)M * 3), nx2(M), xdx(M),
      ec(M);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < M; ++j) {
    const float* y = &yp[3 * (size_t)j];
    float a[3], b2[3], b3[3], b4[3];
    cross_f3(omega, y, a);
    float t2[3], t3[3], t4[3];
    matvec3(W2, y, t2);
    matvec3(W3, y, t3);
    matvec3(W4, y, t4);
    for (int d = 0; d < 3; ++d) {
      a[d] += v[d];
      b2[d] = t2[d] + Wv[d];
      b3[d] = t3[d] + W2v[d];
      b4[d] = t4[d] + W3v[d];
      xiz[3 * (size_t)j + d] = a[d];
      xi2z[3 * (size_t)j + d] = b2[d];
      xi3z[3 * (size_t)j + d] = b3[d];
      xi4z[3 * (size_t)j + d] = b4[d];
    }
    nx2[j] = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    xdx[j] = -(a[0] * b2[0] + a[1] * b2[1] + a[2] * b2[2]);
    ec[j] = (b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]) + 2.0f * (a[0] * b3[0] + a[1] * b3[1] + a[2] * b3[2]);
  }
  const float tc = (float)(1.0 / (2.0 * ell * ell));
  std::vector<double> pb((size_t)N * 4);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) {
    const float* x = src->xyz + 3 * (size_t)i;
    double Bi = 0, Ci = 0, Di = 0, Ei = 0;
    const LegacyRow& R = rows[i];
    for (size_t q = 0; q < R.j.size(); ++q) {
      const int j = R.j[q];
      const float* y = &yp[3 * (size_t)j];
      const float df[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
      auto dot = [&](const std::vector<float>& m) {
        return m[3 * (size_t)j] * df[0] + m[3 * (size_t)j + 1] * df[1] + m[3 * (size_t)j + 2] * df[2];
      };
      const float beta = -2.0f * tc * dot(xiz);
      const float gamma = -tc * (nx2[j] + 2.0f * dot(xi2z));
      const float delta = 2.0f * tc * (xdx[j] + (-dot(xi3z)));
      const float epsil = -tc * (ec[j] + 2.0f * dot(xi4z));
      const float A = R.a[q];
      Bi += (double)(A * beta);
      Ci += (double)A * ((double)gamma + (double)(beta * beta) / 2.0);
      Di += (double)A * ((double)(delta + beta * gamma) + (double)(beta * beta * beta) / 6.0);
      Ei += (double)A * ((double)(epsil + beta * delta) + 0.5 * (double)beta * (double)beta * (double)gamma +
                         0.5 * (double)gamma * (double)gamma +
                         (1.0 / 24.0) * (double)beta * (double)beta * (double)beta * (double)beta);
    }
    pb[4 * (size_t)i] = Bi;
    pb[4 * (size_t)i + 1] = Ci;
    pb[4 * (size_t)i + 2] = Di;
    pb[4 * (size_t)i + 3] = Ei;
  }
  double B = 0, C = 0, D = 0, E = 0;
  for (int i = 0; i < N; ++i) {
    B += pb[4 * (size_t)i];
    C += pb[4 * (size_t)i + 1];
    D += pb[4 * (size_t)i + 2];
    E += pb[4 * (size_t)i + 3];
  }
  bcde_out[0] = B;
  bcde_out[1] = C;
  bcde_out[2] = D;
  bcde_out[3] = E;
  // roots of the float cubic 4E s^3 + 3D s^2 + 2C s + B; a root counts when
  // it is real (the reference's exact imag() == 0 on its companion eigenvalues)
  const double coef[4] = {(double)(4.0f * (float)E), (double)(3.0f * (float)D), (double)(2.0f * (float)C),
                          (double)(float)B};
  double re[3], im[3];
  cubic_roots(coef, re, im);
  float temp = std::numeric_limits<float>::max();
  for (int q = 0; q < 3; ++q) {
    const float rr = (float)re[q];
    if (std::isfinite(re[q]) && rr > 0 && rr < temp && im[q] == 0.0) temp = rr;
  }
  float step = temp == std::numeric_limits<float>::max() ? lp->min_step : temp;
  return step > 0.8f ? 0.8f : step;
}

// compute_indicator (Cvo.cpp:1287-1381) with its flag semantics: a decrease
// request waits while ell <= ell_min, an increase request while ell >= ell_max.
struct LegacyIndicator {
  std::deque<float> qs, qe;
  float ss = 0, se = 0;
  bool decrease = false, increase = false, last_decrease = false;
  void update(float indicator, const orc_legacy_params* lp, float& ell) {
    const float W = lp->ell_reduced_2;
    if ((float)qs.size() < W) {
      qs.push_back(indicator);
      ss += indicator;
    } else if ((float)qe.size() < W) {
      qe.push_back(indicator);
      se += indicator;
    } else if (std::fabs(1 - se / ss) < lp->ell_reduced_3) {
      decrease = true;
      qs.clear();
      qe.clear();
      ss = se = 0;
    } else if (se / ss < 0.7f) {
      increase = true;
      qs.clear();
      qe.clear();
      ss = se = 0;
    } else {
      se -= qe.front();
      ss += qe.front();
      qs.push_back(qe.front());
      qe.pop_front();
      ss -= qs.front();
      qs.pop_front();
      qe.push_back(indicator);
      se += indicator;
    }
    if (decrease && ell > lp->ell_min) {
      ell = ell * lp->ell_reduced_1;
      last_decrease = true;
      decrease = false;
    } else {
      // `~decrease && last_decrease` (Cvo.cpp:1367): ~ of a bool is never 0
      if (last_decrease) last_decrease = false;
    }
    if (increase && ell < lp->ell_max) {
      ell = ell * (1 / lp->ell_reduced_1);
      increase = false;
    }
  }
};

}  // namespace

extern "C" int orc_cvo_legacy_align(const orc_legacy_params* lp, const cvo_cloud_view* src,
                                    const cvo_cloud_view* tgt, const float T_init[16], int32_t threads, float T_out[16],
                                    int32_t* iters_out, int32_t* ret_out, double* seconds, orc_legacy_record* trace,
                                    int32_t trace_cap) {
  if (!lp || !src || !tgt || !T_init || !T_out) return 1;
  for (int i = 0; i < 16; ++i) T_out[i] = (i % 5 == 0) ? 1.0f : 0.0f;
  if (iters_out) *iters_out = 0;
  if (ret_out) *ret_out = 0;
  if (src->n == 0 || tgt->n == 0) return 0;  // set_pcd returns without a problem
  if (src->feat_dim != tgt->feat_dim) return 1;
  const int old = omp_get_max_threads();
  if (threads > 0) omp_set_num_threads(threads);
  const auto t0 = std::chrono::steady_clock::now();
  const int M = tgt->n;
  // set_pcd (Cvo.cpp:1182-1242): R, T from the init guess, ell = ell_init
  float R[9], T[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) R[r * 3 + c] = T_init[r * 4 + c];
    T[r] = T_init[r * 4 + 3];
  }
  float ell = lp->ell_init;
  LegacyIndicator ind;
  int iter = lp->MAX_ITER, ret = 0;
  std::vector<float> yp((size_t)M * 3);
  std::vector<LegacyRow> rows;
  auto applied = [&](float A[9], float b[3]) {  // update_tf: [R^T, -R^T T]
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) A[r * 3 + c] = R[c * 3 + r];
      b[r] = -(R[0 * 3 + r] * T[0] + R[1 * 3 + r] * T[1] + R[2 * 3 + r] * T[2]);
    }
  };
  for (int k = 0; k < lp->MAX_ITER; ++k) {
    float A[9], b[3];
    applied(A, b);
    for (int j = 0; j < M; ++j) {  // transform_pcd
      const float* y = tgt->xyz + 3 * (size_t)j;
      for (int r = 0; r < 3; ++r) yp[3 * (size_t)j + r] = A[r * 3 + 0] * y[0] + A[r * 3 + 1] * y[1] + A[r * 3 + 2] * y[2] + b[r];
    }
    legacy_se_kernel(lp, src, tgt, yp, ell, rows);
    float omega[3], v[3];
    legacy_flow(lp, src, yp, rows, omega, v);
    double bcde[4];
    const float step = legacy_step(lp, src, yp, M, rows, omega, v, ell, bcde);
    int64_t nnz = 0;
    float asum = 0.0f;
    for (const auto& row : rows) {
      nnz += (int64_t)row.j.size();
      for (float a : row.a) asum += a;
    }
    const double on = std::sqrt((double)omega[0] * omega[0] + (double)omega[1] * omega[1] + (double)omega[2] * omega[2]);
    const double vn = std::sqrt((double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2]);
    orc_legacy_record rec{};
    rec.ell = ell;
    rec.step = step;
    rec.nnz = (int32_t)nnz;
    rec.indicator = asum;
    for (int d = 0; d < 3; ++d) {
      rec.omega[d] = omega[d];
      rec.v[d] = v[d];
    }
    for (int c = 0; c < 4; ++c) rec.bcde[c] = (float)bcde[c];
    rec.dist = 0.0f;
    if (on < lp->eps && vn < lp->eps) {
      iter = k;
      if (on < 1e-8 && vn < 1e-8) ret = -1;
      if (trace && k < trace_cap) trace[k] = rec;
      break;
    }
    const float xi[6] = {omega[0], omega[1], omega[2], v[0], v[1], v[2]};
    float dt[12];
    exp_sek3_f32(xi, step, dt);
    float dR[9], dT[3], Rn[9];
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) dR[r * 3 + c] = dt[r * 4 + c];
      dT[r] = dt[r * 4 + 3];
    }
    float Tn[3];
    matvec3(R, dT, Tn);
    for (int r = 0; r < 3; ++r) T[r] = Tn[r] + T[r];
    matmul3(R, dR, Rn);
    std::memcpy(R, Rn, sizeof(R));
    // dist_se3: Frobenius norm of the 4x4 matrix logarithm, sqrt(2|w|^2 + |rho|^2)
    double d12[12], lg[6];
    for (int i = 0; i < 12; ++i) d12[i] = dt[i];
    log_se3(d12, lg);
    const float dist = (float)std::sqrt(2.0 * (lg[3] * lg[3] + lg[4] * lg[4] + lg[5] * lg[5]) +
                                        lg[0] * lg[0] + lg[1] * lg[1] + lg[2] * lg[2]);
    rec.dist = dist;
    if (trace && k < trace_cap) trace[k] = rec;
    if (dist < lp->eps_2) {
      iter = k;
      break;
    }
    ind.update(asum, lp, ell);
  }
  float A[9], b[3];
  applied(A, b);
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) T_out[r * 4 + c] = A[r * 3 + c];
    T_out[r * 4 + 3] = b[r];
  }
  if (iters_out) *iters_out = iter;
  if (ret_out) *ret_out = ret;
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  omp_set_num_threads(old);
  return 0;
}

extern "C" int orc_legacy_indicator_sequence(const orc_legacy_params* lp, int32_t n, const float* ind, float* ell_out) {
  if (!lp || n < 0 || !ind || !ell_out) return 1;
  LegacyIndicator I;
  float ell = lp->ell_init;
  for (int k = 0; k < n; ++k) {
    I.update(ind[k], lp, ell);
    ell_out[k] = ell;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// fill_in_A_mat_gpu_dense_mat_kernel (src/cvo/CvoGPU.cu:719-830) through
// inner_product_non_isotropic_impl (:1873-1952): per source point the metric
// U K^-1 U^T (rotate_kernel_and_inverse :134-161), k = sigma^2 exp(-d2 / 2), no
// distance gate, geometric type off, first K passing targets by index.
// ---------------------------------------------------------------------------

extern "C" int orc_assoc_nonisotropic(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                      const float T[16], const float kernel[9], int32_t K, int32_t* counts,
                                      int32_t* cols, float* vals) {
  if (!p || !src || !tgt || !T || !kernel || !counts || K <= 0) return 1;
  // float inverse of K by cofactors (computed in double, rounded)
  double k[9];
  for (int i = 0; i < 9; ++i) k[i] = kernel[i];
  const double det = k[0] * (k[4] * k[8] - k[5] * k[7]) + k[1] * (k[5] * k[6] - k[3] * k[8]) + k[2] * (k[3] * k[7] - k[4] * k[6]);
  if (!(det != 0.0) || !std::isfinite(det)) return 2;
  const float Ki[9] = {(float)((k[4] * k[8] - k[5] * k[7]) / det), (float)((k[2] * k[7] - k[1] * k[8]) / det),
                       (float)((k[1] * k[5] - k[2] * k[4]) / det), (float)((k[5] * k[6] - k[3] * k[8]) / det),
                       (float)((k[0] * k[8] - k[2] * k[6]) / det), (float)((k[2] * k[3] - k[0] * k[5]) / det),
                       (float)((k[3] * k[7] - k[4] * k[6]) / det), (float)((k[1] * k[6] - k[0] * k[7]) / det),
                       (float)((k[0] * k[4] - k[1] * k[3]) / det)};
  float A[9], b[3];
  applied_transform(pose_from16(T), A, b);
  const int N = src->n, M = tgt->n, F = src->feat_dim, KS = src->num_classes;
  const bool geo = p->is_using_geometry, col = p->is_using_intensity && F > 0, sem = p->is_using_semantics && KS > 0;
  const float sigma2 = p->sigma * p->sigma, c_sigma2 = p->c_sigma * p->c_sigma, s_sigma2 = p->s_sigma * p->s_sigma;
  const float c_ell2 = p->c_ell * p->c_ell, s_ell2 = p->s_ell * p->s_ell;
  const float d2_c_thres = col ? (float)(-2.0 * c_ell2 * std::log((double)(p->sp_thres / c_sigma2))) : 1.0f;
  const float d2_s_thres = sem ? (float)(-2.0 * s_ell2 * std::log((double)(p->sp_thres / s_sigma2))) : 1.0f;
  std::vector<float> yp((size_t)M * 3);
  for (int j = 0; j < M; ++j) xform_f(A, b, tgt->xyz + 3 * (size_t)j, &yp[3 * (size_t)j]);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < N; ++i) {
    const float* x = src->xyz + 3 * (size_t)i;
    float z[3] = {x[0], x[1], x[2]};
    const float n2 = std::fmaf(x[2], x[2], std::fmaf(x[1], x[1], x[0] * x[0]));
    if (n2 > 0.0f) {
      const float nr = std::sqrt(n2);
      for (int d = 0; d < 3; ++d) z[d] = x[d] / nr;
    }
    const float u[3] = {1.0f, 1.0f, (-1.0f) / z[2] * (z[0] + z[1])};
    const float w[3] = {z[1] * u[2] - z[2] * u[1], z[2] * u[0] - z[0] * u[2], z[0] * u[1] - z[1] * u[0]};
    const float U[9] = {u[0], w[0], z[0], u[1], w[1], z[1], u[2], w[2], z[2]};
    float T1[9], Mt[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        T1[r * 3 + c] = std::fmaf(U[r * 3 + 2], Ki[6 + c], std::fmaf(U[r * 3 + 1], Ki[3 + c], U[r * 3] * Ki[c]));
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        Mt[r * 3 + c] = std::fmaf(T1[r * 3 + 2], U[c * 3 + 2], std::fmaf(T1[r * 3 + 1], U[c * 3 + 1], T1[r * 3] * U[c * 3]));
    int n = 0;
    for (int j = 0; j < M && n < K; ++j) {
      const float* y = &yp[3 * (size_t)j];
      float kk = 1.0f, ck = 1.0f, sk = 1.0f;
      if (geo) {
        const float d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
        const float t0 = std::fmaf(d[2], Mt[6], std::fmaf(d[1], Mt[3], d[0] * Mt[0]));
        const float t1 = std::fmaf(d[2], Mt[7], std::fmaf(d[1], Mt[4], d[0] * Mt[1]));
        const float t2 = std::fmaf(d[2], Mt[8], std::fmaf(d[1], Mt[5], d[0] * Mt[2]));
        const float d2 = std::fmaf(t2, d[2], std::fmaf(t1, d[1], t0 * d[0]));
        kk = (float)(sigma2 * std::exp(-(double)d2 / 2.0));
      }
      if (col) {
        const float d2c = sqdist_n_f(src->feat + (size_t)i * F, tgt->feat + (size_t)j * F, F);
        if (!(d2c < d2_c_thres)) continue;
        ck = (float)(c_sigma2 * std::exp(-(double)d2c / (2.0 * c_ell2)));
      }
      if (sem) {
        const float d2s = sqdist_n_f(src->sem + (size_t)i * KS, tgt->sem + (size_t)j * KS, KS);
        if (!(d2s < d2_s_thres)) continue;
        sk = (float)(s_sigma2 * std::exp(-(double)d2s / (2.0 * s_ell2)));
      }
      const float a = ck * kk * sk * 1.0f;
      if (a > p->sp_thres) {
        cols[(size_t)i * K + n] = j;
        vals[(size_t)i * K + n] = a;
        ++n;
      }
    }
    counts[i] = n;
  }
  return 0;
}

#ifndef ORC_CXXFLAGS
#define ORC_CXXFLAGS "(flags not recorded)"
#endif
extern "C" const char* orc_build_info(void) { return "g++ " __VERSION__ ", " ORC_CXXFLAGS; }
