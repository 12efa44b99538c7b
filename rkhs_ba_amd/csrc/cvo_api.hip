// C ABI of the gfx950 CVO library: contexts, device clouds, two-view align,
// inner product / cos angle / association export.
//
// Per align call the device work is: source spatial sort, target grid levels,
// then the iteration loop {assoc_kernel, overflow_kernel, step_kernel}, which is
// captured in a hipGraph of kChunk iterations and replayed until the device
// controller sets `done`.  The host never reads device state inside the loop
// except one 4-byte flag per chunk, double-buffered so the GPU stays busy.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cvo_internal.h"

namespace cvo {

void launch_step(const AssocProblem* probs, int n_probs, int max_src_rows, const KConsts& kc, hipStream_t s,
                 const AssocProblem* host_one, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_init_state(const AssocProblem* probs, int n_probs, const float* T16, int stride16, const KConsts& kc,
                       float ell0, int K0, hipStream_t s);
void launch_grid_count(const float* xyz, int n, const GridLevel& L, uint32_t* key, uint32_t* count, hipStream_t s);
void launch_iota(uint32_t* v, int n, hipStream_t s);
// zero fill of bytes (a multiple of 4) by a kernel: no memset node in a
// caller's graph capture
void launch_zero(void* p, size_t bytes, hipStream_t s);
void launch_math_check(int n_cubic, const double* coef, double* re, double* im, int n_exp, const double* xi,
                       const double* dt, double* T64, double* log_norm, float* T32, hipStream_t s);
void launch_gather(const uint32_t* order, int n, const float* xyz, const float* feat, int F, const float* sem, int K,
                   const float* geo, float4* pos, float* feat_s, float* sem_s, float* geo_s, uint32_t* inv_perm,
                   hipStream_t s);
void launch_morton(const float* xyz, int n, const float* bmin, const float* bmax, uint32_t* key, hipStream_t s);
void launch_export(const AssocProblem* P_dev, int n_src, int64_t* counts_orig, const int64_t* row_ptr, int32_t* col,
                   float* val, int phase, int sort_threshold, hipStream_t s);
void launch_all_done(const AssocProblem* probs, int n_probs, int32_t* out, hipStream_t s);
void launch_rescore(const AssocProblem* probs, int n_probs, const KConsts& kc, float ell, int K, hipStream_t s);
void launch_adaptive_sync(const AssocProblem* probs, const KConsts& kc, hipStream_t s);

cvo_status set_err(cvo_ctx* ctx, cvo_status s, const std::string& msg) {
  if (ctx) ctx->last_error = msg;
  return s;
}

cvo_status hip_check(cvo_ctx* ctx, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory) return set_err(ctx, CVO_ERR_OOM, m);
  return set_err(ctx, CVO_ERR_HIP, m);
}

// Grows one arena buffer (25% headroom); the context's total stays within its
// workspace limit, if one is set.
cvo_status ensure(cvo_ctx* ctx, DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (b.cap >= bytes) return CVO_OK;
  if (ctx->bump) {
    // carve from the call's fixed region; never allocate
    BumpArena& a = *ctx->bump;
    const size_t need = (bytes + 255) / 256 * 256;
    if (a.used + need > a.size)
      return set_err(ctx, CVO_ERR_LIMIT,
                     "scratch: " + std::to_string(a.used + need) + " bytes needed, " + std::to_string(a.size) +
                         " available");
    b.p = a.base + a.used;
    b.cap = need;
    a.used += need;
    return CVO_OK;
  }
  const size_t cap = bytes + bytes / 4;
  if (ctx->ws_limit > 0 && ctx->ws_bytes - b.cap + cap > ctx->ws_limit)
    return set_err(ctx, CVO_ERR_LIMIT,
                   "workspace limit: " + std::to_string(ctx->ws_bytes - b.cap + cap) + " bytes needed, limit " +
                       std::to_string(ctx->ws_limit));
  if (b.p) {
    CVO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CVO_HIP(ctx, hipFree(b.p));
    ctx->ws_bytes -= b.cap;
    b.p = nullptr;
    b.cap = 0;
  }
  CVO_HIP(ctx, hipMalloc(&b.p, cap));
  b.cap = cap;
  ctx->ws_bytes += cap;
  return CVO_OK;
}

template <class T>
static T* as(DevBuf& b) {
  return reinterpret_cast<T*>(b.p);
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ---------------------------------------------------------------------------
// Kernel constants
// ---------------------------------------------------------------------------

KConsts make_kconsts(const cvo_params& p, int F, int Ks, int mode) {
  KConsts k;
  std::memset(&k, 0, sizeof(k));
  k.sigma2 = p.sigma * p.sigma;
  k.c_sigma2 = p.c_sigma * p.c_sigma;
  k.s_sigma2 = p.s_sigma * p.s_sigma;
  k.c2 = p.c_ell * p.c_ell;
  k.s_ell = p.s_ell;
  k.lg = logf(p.sp_thres / k.sigma2);
  float lc = logf(p.sp_thres / k.c_sigma2);
  float ls = logf(p.sp_thres / k.s_sigma2);
  k.use_geo = p.is_using_geometry;
  k.use_col = p.is_using_intensity;
  k.use_sem = p.is_using_semantics;
  k.use_gtype = p.is_using_geometric_type;
  k.d2_c_thres = k.use_col ? (float)(-2.0 * (double)k.c2 * (double)lc) : 1.0f;
  k.d2_s_thres = k.use_sem ? (float)(-2.0 * (double)k.s_ell * (double)k.s_ell * (double)ls) : 1.0f;
  k.q_c = (float)(1.0 / (2.0 * (double)k.c2));
  k.q_s = (float)(1.0 / (2.0 * (double)k.s_ell * (double)k.s_ell));
  k.sp = p.sp_thres;
  float amp = 1.0f;
  if (k.use_geo) amp *= k.sigma2;
  if (k.use_col) amp *= k.c_sigma2;
  if (k.use_sem) amp *= k.s_sigma2;
  k.amp = amp;
  k.F = F;
  k.Ks = Ks;
  k.c = p.c;
  k.d = p.d;
  k.eps = p.eps;
  k.eps_2 = p.eps_2;
  k.min_step = p.min_step;
  k.max_step = p.max_step;
  k.ell_min = p.ell_min;
  k.ell_decay_rate = p.ell_decay_rate;
  k.ell_decay_start = p.ell_decay_start;
  k.max_iter = p.MAX_ITER;
  k.nn_max = p.nearest_neighbors_max;
  k.window = p.indicator_window_size;
  k.ind_thr = p.indicator_stable_threshold;
  k.range_ell_in_step = p.is_using_range_ell;
  k.rq_coef = k.use_geo ? sqrtf(-2.0f * k.lg) * 1.001f : 0.0f;
  k.mode = mode;
  // the K-nearest association serves align and the multiframe states; the
  // inner product, function angle and association export use the ball
  // association in either setting (CvoGPU.cu:1310 is align's only branch)
  k.knn = mode == MODE_IP ? 0 : p.is_using_kdtree;
  return k;
}

// ---------------------------------------------------------------------------
// Workspace of one or more association problems over one source/target pair
// ---------------------------------------------------------------------------


cvo_status validate_pair(cvo_ctx* ctx, const cvo_dcloud* s, const cvo_dcloud* t) {
  const cvo_params& p = ctx->params;
  if (!s || !t) return set_err(ctx, CVO_ERR_INVALID_ARG, "null cloud");
  if (s->n == 0 || t->n == 0) return set_err(ctx, CVO_ERR_EMPTY, "empty cloud");
  if (p.is_using_kdtree < 0 || p.is_using_kdtree > kKnnMaxK)
    return set_err(ctx, CVO_ERR_LIMIT, "is_using_kdtree (the K-nearest multiplier) must be in [0, 512]");
  if (p.is_using_kdtree && !p.is_using_geometry)
    return set_err(ctx, CVO_ERR_UNSUPPORTED, "is_using_kdtree needs is_using_geometry");
  if (p.is_using_intensity && (!s->feat || !t->feat || s->F != t->F || s->F <= 0))
    return set_err(ctx, CVO_ERR_INVALID_ARG, "is_using_intensity needs features of equal dimension on both clouds");
  if (p.is_using_semantics && (!s->sem || !t->sem || s->K != t->K || s->K <= 0))
    return set_err(ctx, CVO_ERR_INVALID_ARG, "is_using_semantics needs class vectors of equal dimension on both clouds");
  if (p.is_using_geometric_type && (!s->geo || !t->geo))
    return set_err(ctx, CVO_ERR_INVALID_ARG, "is_using_geometric_type needs geo_type on both clouds");
  if (p.nearest_neighbors_max < 1 || p.nearest_neighbors_max > 1024)
    return set_err(ctx, CVO_ERR_LIMIT, "nearest_neighbors_max must be in [1, 1024]");
  if (p.indicator_window_size < 1 || p.indicator_window_size > kMaxWindow)
    return set_err(ctx, CVO_ERR_LIMIT, "indicator_window_size must be in [1, 64]");
  return CVO_OK;
}

// Radius needed at lengthscale ell (host mirror of needed_radius).
float radius_for(const KConsts& kc, float ell, float max_norm, float tq) {
  return kc.rq_coef * ell * (1.0f + max_norm / 500.0f) * 1.0001f + 1e-4f * (1.0f + max_norm + tq) + 1e-6f;
}

// Spatially sorted source rows (Morton order over the source bounding box).
cvo_status prepare_source_into(cvo_ctx* ctx, const cvo_dcloud* s, DevBuf& b_pos, DevBuf& b_feat, DevBuf& b_sem,
                               DevBuf& b_geo) {
  hipStream_t st = ctx->stream;
  const int n = s->n;
  CVO_TRY(ensure(ctx, ctx->b_src_key, sizeof(uint32_t) * n * 2));
  CVO_TRY(ensure(ctx, ctx->b_src_perm, sizeof(uint32_t) * n * 2));
  uint32_t* keys = as<uint32_t>(ctx->b_src_key);
  uint32_t* keys2 = keys + n;
  uint32_t* idx = as<uint32_t>(ctx->b_src_perm);
  uint32_t* idx2 = idx + n;
  launch_morton(s->xyz, n, s->bmin, s->bmax, keys, st);
  launch_iota(idx, n, st);
  size_t tmp = 0;
  CVO_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys, keys2, idx, idx2, n, 0, 30, st));
  CVO_TRY(ensure(ctx, ctx->b_cub, tmp));
  CVO_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(ctx->b_cub.p, tmp, keys, keys2, idx, idx2, n, 0, 30, st));
  CVO_TRY(ensure(ctx, b_pos, sizeof(float4) * n));
  float* fs = nullptr;
  float* ss = nullptr;
  float* gs = nullptr;
  if (s->feat) {
    CVO_TRY(ensure(ctx, b_feat, sizeof(float) * (size_t)n * s->F));
    fs = as<float>(b_feat);
  }
  if (s->sem) {
    CVO_TRY(ensure(ctx, b_sem, sizeof(float) * (size_t)n * s->K));
    ss = as<float>(b_sem);
  }
  if (s->geo) {
    CVO_TRY(ensure(ctx, b_geo, sizeof(float) * (size_t)n * 2));
    gs = as<float>(b_geo);
  }
  launch_gather(idx2, n, s->xyz, s->feat, s->F, s->sem, s->K, s->geo, as<float4>(b_pos), fs, ss, gs, nullptr,
                st);
  CVO_HIP(ctx, hipGetLastError());
  return CVO_OK;
}

cvo_status prepare_source(cvo_ctx* ctx, const cvo_dcloud* s) {
  return prepare_source_into(ctx, s, ctx->b_src_pos, ctx->b_src_feat, ctx->b_src_sem, ctx->b_src_geo);
}

// Grid levels over the target for cell sizes hs (descending).
cvo_status build_grid_into(cvo_ctx* ctx, const cvo_dcloud* t, const std::vector<float>& hs, GridHost& gh,
                           DevBuf& b_store, DevBuf& b_hdr) {
  hipStream_t st = ctx->stream;
  const int n = t->n;
  const int L = (int)hs.size();
  GridDev& G = gh.dev;
  std::memset(&G, 0, sizeof(G));
  G.n_levels = L;
  G.n = n;
  // layout
  std::vector<size_t> off_cells(L), off_pos(L), off_feat(L), off_sem(L), off_geo(L), off_inv(L);
  size_t total = 0;
  uint32_t max_cells = 0;
  for (int l = 0; l < L; ++l) {
    GridLevel& lv = G.lv[l];
    float h = hs[l];
    float pad = 1e-3f * h;
    double ext[3];
    for (int d = 0; d < 3; ++d) ext[d] = (double)t->bmax[d] - (double)t->bmin[d] + 2.0 * pad;
    double cells = 1.0;
    for (int d = 0; d < 3; ++d) cells *= std::floor(ext[d] / h) + 1.0;
    const double cell_cap = (double)ctx->grid_max_cells;
    if (cells > cell_cap) h = (float)(h * std::cbrt(cells / cell_cap) * 1.01);
    lv.h = h;
    lv.inv_h = 1.0f / h;
    lv.ox = t->bmin[0] - pad;
    lv.oy = t->bmin[1] - pad;
    lv.oz = t->bmin[2] - pad;
    lv.nx = (int)std::floor(ext[0] / h) + 1;
    lv.ny = (int)std::floor(ext[1] / h) + 1;
    lv.nz = (int)std::floor(ext[2] / h) + 1;
    uint32_t nc = (uint32_t)lv.nx * (uint32_t)lv.ny * (uint32_t)lv.nz;
    max_cells = std::max(max_cells, nc);
    off_cells[l] = total; total = align_up(total + sizeof(uint32_t) * (nc + 1), 256);
    off_pos[l] = total; total = align_up(total + sizeof(float4) * n, 256);
    off_feat[l] = total; total = align_up(total + (t->feat ? sizeof(float) * (size_t)n * t->F : 0), 256);
    off_sem[l] = total; total = align_up(total + (t->sem ? sizeof(float) * (size_t)n * t->K : 0), 256);
    off_geo[l] = total; total = align_up(total + (t->geo ? sizeof(float) * (size_t)n * 2 : 0), 256);
    off_inv[l] = total; total = align_up(total + sizeof(uint32_t) * n, 256);
  }
  CVO_TRY(ensure(ctx, b_store, total));
  char* base = as<char>(b_store);
  // temporaries: keys, keys2, idx, idx2, counts
  CVO_TRY(ensure(ctx, ctx->b_tmp0, sizeof(uint32_t) * (size_t)n * 4));
  CVO_TRY(ensure(ctx, ctx->b_tmp1, sizeof(uint32_t) * ((size_t)max_cells + 1)));
  uint32_t* keys = as<uint32_t>(ctx->b_tmp0);
  uint32_t* keys2 = keys + n;
  uint32_t* idx = keys2 + n;
  uint32_t* idx2 = idx + n;
  uint32_t* counts = as<uint32_t>(ctx->b_tmp1);
  for (int l = 0; l < L; ++l) {
    GridLevel& lv = G.lv[l];
    uint32_t nc = (uint32_t)lv.nx * (uint32_t)lv.ny * (uint32_t)lv.nz;
    uint32_t* cell_start = reinterpret_cast<uint32_t*>(base + off_cells[l]);
    launch_zero(counts, sizeof(uint32_t) * (nc + 1), st);
    launch_grid_count(t->xyz, n, lv, keys, counts, st);
    size_t tmp = 0;
    CVO_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, counts, cell_start, (int)nc + 1, st));
    size_t tmp2 = 0;
    int bits = 1;
    while ((1u << bits) < nc && bits < 32) ++bits;
    CVO_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp2, keys, keys2, idx, idx2, n, 0, bits, st));
    CVO_TRY(ensure(ctx, ctx->b_cub, std::max(tmp, tmp2)));
    CVO_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->b_cub.p, tmp, counts, cell_start, (int)nc + 1, st));
    launch_iota(idx, n, st);
    CVO_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(ctx->b_cub.p, tmp2, keys, keys2, idx, idx2, n, 0, bits, st));
    lv.cell_start = cell_start;
    lv.pos = reinterpret_cast<const float4*>(base + off_pos[l]);
    lv.feat = t->feat ? reinterpret_cast<const float*>(base + off_feat[l]) : nullptr;
    lv.sem = t->sem ? reinterpret_cast<const float*>(base + off_sem[l]) : nullptr;
    lv.geo = t->geo ? reinterpret_cast<const float*>(base + off_geo[l]) : nullptr;
    lv.inv_perm = reinterpret_cast<const uint32_t*>(base + off_inv[l]);
    launch_gather(idx2, n, t->xyz, t->feat, t->F, t->sem, t->K, t->geo, const_cast<float4*>(lv.pos),
                  const_cast<float*>(lv.feat), const_cast<float*>(lv.sem), const_cast<float*>(lv.geo),
                  const_cast<uint32_t*>(lv.inv_perm), st);
  }
  CVO_HIP(ctx, hipGetLastError());
  CVO_TRY(ensure(ctx, b_hdr, sizeof(GridDev)));
  gh.dev_ptr = as<GridDev>(b_hdr);
  CVO_HIP(ctx, hipMemcpyAsync(gh.dev_ptr, &G, sizeof(GridDev), hipMemcpyHostToDevice, st));
  return CVO_OK;
}

cvo_status build_grid(cvo_ctx* ctx, const cvo_dcloud* t, const std::vector<float>& hs, GridHost& gh) {
  return build_grid_into(ctx, t, hs, gh, ctx->b_grid_store, ctx->b_gridhdr);
}

// Cell sizes for lengthscales from ell_hi down to ell_lo, ratio 0.7, the finest
// reaching cell_scale x the radius at ell_lo.
std::vector<float> level_sizes(const KConsts& kc, float ell_hi, float ell_lo, float max_norm, float tq,
                               float cell_scale) {
  std::vector<float> hs;
  // without the geometric gate every row takes the full scan; one coarse level
  if (!kc.use_geo) return {1.0e4f};
  float r_hi = radius_for(kc, ell_hi, max_norm, tq);
  float r_lo = radius_for(kc, ell_lo, max_norm, tq) * cell_scale;
  if (!(r_hi > 0)) r_hi = 1.0f;
  float h = r_hi * 1.02f;
  for (int l = 0; l < kMaxLevels; ++l) {
    hs.push_back(h);
    if (h * 0.7f < r_lo) break;
    h *= 0.7f;
  }
  return hs;
}

// An association launch shape (tuning) as kernel constants.  batched: the
// launches carry several problems (their kernel has no four-lane path).
void apply_tuning(const cvo_tuning& t, KConsts& kc, bool batched) {
  const bool split = t.split_ell > 0.0f;
  kc.pol.split_ell = split ? t.split_ell : 0.0f;
  kc.pol.group_big = t.lanes_per_row;
  kc.pol.group_small = split ? t.lanes_small_ell : t.lanes_per_row;
  kc.pol.scale_big = t.cell_scale;
  kc.pol.scale_small = split ? t.cell_scale_small_ell : t.cell_scale;
  if (kc.knn > 0 || batched) {
    // the K-nearest pass keeps one 256-bin histogram per row, a batched
    // launch has no four-lane path: 8 or 16 lanes
    kc.pol.group_big = std::max(kc.pol.group_big, 8);
    kc.pol.group_small = std::max(kc.pol.group_small, 8);
  }
  kc.max_group = std::max(kc.pol.group_big, kc.pol.group_small);
  kc.min_scale = std::min(kc.pol.scale_big, kc.pol.scale_small);
  kc.xcd_remap = t.xcd_remap;
}

void apply_tuning(const cvo_ctx* ctx, KConsts& kc, bool batched) { apply_tuning(ctx->tuning, kc, batched); }

// Buffers and descriptors of n_probs problems sharing one source and target;
// each problem owns the association blocks of its rows at max_group lanes per
// row (a level with fewer lanes per row uses a prefix of them).
cvo_status setup_problems(cvo_ctx* ctx, const cvo_dcloud* s, const cvo_dcloud* t, const GridHost& gh, int n_probs,
                          int cap_alloc, int trace_cap, int max_group, Problems& pr) {
  const int n = s->n;
  const int rb = rows_per_block(max_group);
  const int nb = (n + rb - 1) / rb;
  pr.n_probs = n_probs;
  pr.total_blocks = nb * n_probs;
  pr.max_rows = n;
  // cap_alloc 0: the passes only sum (no pair array)
  if (cap_alloc > 0) CVO_TRY(ensure(ctx, ctx->b_pairs, sizeof(uint2) * (size_t)n * cap_alloc * n_probs));
  CVO_TRY(ensure(ctx, ctx->b_counts, sizeof(uint32_t) * (size_t)n * n_probs));
  CVO_TRY(ensure(ctx, ctx->b_ovf, sizeof(uint32_t) * (size_t)n * n_probs));
  CVO_TRY(ensure(ctx, ctx->b_state, sizeof(ProblemState) * n_probs));
  CVO_TRY(ensure(ctx, ctx->b_probs, sizeof(AssocProblem) * n_probs));
  CVO_TRY(ensure(ctx, ctx->b_part_a, sizeof(double) * 8 * (size_t)nb * n_probs));
  CVO_TRY(ensure(ctx, ctx->b_part_ovf, sizeof(double) * 10 * (size_t)kOvfBlocks * n_probs));
  CVO_TRY(ensure(ctx, ctx->b_part_s, sizeof(double) * 4 * (size_t)nb * n_probs));
  CVO_TRY(ensure(ctx, ctx->b_part_cnt, sizeof(uint32_t) * 2 * (size_t)nb * n_probs));
  pr.trace_cap = trace_cap;
  if (trace_cap > 0) {
    CVO_TRY(ensure(ctx, ctx->b_trace, sizeof(cvo_iter_record) * (size_t)trace_cap * n_probs));
    pr.trace = as<cvo_iter_record>(ctx->b_trace);
  } else {
    pr.trace = nullptr;
  }
  // ovf stamps start at 0 (iteration stamps are iter+1 >= 1)
  launch_zero(ctx->b_ovf.p, sizeof(uint32_t) * (size_t)n * n_probs, ctx->stream);
  launch_zero(ctx->b_state.p, sizeof(ProblemState) * n_probs, ctx->stream);
  pr.states = as<ProblemState>(ctx->b_state);
  pr.host.resize(n_probs);
  for (int q = 0; q < n_probs; ++q) {
    AssocProblem& P = pr.host[q];
    std::memset(&P, 0, sizeof(P));
    P.src_pos = as<float4>(ctx->b_src_pos);
    P.src_feat = s->feat ? as<float>(ctx->b_src_feat) : nullptr;
    P.src_sem = s->sem ? as<float>(ctx->b_src_sem) : nullptr;
    P.src_geo = s->geo ? as<float>(ctx->b_src_geo) : nullptr;
    P.n_src = n;
    P.n_tgt = t->n;
    P.grid = gh.dev_ptr;
    P.tgt_xyz = t->xyz;
    P.tgt_feat = t->feat;
    P.tgt_sem = t->sem;
    P.tgt_geo = t->geo;
    P.st = pr.states + q;
    P.pairs = cap_alloc > 0 ? as<uint2>(ctx->b_pairs) + (size_t)q * n * cap_alloc : nullptr;
    P.counts = as<uint32_t>(ctx->b_counts) + (size_t)q * n;
    P.ovf_stamp = as<uint32_t>(ctx->b_ovf) + (size_t)q * n;
    P.cap_alloc = cap_alloc;
    P.block_begin = q * nb;
    P.n_blocks = nb;
    P.trace_cap = trace_cap;
    P.part_a = as<double>(ctx->b_part_a) + (size_t)q * nb * 8;
    P.part_ovf = as<double>(ctx->b_part_ovf) + (size_t)q * kOvfBlocks * 10;
    P.part_s = as<double>(ctx->b_part_s) + (size_t)q * nb * 4;
    P.part_cnt = as<uint32_t>(ctx->b_part_cnt) + (size_t)q * nb * 2;
    P.trace = pr.trace ? pr.trace + (size_t)q * trace_cap : nullptr;
  }
  pr.dev = as<AssocProblem>(ctx->b_probs);
  CVO_HIP(ctx, hipMemcpyAsync(pr.dev, pr.host.data(), sizeof(AssocProblem) * n_probs, hipMemcpyHostToDevice,
                              ctx->stream));
  return CVO_OK;
}

cvo_status pinned(cvo_ctx* ctx, size_t bytes) {
  if (ctx->pinned_cap >= bytes) return CVO_OK;
  if (ctx->pinned) CVO_HIP(ctx, hipHostFree(ctx->pinned));
  ctx->pinned = nullptr;
  CVO_HIP(ctx, hipHostMalloc(&ctx->pinned, bytes));
  ctx->pinned_cap = bytes;
  return CVO_OK;
}

cvo_status ensure_events(cvo_ctx* ctx, size_t n) {
  while (ctx->ev_pool.size() < n) {
    hipEvent_t e;
    CVO_HIP(ctx, hipEventCreate(&e));
    ctx->ev_pool.push_back(e);
  }
  return CVO_OK;
}

// Iteration loop: replays chunks of kChunk iterations until every problem is
// done.  Profiling mode launches kernel by kernel with events around each.
static constexpr int kChunk = 24;

static cvo_status run_loop(cvo_ctx* ctx, const Problems& pr, const KConsts& kc, int max_iter, bool with_step,
                           double* device_seconds) {
  hipStream_t st = ctx->stream;
  CVO_TRY(pinned(ctx, 64));
  volatile int32_t* flags = reinterpret_cast<volatile int32_t*>(ctx->pinned);
  // The done flags of all problems are gathered with one strided copy of the
  // states' `done` fields.
  const size_t pitch = sizeof(ProblemState);
  CVO_TRY(ensure(ctx, ctx->b_tmp3, sizeof(int32_t) * pr.n_probs + 64));
  auto read_all_done = [&](bool* all) -> cvo_status {
    std::vector<int32_t> fl(pr.n_probs);
    CVO_HIP(ctx, hipMemcpy2DAsync((void*)(ctx->b_tmp3.p), sizeof(int32_t), &pr.states[0].done, pitch, sizeof(int32_t),
                                  pr.n_probs, hipMemcpyDeviceToDevice, st));
    CVO_HIP(ctx, hipMemcpyAsync(fl.data(), ctx->b_tmp3.p, sizeof(int32_t) * pr.n_probs, hipMemcpyDeviceToHost, st));
    CVO_HIP(ctx, hipStreamSynchronize(st));
    *all = true;
    for (int q = 0; q < pr.n_probs; ++q) *all = *all && fl[q];
    return CVO_OK;
  };
  // the step pass and controller; an adaptive-lengthscale run steps its
  // registration problem only, then its self-association problems follow it
  auto step_phase = [&](hipEvent_t e0, hipEvent_t e1) {
    launch_step(pr.dev, kc.adaptive ? 1 : pr.n_probs, pr.max_rows, kc, st, pr.host.data(), e0, e1);
    if (kc.adaptive) launch_adaptive_sync(pr.dev, kc, st);
  };
  CVO_TRY(ensure_events(ctx, 8));
  hipEvent_t t0 = ctx->ev_pool[0], t1 = ctx->ev_pool[1];
  hipEvent_t done_ev[2] = {ctx->ev_pool[2], ctx->ev_pool[3]};
  CVO_HIP(ctx, hipEventRecord(t0, st));

  if (ctx->profiling) {
    // kernel by kernel, each timed by events at its own start and end (the
    // kernel alone, as a kernel tracer times it)
    CVO_TRY(ensure_events(ctx, 8 + 6));
    hipEvent_t e[6];
    for (int i = 0; i < 6; ++i) e[i] = ctx->ev_pool[8 + i];
    for (int it = 0; it < max_iter; ++it) {
      launch_assoc(pr.dev, pr.n_probs, pr.total_blocks, kc, st, pr.host.data(), e[0], e[1]);
      launch_overflow(pr.dev, pr.n_probs, kc, st, pr.host.data(), e[2], e[3]);
      if (with_step) step_phase(e[4], e[5]);
      bool all = true;
      CVO_TRY(read_all_done(&all));
      float ms[3] = {0.0f, 0.0f, 0.0f};
      CVO_HIP(ctx, hipEventElapsedTime(&ms[0], e[0], e[1]));
      CVO_HIP(ctx, hipEventElapsedTime(&ms[1], e[2], e[3]));
      if (with_step) CVO_HIP(ctx, hipEventElapsedTime(&ms[2], e[4], e[5]));
      ctx->stats.ms[CVO_KSTAT_ASSOC] += ms[0];
      ctx->stats.ms[CVO_KSTAT_OVERFLOW] += ms[1];
      ctx->stats.ms[CVO_KSTAT_STEP] += with_step ? ms[2] : 0.0;
      ctx->stats.launches[CVO_KSTAT_ASSOC] += 1;
      ctx->stats.launches[CVO_KSTAT_OVERFLOW] += 1;
      if (with_step) ctx->stats.launches[CVO_KSTAT_STEP] += 1;
      if (!with_step || all) break;
    }
  } else if (!with_step) {
    launch_assoc(pr.dev, pr.n_probs, pr.total_blocks, kc, st, pr.host.data());
    launch_overflow(pr.dev, pr.n_probs, kc, st, pr.host.data());
    CVO_HIP(ctx, hipGetLastError());
  } else {
    // one chunk of iterations, captured in a graph unless CVO_NO_GRAPH is set
    // (direct launches, e.g. under a kernel tracer)
    auto launch_chunk = [&]() {
      for (int it = 0; it < kChunk; ++it) {
        launch_assoc(pr.dev, pr.n_probs, pr.total_blocks, kc, st, pr.host.data());
        launch_overflow(pr.dev, pr.n_probs, kc, st, pr.host.data());
        step_phase(nullptr, nullptr);
      }
    };
    const char* ng = std::getenv("CVO_NO_GRAPH");
    const bool use_graph = !(ng && ng[0] && ng[0] != '0');
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
      CVO_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      launch_chunk();
      hipError_t ce = hipStreamEndCapture(st, &graph);
      if (ce != hipSuccess) return hip_check(ctx, ce, "hipStreamEndCapture");
      hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (ie != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return hip_check(ctx, ie, "hipGraphInstantiate");
      }
    }
    const int n_chunks = (max_iter + kChunk - 1) / kChunk;
    bool finished = false;
    cvo_status status = CVO_OK;
    int32_t* dflag = as<int32_t>(ctx->b_tmp3);
    // Chunk c writes its all-done flag into pinned slot c % 2; the host waits
    // for chunk c-1 while chunk c runs.
    for (int c = 0; c < n_chunks && !finished; ++c) {
      hipError_t le = hipSuccess;
      if (use_graph) le = hipGraphLaunch(exec, st);
      else {
        launch_chunk();
        le = hipGetLastError();
      }
      if (le != hipSuccess) { status = hip_check(ctx, le, "chunk launch"); break; }
      launch_all_done(pr.dev, pr.n_probs, dflag + (c & 1), st);
      hipError_t me = hipMemcpyAsync((void*)(flags + (c & 1)), dflag + (c & 1), sizeof(int32_t),
                                     hipMemcpyDeviceToHost, st);
      if (me != hipSuccess) { status = hip_check(ctx, me, "flag copy"); break; }
      me = hipEventRecord(done_ev[c & 1], st);
      if (me != hipSuccess) { status = hip_check(ctx, me, "hipEventRecord"); break; }
      if (c >= 1) {
        me = hipEventSynchronize(done_ev[(c - 1) & 1]);
        if (me != hipSuccess) { status = hip_check(ctx, me, "hipEventSynchronize"); break; }
        if (flags[(c - 1) & 1]) finished = true;
      }
    }
    hipError_t se = hipStreamSynchronize(st);
    if (use_graph) {
      (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
    }
    if (status != CVO_OK) return status;
    if (se != hipSuccess) return hip_check(ctx, se, "hipStreamSynchronize");
  }
  CVO_HIP(ctx, hipEventRecord(t1, st));
  CVO_HIP(ctx, hipEventSynchronize(t1));
  float ms = 0;
  CVO_HIP(ctx, hipEventElapsedTime(&ms, t0, t1));
  if (device_seconds) *device_seconds = ms / 1000.0;
  CVO_HIP(ctx, hipGetLastError());
  return CVO_OK;
}

static void mat_inverse_rigid(const float T[16], float out[16]) {
  // out = [R^T, -R^T t]
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) out[r * 4 + c] = T[c * 4 + r];
    out[r * 4 + 3] = -(T[0 * 4 + r] * T[3] + T[1 * 4 + r] * T[7] + T[2 * 4 + r] * T[11]);
  }
  out[12] = out[13] = out[14] = 0.0f;
  out[15] = 1.0f;
}

static cvo_status export_assoc(cvo_ctx* ctx, const cvo_dcloud* s, const cvo_dcloud* t, const Problems& pr,
                               cvo_assoc* out) {
  return export_problem(ctx, pr.dev, s->n, t->n, out);
}

// The stored pairs of one problem (device descriptor P_dev, n_src rows) as a
// host CSR: rows in input order, target ids ascending (or by value, see
// export_rows).
cvo_status export_problem(cvo_ctx* ctx, const AssocProblem* P_dev, int n_src, int n_tgt, cvo_assoc* out) {
  hipStream_t st = ctx->stream;
  const int n = n_src;
  CVO_TRY(ensure(ctx, ctx->b_export, sizeof(int64_t) * (size_t)(n + 1) * 2));
  int64_t* cnt = as<int64_t>(ctx->b_export);
  int64_t* rp = cnt + (n + 1);
  CVO_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(int64_t) * (n + 1), st));
  launch_export(P_dev, n, cnt, nullptr, nullptr, nullptr, 0, 0, st);
  size_t tmp = 0;
  CVO_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt, rp, n + 1, st));
  CVO_TRY(ensure(ctx, ctx->b_cub, tmp));
  CVO_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->b_cub.p, tmp, cnt, rp, n + 1, st));
  int64_t nnz = 0;
  CVO_HIP(ctx, hipMemcpyAsync(&nnz, rp + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  CVO_HIP(ctx, hipStreamSynchronize(st));
  CVO_TRY(ensure(ctx, ctx->b_tmp2, (sizeof(int32_t) + sizeof(float)) * (size_t)std::max<int64_t>(nnz, 1)));
  int32_t* dcol = as<int32_t>(ctx->b_tmp2);
  float* dval = reinterpret_cast<float*>(dcol + std::max<int64_t>(nnz, 1));
  launch_export(P_dev, n, nullptr, rp, dcol, dval, 1, ctx->params.multiframe_is_sorting_inner_product, st);
  CVO_HIP(ctx, hipGetLastError());
  out->rows = n;
  out->cols = n_tgt;
  out->nnz = nnz;
  out->row_ptr = static_cast<int64_t*>(std::malloc(sizeof(int64_t) * (n + 1)));
  out->col = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
  out->val = static_cast<float*>(std::malloc(sizeof(float) * std::max<int64_t>(nnz, 1)));
  if (!out->row_ptr || !out->col || !out->val) return set_err(ctx, CVO_ERR_OOM, "host allocation of association");
  CVO_HIP(ctx, hipMemcpyAsync(out->row_ptr, rp, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, st));
  if (nnz > 0) {
    CVO_HIP(ctx, hipMemcpyAsync(out->col, dcol, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, st));
    CVO_HIP(ctx, hipMemcpyAsync(out->val, dval, sizeof(float) * nnz, hipMemcpyDeviceToHost, st));
  }
  CVO_HIP(ctx, hipStreamSynchronize(st));
  return CVO_OK;
}

}  // namespace cvo

using namespace cvo;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" const char* cvo_version(void) { return "rkhs_ba_amd 0.1 (gfx950)"; }

extern "C" cvo_status cvo_device_math_check(int32_t n_cubic, const double* coef, double* re, double* im,
                                            int32_t n_exp, const double* xi, const double* dt, double* T64,
                                            double* log_norm, float* T32) {
  if (n_cubic < 0 || n_exp < 0 || (n_cubic > 0 && (!coef || !re || !im)) ||
      (n_exp > 0 && (!xi || !dt || !T64 || !log_norm || !T32)))
    return CVO_ERR_INVALID_ARG;
  if (n_cubic == 0 && n_exp == 0) return CVO_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CVO_ERR_NO_DEVICE;
  const size_t nc = (size_t)n_cubic, ne = (size_t)n_exp;
  // one buffer: coef 4, re 3, im 3 per cubic; xi 6, dt 1, T64 12, log 1 (doubles), T32 12 (floats) per twist
  const size_t nd = 10 * nc + 20 * ne;
  const size_t bytes = nd * sizeof(double) + 12 * ne * sizeof(float);
  void* buf = nullptr;
  if (hipMalloc(&buf, bytes) != hipSuccess) return CVO_ERR_OOM;
  double* d = static_cast<double*>(buf);
  double *d_coef = d, *d_re = d + 4 * nc, *d_im = d + 7 * nc, *d_xi = d + 10 * nc, *d_dt = d_xi + 6 * ne;
  double *d_T64 = d_dt + ne, *d_log = d_T64 + 12 * ne;
  float* d_T32 = reinterpret_cast<float*>(d + nd);
  bool ok = true;
  ok = ok && (nc == 0 || hipMemcpy(d_coef, coef, 4 * nc * sizeof(double), hipMemcpyHostToDevice) == hipSuccess);
  ok = ok && (ne == 0 || hipMemcpy(d_xi, xi, 6 * ne * sizeof(double), hipMemcpyHostToDevice) == hipSuccess);
  ok = ok && (ne == 0 || hipMemcpy(d_dt, dt, ne * sizeof(double), hipMemcpyHostToDevice) == hipSuccess);
  if (ok) {
    cvo::launch_math_check(n_cubic, d_coef, d_re, d_im, n_exp, d_xi, d_dt, d_T64, d_log, d_T32, nullptr);
    ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess;
  }
  ok = ok && (nc == 0 || (hipMemcpy(re, d_re, 3 * nc * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
                          hipMemcpy(im, d_im, 3 * nc * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess));
  ok = ok && (ne == 0 || (hipMemcpy(T64, d_T64, 12 * ne * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
                          hipMemcpy(log_norm, d_log, ne * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
                          hipMemcpy(T32, d_T32, 12 * ne * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess));
  (void)hipFree(buf);
  return ok ? CVO_OK : CVO_ERR_HIP;
}

extern "C" const char* cvo_status_string(cvo_status s) {
  switch (s) {
    case CVO_OK: return "ok";
    case CVO_ERR_INVALID_ARG: return "invalid argument";
    case CVO_ERR_EMPTY: return "empty point cloud";
    case CVO_ERR_IO: return "file could not be opened";
    case CVO_ERR_PARSE: return "YAML value could not be parsed";
    case CVO_ERR_HIP: return "HIP runtime error";
    case CVO_ERR_NO_DEVICE: return "no HIP device";
    case CVO_ERR_OOM: return "out of device memory";
    case CVO_ERR_UNSUPPORTED: return "unsupported option";
    case CVO_ERR_LIMIT: return "size beyond a documented limit";
  }
  return "unknown status";
}

extern "C" const char* cvo_last_error(const cvo_ctx* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

// ---------------------------------------------------------------------------
// Scratch memory of the no-allocation entry point: a fixed device array per
// GPU.  In host code its name stands for a host-side shadow, not a device
// address; the device address comes from hipGetSymbolAddress, once per device
// with that device current, and only that address is handed to the runtime
// and the kernels.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = 64ull << 20;
constexpr int kNoAllocMaxPoints = 1 << 17;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[kMaxDevices] = {};

// The device address of this device's scratch array (ctx->device current).
static cvo_status scratch_address(cvo_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return set_err(ctx, CVO_ERR_LIMIT, "device ordinal above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[ctx->device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return hip_check(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[ctx->device] = static_cast<char*>(ptr);
  }
  ctx->scratch = g_scratch_ptr[ctx->device];
  return CVO_OK;
}

extern "C" cvo_status cvo_create_from_params(const cvo_params* p, cvo_ctx** out) {
  if (!p || !out) return CVO_ERR_INVALID_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CVO_ERR_NO_DEVICE;
  cvo_ctx* ctx = new cvo_ctx();
  ctx->params = *p;
  if (hipGetDevice(&ctx->device) != hipSuccess) {
    delete ctx;
    return CVO_ERR_HIP;
  }
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return CVO_ERR_HIP;
  }
  if (scratch_address(ctx) != CVO_OK) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CVO_ERR_HIP;
  }
  *out = ctx;
  return CVO_OK;
}

extern "C" cvo_status cvo_create(const char* yaml_path, cvo_ctx** out) {
  if (!out) return CVO_ERR_INVALID_ARG;
  cvo_params p;
  cvo_params_default(&p);
  if (yaml_path) {
    cvo_status s = cvo_params_read_yaml(yaml_path, &p);
    if (s != CVO_OK) return s;
  }
  return cvo_create_from_params(&p, out);
}

extern "C" cvo_status cvo_get_params(const cvo_ctx* ctx, cvo_params* out) {
  if (!ctx || !out) return CVO_ERR_INVALID_ARG;
  *out = ctx->params;
  return CVO_OK;
}

extern "C" cvo_status cvo_set_params(cvo_ctx* ctx, const cvo_params* p) {
  if (!ctx || !p) return CVO_ERR_INVALID_ARG;
  ctx->params = *p;
  return CVO_OK;
}

static void free_buf(cvo_ctx* ctx, DevBuf& b) {
  if (b.p) (void)hipFree(b.p);
  ctx->ws_bytes -= b.cap;
  b.p = nullptr;
  b.cap = 0;
}

// Every arena buffer of a context.
static std::vector<DevBuf*> all_bufs(cvo_ctx* ctx) {
  std::vector<DevBuf*> v = {&ctx->b_src_pos, &ctx->b_src_feat, &ctx->b_src_sem, &ctx->b_src_geo, &ctx->b_src_key,
                            &ctx->b_src_perm, &ctx->b_grid_store, &ctx->b_pairs, &ctx->b_counts, &ctx->b_ovf,
                            &ctx->b_state, &ctx->b_probs, &ctx->b_gridhdr, &ctx->b_part_a, &ctx->b_part_ovf,
                            &ctx->b_part_s, &ctx->b_part_cnt, &ctx->b_trace, &ctx->b_tmp0, &ctx->b_tmp1,
                            &ctx->b_tmp2, &ctx->b_tmp3, &ctx->b_cub, &ctx->b_export};
  for (auto& b : ctx->b_ba) v.push_back(&b);
  for (auto& b : ctx->b_frames) v.push_back(&b);
  for (auto& b : ctx->b_mlp) v.push_back(&b);
  for (auto& b : ctx->b_acvo) v.push_back(&b);
  return v;
}

static void for_each_buf(cvo_ctx* ctx, void (*fn)(cvo_ctx*, DevBuf&)) {
  for (DevBuf* b : all_bufs(ctx)) fn(ctx, *b);
}

extern "C" void cvo_destroy(cvo_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for_each_buf(ctx, free_buf);
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" cvo_status cvo_cloud_upload(cvo_ctx* ctx, const cvo_cloud_view* v, cvo_dcloud** out) {
  if (!ctx || !v || !out) return CVO_ERR_INVALID_ARG;
  *out = nullptr;
  if (v->n < 0 || (v->n > 0 && !v->xyz)) return set_err(ctx, CVO_ERR_INVALID_ARG, "bad cloud view");
  if ((v->feat && v->feat_dim <= 0) || (v->sem && v->num_classes <= 0) || v->feat_dim > 64 || v->num_classes > 64)
    return set_err(ctx, CVO_ERR_INVALID_ARG, "feature/class dimension must be in [1, 64]");
  cvo_dcloud* c = new cvo_dcloud();
  c->n = v->n;
  c->F = v->feat ? v->feat_dim : 0;
  c->K = v->sem ? v->num_classes : 0;
  c->has_geo = v->geo_type != nullptr;
  for (int d = 0; d < 3; ++d) {
    c->bmin[d] = 0.0f;
    c->bmax[d] = 0.0f;
  }
  if (c->n > 0) {
    for (int d = 0; d < 3; ++d) c->bmin[d] = c->bmax[d] = v->xyz[d];
    double mx = 0;
    for (int j = 0; j < c->n; ++j) {
      for (int d = 0; d < 3; ++d) {
        float x = v->xyz[3 * (size_t)j + d];
        if (!std::isfinite(x)) {
          delete c;
          return set_err(ctx, CVO_ERR_INVALID_ARG, "non-finite coordinate in cloud");
        }
        c->bmin[d] = std::min(c->bmin[d], x);
        c->bmax[d] = std::max(c->bmax[d], x);
      }
      const float* x = v->xyz + 3 * (size_t)j;
      mx = std::max(mx, std::sqrt((double)x[0] * x[0] + (double)x[1] * x[1] + (double)x[2] * x[2]));
    }
    c->max_norm = (float)mx;
    auto up = [&](float** dst, const float* src, size_t count) -> cvo_status {
      if (!src) return CVO_OK;
      CVO_HIP(ctx, hipMalloc(dst, sizeof(float) * count));
      CVO_HIP(ctx, hipMemcpyAsync(*dst, src, sizeof(float) * count, hipMemcpyHostToDevice, ctx->stream));
      return CVO_OK;
    };
    cvo_status s = up(&c->xyz, v->xyz, 3 * (size_t)c->n);
    if (s == CVO_OK) s = up(&c->feat, v->feat, (size_t)c->n * c->F);
    if (s == CVO_OK) s = up(&c->sem, v->sem, (size_t)c->n * c->K);
    if (s == CVO_OK) s = up(&c->geo, v->geo_type, (size_t)c->n * 2);
    if (s == CVO_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) s = set_err(ctx, CVO_ERR_HIP, "upload sync");
    if (s != CVO_OK) {
      cvo_cloud_free(ctx, c);
      return s;
    }
  }
  *out = c;
  return CVO_OK;
}

extern "C" void cvo_cloud_free(cvo_ctx* ctx, cvo_dcloud* c) {
  if (!c) return;
  if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (c->xyz) (void)hipFree(c->xyz);
  if (c->feat) (void)hipFree(c->feat);
  if (c->sem) (void)hipFree(c->sem);
  if (c->geo) (void)hipFree(c->geo);
  delete c;
}

extern "C" int32_t cvo_cloud_size(const cvo_dcloud* c) { return c ? c->n : 0; }

extern "C" cvo_status cvo_set_trace(cvo_ctx* ctx, int32_t cap) {
  if (!ctx || cap < 0) return CVO_ERR_INVALID_ARG;
  ctx->trace_cap = cap;
  ctx->trace_n = 0;
  return CVO_OK;
}

extern "C" cvo_status cvo_get_trace(cvo_ctx* ctx, cvo_iter_record* out, int32_t cap, int32_t* n_out) {
  if (!ctx || (!out && cap > 0)) return CVO_ERR_INVALID_ARG;
  int n = std::min<int>(cap, ctx->trace_n);
  for (int i = 0; i < n; ++i) out[i] = ctx->trace[i];
  if (n_out) *n_out = ctx->trace_n;
  return CVO_OK;
}

extern "C" cvo_status cvo_set_profiling(cvo_ctx* ctx, int32_t on) {
  if (!ctx) return CVO_ERR_INVALID_ARG;
  ctx->profiling = on != 0;
  return CVO_OK;
}

extern "C" cvo_status cvo_get_kernel_stats(cvo_ctx* ctx, cvo_kernel_stats* out) {
  if (!ctx || !out) return CVO_ERR_INVALID_ARG;
  *out = ctx->stats;
  return CVO_OK;
}

extern "C" cvo_status cvo_reset_kernel_stats(cvo_ctx* ctx) {
  if (!ctx) return CVO_ERR_INVALID_ARG;
  std::memset(&ctx->stats, 0, sizeof(ctx->stats));
  return CVO_OK;
}

// Algorithmic bytes of one association launch and one step launch.
static void account_bytes(cvo_ctx* ctx, const cvo_dcloud* s, const cvo_dcloud* t, const KConsts& kc,
                          const std::vector<cvo_iter_record>& tr, int n_it, bool with_step) {
  const double per_pt = 12.0 + (kc.use_col ? 4.0 * kc.F : 0.0) + (kc.use_sem ? 4.0 * kc.Ks : 0.0);
  for (int i = 0; i < n_it && i < (int)tr.size(); ++i) {
    const double S = tr[i].nnz;
    ctx->stats.bytes[CVO_KSTAT_ASSOC] += per_pt * (s->n + t->n) + 8.0 * S + 4.0 * s->n;
    if (with_step) ctx->stats.bytes[CVO_KSTAT_STEP] += 8.0 * S + 4.0 * s->n + 12.0 * (s->n + t->n);
    ctx->stats.pairs_stored += S;
  }
}

// The adaptive-lengthscale association covers three problems (x against
// the moved y, x against x, y against y): every point's data three times per
// launch; the stored pairs counted are the registration problem's (the self
// problems' counts are not traced), so the figure is a lower bound.
static void account_bytes_adaptive(cvo_ctx* ctx, const cvo_dcloud* x, const cvo_dcloud* y, const KConsts& kc,
                                   const std::vector<cvo_iter_record>& tr, int n_it) {
  const double per_pt = 12.0 + (kc.use_col ? 4.0 * kc.F : 0.0) + (kc.use_sem ? 4.0 * kc.Ks : 0.0);
  for (int i = 0; i < n_it && i < (int)tr.size(); ++i) {
    const double S = tr[i].nnz;
    ctx->stats.bytes[CVO_KSTAT_ASSOC] += per_pt * 3.0 * (x->n + y->n) + 8.0 * S + 4.0 * (2.0 * x->n + y->n);
    ctx->stats.bytes[CVO_KSTAT_STEP] += 8.0 * S + 4.0 * x->n + 12.0 * (x->n + y->n);
    ctx->stats.pairs_stored += S;
  }
}

// The registration loop of both align overloads (align_impl, CvoGPU.cu:
// 1239-1477) from the init transform as given.
static cvo_status align_impl_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, const float T_init[16],
                               float T_out[16], cvo_assoc* assoc_out, double* seconds, int32_t* iters,
                               int32_t* ret_code) {
  if (!ctx || !T_init || !T_out) return CVO_ERR_INVALID_ARG;
  if (ret_code) *ret_code = 0;
  if (iters) *iters = 0;
  if (seconds) *seconds = 0.0;
  if (!src || !tgt || src->n == 0 || tgt->n == 0) {
    // the int overload returns 0 on empty input and leaves transform untouched
    for (int i = 0; i < 16; ++i) T_out[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    return set_err(ctx, CVO_ERR_EMPTY, "empty point cloud");
  }
  CVO_TRY(validate_pair(ctx, src, tgt));
  const cvo_params& p = ctx->params;
  KConsts kc = make_kconsts(p, src->F, src->K, MODE_ALIGN);
  apply_tuning(ctx, kc);
  kc.mirror_pose = ctx->pose_arith == CVO_POSE_REFERENCE ? 1 : 0;
  kc.n_src_f = (float)src->n;
  kc.n_tgt_f = (float)tgt->n;
  kc.max_src_norm = src->max_norm;
  CVO_TRY(prepare_source(ctx, src));
  float tq = std::max(std::fabs(T_init[3]), std::max(std::fabs(T_init[7]), std::fabs(T_init[11])));
  GridHost gh;
  float ell_lo = std::min(p.ell_min, p.ell_init);
  CVO_TRY(build_grid(ctx, tgt, level_sizes(kc, std::max(p.ell_init, p.ell_min), ell_lo, src->max_norm, tq, kc.min_scale),
                     gh));
  const bool want_trace = ctx->trace_cap > 0 || ctx->profiling;
  const int trace_cap = want_trace ? std::max(ctx->trace_cap, ctx->profiling ? p.MAX_ITER : 0) : 0;
  Problems pr;
  // the first pass takes KDTREE_K_SIZE nearest with is_using_kdtree (CvoGPU.cu:1287)
  const int K0 = kc.knn > 0 ? kKnnMaxK : p.nearest_neighbors_max;
  CVO_TRY(setup_problems(ctx, src, tgt, gh, 1, std::max(K0, p.nearest_neighbors_max), trace_cap, kc.max_group, pr));
  CVO_TRY(ensure(ctx, ctx->b_tmp2, 64 * sizeof(float)));
  CVO_HIP(ctx, hipMemcpyAsync(ctx->b_tmp2.p, T_init, sizeof(float) * 16, hipMemcpyHostToDevice, ctx->stream));
  launch_init_state(pr.dev, 1, as<float>(ctx->b_tmp2), 16, kc, p.ell_init, K0, ctx->stream);
  CVO_HIP(ctx, hipGetLastError());
  double dev_s = 0.0;
  if (p.MAX_ITER > 0) CVO_TRY(run_loop(ctx, pr, kc, p.MAX_ITER, true, &dev_s));
  ProblemState hs;
  CVO_HIP(ctx, hipMemcpyAsync(&hs, pr.states, sizeof(ProblemState), hipMemcpyDeviceToHost, ctx->stream));
  CVO_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // T_out = [R^T, -R^T T] from the FP64 pose
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) T_out[r * 4 + c] = (float)hs.R[c * 3 + r];
    T_out[r * 4 + 3] = (float)(-(hs.R[0 * 3 + r] * hs.T[0] + hs.R[1 * 3 + r] * hs.T[1] + hs.R[2 * 3 + r] * hs.T[2]));
  }
  T_out[12] = T_out[13] = T_out[14] = 0.0f;
  T_out[15] = 1.0f;
  int n_it = p.MAX_ITER <= 0 ? 0 : (hs.done ? hs.iters_out : p.MAX_ITER);
  if (iters) *iters = n_it;
  if (ret_code) *ret_code = hs.ret;
  if (seconds) *seconds = dev_s;
  if (want_trace) {
    int n_rec = std::min(trace_cap, n_it + 1);
    ctx->trace.resize(n_rec);
    if (n_rec > 0)
      CVO_HIP(ctx, hipMemcpy(ctx->trace.data(), pr.trace, sizeof(cvo_iter_record) * n_rec, hipMemcpyDeviceToHost));
    ctx->trace_n = n_rec;
    if (ctx->profiling) account_bytes(ctx, src, tgt, kc, ctx->trace, n_rec, true);
  }
  if (assoc_out && p.is_exporting_association) CVO_TRY(export_assoc(ctx, src, tgt, pr, assoc_out));
  return CVO_OK;
}

// General 4x4 inverse (cofactors, FP64) rounded to float: the Eigen
// Matrix4f::inverse() of the reference's align overloads.
static void inverse4(const float m[16], float out[16]) {
  double a[16], inv[16];
  for (int i = 0; i < 16; ++i) a[i] = (double)m[i];
  inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] +
           a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
  inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] -
           a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
  inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] +
           a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
  inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] -
            a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
  inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] -
           a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
  inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] +
           a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
  inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] -
           a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
  inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] +
            a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
  inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] +
           a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
  inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] -
           a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
  inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] +
            a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
  inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] -
            a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
  inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] -
           a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
  inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] +
           a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
  inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] -
            a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
  inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] +
            a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
  const double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
  for (int i = 0; i < 16; ++i) out[i] = (float)(inv[i] / det);
}

// Approximate cos angles of one source / target pair at n poses (row-major
// 4x4 each, the inner-product convention), all in one batched association
// pass at lengthscale ell with the row cap nearest_neighbors_max: the
// function_angle scores of get_nearest_init_pose (CvoGPU.cpp:381-389).
static cvo_status score_poses(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, int n, const float* T16,
                              float ell, std::vector<double>& ip) {
  CVO_TRY(validate_pair(ctx, src, tgt));
  const cvo_params& p = ctx->params;
  KConsts kc = make_kconsts(p, src->F, src->K, MODE_IP);
  kc.knn = 0;
  apply_tuning(ctx, kc, n > 1);
  kc.n_src_f = (float)src->n;
  kc.n_tgt_f = (float)tgt->n;
  kc.max_src_norm = src->max_norm;
  kc.max_iter = 1;
  CVO_TRY(prepare_source(ctx, src));
  float tq = 0.0f;
  for (int s = 0; s < n; ++s)
    tq = std::max(tq, std::max(std::fabs(T16[16 * s + 3]), std::max(std::fabs(T16[16 * s + 7]), std::fabs(T16[16 * s + 11]))));
  GridHost gh;
  CVO_TRY(build_grid(ctx, tgt, level_sizes(kc, ell, ell, src->max_norm, tq, kc.min_scale), gh));
  Problems pr;
  CVO_TRY(setup_problems(ctx, src, tgt, gh, n, 0, 0, kc.max_group, pr));
  CVO_TRY(ensure(ctx, ctx->b_tmp2, sizeof(float) * 16 * (size_t)n + 64));
  CVO_HIP(ctx, hipMemcpyAsync(ctx->b_tmp2.p, T16, sizeof(float) * 16 * n, hipMemcpyHostToDevice, ctx->stream));
  launch_init_state(pr.dev, n, as<float>(ctx->b_tmp2), 16, kc, ell, p.nearest_neighbors_max, ctx->stream);
  CVO_TRY(run_loop(ctx, pr, kc, 1, false, nullptr));
  std::vector<ProblemState> hs(n);
  CVO_HIP(ctx, hipMemcpyAsync(hs.data(), pr.states, sizeof(ProblemState) * n, hipMemcpyDeviceToHost, ctx->stream));
  CVO_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ip.resize(n);
  for (int s = 0; s < n; ++s) ip[s] = hs[s].ip;
  return CVO_OK;
}

// get_nearest_init_pose (CvoGPU.cpp:367-398): the table rotation whose
// approximate cos angle is the first strict maximum above 0, as [R 0; 0 1];
// the identity when none is positive.  secs += the search's wall time.
static cvo_status nearest_init_pose(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, float ell,
                                    float T_out[16], double* secs) {
  const auto t0 = std::chrono::steady_clock::now();
  float rots[540];
  CVO_TRY(cvo_icosahedral_rotations(rots));
  // each candidate is scored at the inverse of [R 0; 0 1], i.e. [R^T 0; 0 1]
  std::vector<float> T(60 * 16, 0.0f);
  for (int i = 0; i < 60; ++i) {
    float* t = &T[16 * i];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) t[r * 4 + c] = rots[9 * i + c * 3 + r];
    t[15] = 1.0f;
  }
  std::vector<double> ip;
  CVO_TRY(score_poses(ctx, src, tgt, 60, T.data(), ell, ip));
  // function_angle in float: fxfz / (sqrt(N) sqrt(M)) (CvoGPU.cu:1785-1802)
  const float fx = (float)std::sqrt((double)src->n), fz = (float)std::sqrt((double)tgt->n);
  float best = 0.0f;
  int best_i = -1;
  for (int i = 0; i < 60; ++i) {
    const float c = (float)ip[i] / (fx * fz);
    if (c > best) {
      best = c;
      best_i = i;
    }
  }
  for (int i = 0; i < 16; ++i) T_out[i] = (i % 5 == 0) ? 1.0f : 0.0f;
  if (best_i >= 0)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) T_out[r * 4 + c] = rots[9 * best_i + r * 3 + c];
  if (secs) *secs += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return CVO_OK;
}

extern "C" cvo_status cvo_get_nearest_init_pose_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt,
                                                  float ell, int32_t num_guess, float T_out[16], double* seconds) {
  (void)num_guess;
  if (!ctx || !T_out) return CVO_ERR_INVALID_ARG;
  for (int i = 0; i < 16; ++i) T_out[i] = (i % 5 == 0) ? 1.0f : 0.0f;
  if (!(ell > 0.0f)) return set_err(ctx, CVO_ERR_INVALID_ARG, "ell must be positive");
  if (!src || !tgt || src->n == 0 || tgt->n == 0) return set_err(ctx, CVO_ERR_EMPTY, "empty point cloud");
  return nearest_init_pose(ctx, src, tgt, ell, T_out, seconds);
}

extern "C" cvo_status cvo_align_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, const float T_init[16],
                                  float T_out[16], cvo_assoc* assoc_out, double* seconds, int32_t* iters,
                                  int32_t* ret_code) {
  if (!ctx || !T_init || !T_out) return CVO_ERR_INVALID_ARG;
  const cvo_params& p = ctx->params;
  if (p.is_global_angle_registration > 0 && src && tgt && src->n > 0 && tgt->n > 0) {
    // init = get_nearest_init_pose(...).inverse() (CvoGPU.cu:1531-1543)
    double search = 0.0;
    float Tg[16], Ti[16];
    CVO_TRY(nearest_init_pose(ctx, src, tgt, p.ell_init_first_frame, Tg, &search));
    inverse4(Tg, Ti);
    CVO_TRY(align_impl_d(ctx, src, tgt, Ti, T_out, assoc_out, seconds, iters, ret_code));
    if (seconds) *seconds += search;
    return CVO_OK;
  }
  return align_impl_d(ctx, src, tgt, T_init, T_out, assoc_out, seconds, iters, ret_code);
}

extern "C" cvo_status cvo_align_s2t_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt,
                                      const float T_init_s2t[16], float T_out[16], cvo_assoc* assoc_out,
                                      double* seconds, int32_t* iters, int32_t* ret_code) {
  if (!ctx || !T_init_s2t || !T_out) return CVO_ERR_INVALID_ARG;
  if (!src || !tgt || src->n == 0 || tgt->n == 0) {
    if (ret_code) *ret_code = -1;
    if (iters) *iters = 0;
    if (seconds) *seconds = 0.0;
    for (int i = 0; i < 16; ++i) T_out[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    return set_err(ctx, CVO_ERR_EMPTY, "empty point cloud");
  }
  const cvo_params& p = ctx->params;
  float T0[16];
  double search = 0.0;
  if (p.is_global_angle_registration > 0) {
    // init = get_nearest_init_pose(...) itself, not its inverse (CvoGPU.cu:1580-1591)
    CVO_TRY(nearest_init_pose(ctx, src, tgt, p.ell_init_first_frame, T0, &search));
  } else {
    inverse4(T_init_s2t, T0);  // init_guess_T_s2t.inverse() (:1578)
  }
  CVO_TRY(align_impl_d(ctx, src, tgt, T0, T_out, assoc_out, seconds, iters, ret_code));
  if (seconds) *seconds += search;
  return CVO_OK;
}

// RAII holder for uploaded clouds of a host-view call.
struct Uploaded {
  cvo_ctx* ctx;
  cvo_dcloud* c = nullptr;
  ~Uploaded() {
    if (c) cvo_cloud_free(ctx, c);
  }
};

extern "C" cvo_status cvo_align(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                const float T_init[16], float T_out[16], cvo_assoc* assoc_out, double* seconds,
                                int32_t* iters, int32_t* ret_code) {
  if (!ctx || !src || !tgt || !T_init || !T_out) return CVO_ERR_INVALID_ARG;
  if (src->n == 0 || tgt->n == 0) return cvo_align_d(ctx, nullptr, nullptr, T_init, T_out, assoc_out, seconds, iters, ret_code);
  Uploaded a{ctx}, b{ctx};
  CVO_TRY(cvo_cloud_upload(ctx, src, &a.c));
  CVO_TRY(cvo_cloud_upload(ctx, tgt, &b.c));
  return cvo_align_d(ctx, a.c, b.c, T_init, T_out, assoc_out, seconds, iters, ret_code);
}

extern "C" cvo_status cvo_align_s2t(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                    const float T_init_s2t[16], float T_out[16], cvo_assoc* assoc_out,
                                    double* seconds, int32_t* iters, int32_t* ret_code) {
  if (!ctx || !src || !tgt || !T_init_s2t || !T_out) return CVO_ERR_INVALID_ARG;
  if (src->n == 0 || tgt->n == 0)
    return cvo_align_s2t_d(ctx, nullptr, nullptr, T_init_s2t, T_out, assoc_out, seconds, iters, ret_code);
  Uploaded a{ctx}, b{ctx};
  CVO_TRY(cvo_cloud_upload(ctx, src, &a.c));
  CVO_TRY(cvo_cloud_upload(ctx, tgt, &b.c));
  return cvo_align_s2t_d(ctx, a.c, b.c, T_init_s2t, T_out, assoc_out, seconds, iters, ret_code);
}

extern "C" cvo_status cvo_get_nearest_init_pose(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                                float ell, int32_t num_guess, float T_out[16], double* seconds) {
  if (!ctx || !src || !tgt || !T_out) return CVO_ERR_INVALID_ARG;
  if (src->n == 0 || tgt->n == 0)
    return cvo_get_nearest_init_pose_d(ctx, nullptr, nullptr, ell, num_guess, T_out, seconds);
  Uploaded a{ctx}, b{ctx};
  CVO_TRY(cvo_cloud_upload(ctx, src, &a.c));
  CVO_TRY(cvo_cloud_upload(ctx, tgt, &b.c));
  return cvo_get_nearest_init_pose_d(ctx, a.c, b.c, ell, num_guess, T_out, seconds);
}

// Single association pass at pose T and lengthscale ell; the problem stays in
// the context's buffers for an export.
// knn_K > 0: the K-nearest pass with row cap knn_K instead of the ball pass.
static KConsts adaptive_kconsts(const cvo_params& p, int F, int Ks);

// kc_in: the pass's kernel constants (mode MODE_IP) instead of the
// reference's association.
// keep_pairs false: the pass only sums (nothing to export afterwards).
static cvo_status single_pass(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, const float T[16], float ell,
                              Problems& pr, KConsts& kc, double* ip, int knn_K = 0, const KConsts* kc_in = nullptr,
                              bool keep_pairs = true) {
  CVO_TRY(validate_pair(ctx, src, tgt));
  const cvo_params& p = ctx->params;
  kc = kc_in ? *kc_in : make_kconsts(p, src->F, src->K, MODE_IP);
  kc.knn = knn_K > 0 ? 1 : 0;
  const int K0 = knn_K > 0 ? knn_K : p.nearest_neighbors_max;
  apply_tuning(ctx, kc);
  kc.n_src_f = (float)src->n;
  kc.n_tgt_f = (float)tgt->n;
  kc.max_src_norm = src->max_norm;
  kc.max_iter = 1;
  CVO_TRY(prepare_source(ctx, src));
  float tq = std::max(std::fabs(T[3]), std::max(std::fabs(T[7]), std::fabs(T[11])));
  GridHost gh;
  std::vector<float> hs = level_sizes(kc, ell, ell, src->max_norm, tq, kc.min_scale);
  CVO_TRY(build_grid(ctx, tgt, hs, gh));
  CVO_TRY(setup_problems(ctx, src, tgt, gh, 1, keep_pairs ? K0 : 0, 0, kc.max_group, pr));
  CVO_TRY(ensure(ctx, ctx->b_tmp2, 64 * sizeof(float)));
  CVO_HIP(ctx, hipMemcpyAsync(ctx->b_tmp2.p, T, sizeof(float) * 16, hipMemcpyHostToDevice, ctx->stream));
  launch_init_state(pr.dev, 1, as<float>(ctx->b_tmp2), 16, kc, ell, K0, ctx->stream);
  CVO_TRY(run_loop(ctx, pr, kc, 1, false, nullptr));
  ProblemState hs2;
  CVO_HIP(ctx, hipMemcpyAsync(&hs2, pr.states, sizeof(ProblemState), hipMemcpyDeviceToHost, ctx->stream));
  CVO_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (ip) *ip = hs2.ip;
  return CVO_OK;
}

extern "C" cvo_status cvo_inner_product_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, const float T[16],
                                          float ell, double* ip) {
  if (!ctx || !T || !ip) return CVO_ERR_INVALID_ARG;
  *ip = 0.0;
  if (!src || !tgt || src->n == 0 || tgt->n == 0) return set_err(ctx, CVO_ERR_EMPTY, "empty point cloud");
  Problems pr;
  KConsts kc;
  return single_pass(ctx, src, tgt, T, ell, pr, kc, ip, 0, nullptr, false);
}

// Kernel constants of inner_product_cpu (se_kernel_init_ell_cpu,
// CvoGPU.cpp:114-186): the radius gate always (nanoflann radius search with
// squared radius -2 l^2 log(sp / sigma^2)), the fixed lengthscale, every pair
// inside it summed (no row cap), the factors of the enabled terms without
// their gates, no geometric type.
static KConsts cpu_ip_kconsts(const cvo_params& p, int F, int Ks) {
  KConsts k = make_kconsts(p, F, Ks, MODE_IP);
  k.knn = 0;
  k.fixed_ell = 1;
  k.uncapped = 1;
  k.use_gtype = 0;
  k.d2_c_thres = INFINITY;
  k.d2_s_thres = INFINITY;
  if (!k.use_geo) {
    // the radius search still gates; k stays 1
    k.use_geo = 1;
    k.no_geo_factor = 1;
    k.rq_coef = sqrtf(-2.0f * k.lg) * 1.001f;
  }
  return k;
}

static cvo_status inner_product_cpu_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, const float T[16],
                                      float ell, double* ip) {
  const KConsts kc_in = cpu_ip_kconsts(ctx->params, src->F, src->K);
  Problems pr;
  KConsts kc;
  return single_pass(ctx, src, tgt, T, ell, pr, kc, ip, 0, &kc_in, false);
}

extern "C" cvo_status cvo_inner_product_cpu(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                            const float T[16], float ell, double* ip) {
  if (!ctx || !src || !tgt || !T || !ip) return CVO_ERR_INVALID_ARG;
  *ip = 0.0;
  if (src->n == 0 || tgt->n == 0) return CVO_OK;  // inner_product_cpu returns 0 (CvoGPU.cpp:209-211)
  Uploaded a{ctx}, b{ctx};
  CVO_TRY(cvo_cloud_upload(ctx, src, &a.c));
  CVO_TRY(cvo_cloud_upload(ctx, tgt, &b.c));
  return inner_product_cpu_d(ctx, a.c, b.c, T, ell, ip);
}

extern "C" cvo_status cvo_inner_product(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                        const float T[16], float ell, double* ip) {
  if (!ctx || !src || !tgt || !T || !ip) return CVO_ERR_INVALID_ARG;
  *ip = 0.0;
  if (src->n == 0 || tgt->n == 0) return set_err(ctx, CVO_ERR_EMPTY, "empty point cloud");
  Uploaded a{ctx}, b{ctx};
  CVO_TRY(cvo_cloud_upload(ctx, src, &a.c));
  CVO_TRY(cvo_cloud_upload(ctx, tgt, &b.c));
  return cvo_inner_product_d(ctx, a.c, b.c, T, ell, ip);
}

extern "C" cvo_status cvo_function_angle(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                         const float T[16], float ell, int32_t approx, double* cos_out) {
  return cvo_function_angle_ex(ctx, src, tgt, T, ell, approx, 1, cos_out);
}

extern "C" cvo_status cvo_function_angle_ex(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                            const float T[16], float ell, int32_t approx, int32_t is_gpu,
                                            double* cos_out) {
  if (!ctx || !src || !tgt || !T || !cos_out) return CVO_ERR_INVALID_ARG;
  *cos_out = 0.0;
  // function_angle returns 0 on empty input (CvoGPU.cu:1775-1777)
  if (src->n == 0 || tgt->n == 0) return CVO_OK;
  Uploaded a{ctx}, b{ctx};
  CVO_TRY(cvo_cloud_upload(ctx, src, &a.c));
  CVO_TRY(cvo_cloud_upload(ctx, tgt, &b.c));
  // is_gpu selects inner_product_gpu or inner_product_cpu (CvoGPU.cu:1780-1783)
  auto inner = [&](const cvo_dcloud* x, const cvo_dcloud* y, const float* Tm, double* out) -> cvo_status {
    return is_gpu ? cvo_inner_product_d(ctx, x, y, Tm, ell, out) : inner_product_cpu_d(ctx, x, y, Tm, ell, out);
  };
  double fxfz = 0.0;
  CVO_TRY(inner(a.c, b.c, T, &fxfz));
  double fx, fz;
  if (approx) {
    fx = std::sqrt((double)src->n);
    fz = std::sqrt((double)tgt->n);
  } else {
    const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    double xx = 0, zz = 0;
    CVO_TRY(inner(a.c, a.c, I, &xx));
    CVO_TRY(inner(b.c, b.c, I, &zz));
    fx = std::sqrt(xx);
    fz = std::sqrt(zz);
  }
  *cos_out = fxfz / (fx * fz);
  return CVO_OK;
}

extern "C" cvo_status cvo_compute_association(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                              const float T[16], float ell, cvo_assoc* out) {
  if (!ctx || !src || !tgt || !T || !out) return CVO_ERR_INVALID_ARG;
  std::memset(out, 0, sizeof(*out));
  if (src->n == 0 || tgt->n == 0) return CVO_OK;  // the reference returns without output
  Uploaded a{ctx}, b{ctx};
  CVO_TRY(cvo_cloud_upload(ctx, src, &a.c));
  CVO_TRY(cvo_cloud_upload(ctx, tgt, &b.c));
  Problems pr;
  KConsts kc;
  double ip = 0.0;
  CVO_TRY(single_pass(ctx, a.c, b.c, T, ell, pr, kc, &ip));
  return export_assoc(ctx, a.c, b.c, pr, out);
}

extern "C" cvo_status cvo_compute_association_knn(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                                  const float T[16], float ell, int32_t K, cvo_assoc* out) {
  if (!ctx || !src || !tgt || !T || !out) return CVO_ERR_INVALID_ARG;
  std::memset(out, 0, sizeof(*out));
  if (K < 1 || K > kKnnMaxK) return set_err(ctx, CVO_ERR_LIMIT, "K must be in [1, 512]");
  if (!(ell > 0.0f)) return set_err(ctx, CVO_ERR_INVALID_ARG, "ell must be positive");
  if (!ctx->params.is_using_geometry)
    return set_err(ctx, CVO_ERR_UNSUPPORTED, "the K-nearest association needs is_using_geometry");
  if (src->n == 0 || tgt->n == 0) return CVO_OK;
  Uploaded a{ctx}, b{ctx};
  CVO_TRY(cvo_cloud_upload(ctx, src, &a.c));
  CVO_TRY(cvo_cloud_upload(ctx, tgt, &b.c));
  Problems pr;
  KConsts kc;
  double ip = 0.0;
  CVO_TRY(single_pass(ctx, a.c, b.c, T, ell, pr, kc, &ip, K));
  return export_assoc(ctx, a.c, b.c, pr, out);
}

extern "C" void cvo_assoc_free(cvo_assoc* a) {
  if (!a) return;
  std::free(a->row_ptr);
  std::free(a->col);
  std::free(a->val);
  a->row_ptr = nullptr;
  a->col = nullptr;
  a->val = nullptr;
  a->nnz = 0;
}

// ---------------------------------------------------------------------------
// Batched multi-start global registration
// ---------------------------------------------------------------------------

static void pose_to_output(const ProblemState& hs, float* T_out) {
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) T_out[r * 4 + c] = (float)hs.R[c * 3 + r];
    T_out[r * 4 + 3] = (float)(-(hs.R[0 * 3 + r] * hs.T[0] + hs.R[1 * 3 + r] * hs.T[1] + hs.R[2 * 3 + r] * hs.T[2]));
  }
  T_out[12] = T_out[13] = T_out[14] = 0.0f;
  T_out[15] = 1.0f;
}

extern "C" cvo_status cvo_align_multistart(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, int32_t n_starts,
                                           const float* T_starts, float score_ell, int32_t mode, float T_best[16],
                                           int32_t* best_index, float* per_start_T, double* per_start_cos,
                                           int32_t* per_start_iters, double* seconds) {
  if (!ctx || !T_starts || !T_best || n_starts <= 0 || !(score_ell > 0)) return CVO_ERR_INVALID_ARG;
  if (n_starts > 4096) return set_err(ctx, CVO_ERR_LIMIT, "at most 4096 starts per call");
  if (!src || !tgt || src->n == 0 || tgt->n == 0) return set_err(ctx, CVO_ERR_EMPTY, "empty point cloud");
  CVO_TRY(validate_pair(ctx, src, tgt));
  const cvo_params& p = ctx->params;
  CVO_TRY(ensure_events(ctx, 8));
  hipEvent_t e0 = ctx->ev_pool[4], e1 = ctx->ev_pool[5];
  CVO_HIP(ctx, hipEventRecord(e0, ctx->stream));
  KConsts kc = make_kconsts(p, src->F, src->K, MODE_ALIGN);
  apply_tuning(ctx, kc, true);  // batched launches
  kc.mirror_pose = ctx->pose_arith == CVO_POSE_REFERENCE ? 1 : 0;
  kc.n_src_f = (float)src->n;
  kc.n_tgt_f = (float)tgt->n;
  kc.max_src_norm = src->max_norm;
  KConsts kci = kc;
  kci.mode = MODE_IP;
  kci.knn = 0;  // scoring is an inner product: the ball association
  CVO_TRY(prepare_source(ctx, src));
  float tq = 0.0f;
  for (int s = 0; s < n_starts; ++s)
    tq = std::max(tq, std::max(std::fabs(T_starts[16 * s + 3]),
                               std::max(std::fabs(T_starts[16 * s + 7]), std::fabs(T_starts[16 * s + 11]))));
  const float ell_hi = std::max(std::max(p.ell_init, p.ell_min), score_ell);
  const float ell_lo = std::min(std::min(p.ell_min, p.ell_init), score_ell);
  GridHost gh;
  CVO_TRY(build_grid(ctx, tgt, level_sizes(kc, ell_hi, ell_lo, src->max_norm, tq, kc.min_scale), gh));
  const int n_run = (mode == 1) ? n_starts : n_starts;
  Problems pr;
  const int K0 = kc.knn > 0 ? kKnnMaxK : p.nearest_neighbors_max;
  CVO_TRY(setup_problems(ctx, src, tgt, gh, n_run, std::max(K0, p.nearest_neighbors_max), 0, kc.max_group, pr));
  CVO_TRY(ensure(ctx, ctx->b_tmp2, sizeof(float) * 16 * (size_t)n_starts + 64));
  CVO_HIP(ctx, hipMemcpyAsync(ctx->b_tmp2.p, T_starts, sizeof(float) * 16 * n_starts, hipMemcpyHostToDevice,
                              ctx->stream));
  launch_init_state(pr.dev, n_starts, as<float>(ctx->b_tmp2), 16, kc, p.ell_init, K0, ctx->stream);
  CVO_HIP(ctx, hipGetLastError());
  std::vector<ProblemState> hs(n_starts);
  std::vector<double> cosv(n_starts, 0.0);
  std::vector<int32_t> its(n_starts, 0);
  const double norm = std::sqrt((double)src->n) * std::sqrt((double)tgt->n);
  int best = 0;
  if (mode == 0) {
    if (p.MAX_ITER > 0) CVO_TRY(run_loop(ctx, pr, kc, p.MAX_ITER, true, nullptr));
    CVO_HIP(ctx, hipMemcpyAsync(hs.data(), pr.states, sizeof(ProblemState) * n_starts, hipMemcpyDeviceToHost,
                                ctx->stream));
    CVO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int s = 0; s < n_starts; ++s) its[s] = p.MAX_ITER <= 0 ? 0 : (hs[s].done ? hs[s].iters_out : p.MAX_ITER);
    launch_rescore(pr.dev, n_starts, kci, score_ell, p.nearest_neighbors_max, ctx->stream);
    CVO_TRY(run_loop(ctx, pr, kci, 1, false, nullptr));
    CVO_HIP(ctx, hipMemcpyAsync(hs.data(), pr.states, sizeof(ProblemState) * n_starts, hipMemcpyDeviceToHost,
                                ctx->stream));
    CVO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int s = 0; s < n_starts; ++s) {
      cosv[s] = hs[s].ip / norm;
      if (cosv[s] > cosv[best]) best = s;
      if (per_start_T) pose_to_output(hs[s], per_start_T + 16 * (size_t)s);
    }
    pose_to_output(hs[best], T_best);
  } else {
    // the reference's get_nearest_init_pose (CvoGPU.cpp:367-398): score the
    // initial poses only, keep the first maximum, then one align from it
    launch_rescore(pr.dev, n_starts, kci, score_ell, p.nearest_neighbors_max, ctx->stream);
    CVO_TRY(run_loop(ctx, pr, kci, 1, false, nullptr));
    CVO_HIP(ctx, hipMemcpyAsync(hs.data(), pr.states, sizeof(ProblemState) * n_starts, hipMemcpyDeviceToHost,
                                ctx->stream));
    CVO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double best_ip = 0.0;
    best = 0;
    for (int s = 0; s < n_starts; ++s) {
      cosv[s] = hs[s].ip / norm;
      if (hs[s].ip > best_ip) {
        best_ip = hs[s].ip;
        best = s;
      }
      if (per_start_T) pose_to_output(hs[s], per_start_T + 16 * (size_t)s);
    }
    int32_t it = 0, ret = 0;
    double secs = 0;
    CVO_TRY(align_impl_d(ctx, src, tgt, T_starts + 16 * (size_t)best, T_best, nullptr, &secs, &it, &ret));
    its[best] = it;
  }
  CVO_HIP(ctx, hipEventRecord(e1, ctx->stream));
  CVO_HIP(ctx, hipEventSynchronize(e1));
  float ms = 0;
  CVO_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
  if (seconds) *seconds = ms / 1000.0;
  if (best_index) *best_index = best;
  if (per_start_cos)
    for (int s = 0; s < n_starts; ++s) per_start_cos[s] = cosv[s];
  if (per_start_iters)
    for (int s = 0; s < n_starts; ++s) per_start_iters[s] = its[s];
  return CVO_OK;
}

// ---------------------------------------------------------------------------
// Adaptive-lengthscale registration (AdaptiveCvoGPU::align)
// ---------------------------------------------------------------------------

// Kernel constants of AdaptiveCvoGPU's association (AdaptiveCvoGPU.cu:263-408):
// the fixed-lengthscale distance gate, the colour gate always on (its
// threshold as written there), the class factor without a gate, no geometric
// type; row cap min(KDTREE_K_SIZE, nearest_neighbors_max).
static KConsts adaptive_kconsts(const cvo_params& p, int F, int Ks) {
  KConsts k = make_kconsts(p, F, Ks, MODE_ALIGN);
  k.knn = 0;
  k.adaptive = 1;
  k.fixed_ell = 1;
  k.ip_ad2 = 1;
  k.dl_step = p.dl_step;
  k.ell_max0 = p.ell_max;
  k.use_geo = 1;
  k.use_col = 1;
  k.use_gtype = 0;
  k.d2_c_thres = (float)(-2.0 * (double)p.c_ell * (double)p.c_ell *
                         (double)logf(p.sp_thres / p.c_sigma / p.c_sigma));
  k.d2_s_thres = INFINITY;
  k.lg = logf(p.sp_thres / k.sigma2);
  k.rq_coef = sqrtf(-2.0f * k.lg) * 1.001f;
  k.amp = k.sigma2 * k.c_sigma2 * (k.use_sem ? k.s_sigma2 : 1.0f);
  k.range_ell_in_step = 0;
  return k;
}

extern "C" cvo_status cvo_align_adaptive_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt,
                                           const float T_init[16], float T_out[16], double* seconds, int32_t* iters,
                                           int32_t* ret_code) {
  if (!ctx || !T_init || !T_out) return CVO_ERR_INVALID_ARG;
  if (ret_code) *ret_code = 0;
  if (iters) *iters = 0;
  if (seconds) *seconds = 0.0;
  for (int i = 0; i < 16; ++i) T_out[i] = (i % 5 == 0) ? 1.0f : 0.0f;
  if (!src || !tgt || src->n == 0 || tgt->n == 0) return set_err(ctx, CVO_ERR_EMPTY, "empty point cloud");
  CVO_TRY(validate_pair(ctx, src, tgt));
  if (!src->feat || !tgt->feat || src->F != tgt->F || src->F <= 0)
    return set_err(ctx, CVO_ERR_INVALID_ARG, "the adaptive-lengthscale association needs features on both clouds");
  const cvo_params& p = ctx->params;
  if (!(p.ell_init > 0.0f) || !(p.ell_min > 0.0f) || !(p.ell_max > 0.0f))
    return set_err(ctx, CVO_ERR_INVALID_ARG, "lengthscales must be positive");
  if (p.is_using_semantics && (!src->sem || !tgt->sem || src->K != tgt->K))
    return set_err(ctx, CVO_ERR_INVALID_ARG, "is_using_semantics needs class vectors on both clouds");
  KConsts kc = adaptive_kconsts(p, src->F, p.is_using_semantics ? src->K : 0);
  if (!p.is_using_semantics) kc.use_sem = 0;
  kc.amp = kc.sigma2 * kc.c_sigma2 * (kc.use_sem ? kc.s_sigma2 : 1.0f);
  apply_tuning(ctx, kc, true);  // batched launches
  kc.mirror_pose = ctx->pose_arith == CVO_POSE_REFERENCE ? 1 : 0;
  kc.n_src_f = (float)src->n;
  kc.n_tgt_f = (float)tgt->n;
  kc.max_src_norm = std::max(src->max_norm, tgt->max_norm);
  const int K = std::min(kKnnMaxK, p.nearest_neighbors_max);
  hipStream_t st = ctx->stream;
  // sorted rows: source into the context's buffers, target into its own
  CVO_TRY(prepare_source(ctx, src));
  DevBuf* B = ctx->b_acvo;
  CVO_TRY(prepare_source_into(ctx, tgt, B[0], B[1], B[2], B[3]));
  const float tq = std::max(std::fabs(T_init[3]), std::max(std::fabs(T_init[7]), std::fabs(T_init[11])));
  const float ell_hi = std::max(std::max(p.ell_init, p.ell_max), p.ell_min);
  const float ell_lo = std::min(p.ell_min, p.ell_init);
  GridHost gy, gx;
  CVO_TRY(build_grid(ctx, tgt, level_sizes(kc, ell_hi, ell_lo, kc.max_src_norm, tq, kc.min_scale), gy));
  CVO_TRY(build_grid_into(ctx, src, level_sizes(kc, ell_hi, ell_lo, kc.max_src_norm, 0.0f, kc.min_scale), gx, B[4],
                          B[5]));
  // three problems: x against T^-1 y (the registration), x against x, and
  // T^-1 y against itself
  const int rb = rows_per_block(kc.max_group);
  const int rows[3] = {src->n, src->n, tgt->n};
  int nb[3], blk0[3], total = 0;
  size_t row0[3], nrows = 0;
  for (int q = 0; q < 3; ++q) {
    nb[q] = (rows[q] + rb - 1) / rb;
    blk0[q] = total;
    total += nb[q];
    row0[q] = nrows;
    nrows += (size_t)rows[q];
  }
  const bool want_trace = ctx->trace_cap > 0 || ctx->profiling;
  const int trace_cap = want_trace ? std::max(ctx->trace_cap, ctx->profiling ? p.MAX_ITER : 0) : 0;
  CVO_TRY(ensure(ctx, ctx->b_pairs, sizeof(uint2) * nrows * K));
  CVO_TRY(ensure(ctx, ctx->b_counts, sizeof(uint32_t) * nrows));
  CVO_TRY(ensure(ctx, ctx->b_ovf, sizeof(uint32_t) * nrows));
  CVO_TRY(ensure(ctx, ctx->b_state, sizeof(ProblemState) * 3));
  CVO_TRY(ensure(ctx, ctx->b_probs, sizeof(AssocProblem) * 3));
  CVO_TRY(ensure(ctx, ctx->b_part_a, sizeof(double) * 8 * (size_t)total));
  CVO_TRY(ensure(ctx, ctx->b_part_ovf, sizeof(double) * 10 * (size_t)kOvfBlocks * 3));
  CVO_TRY(ensure(ctx, ctx->b_part_s, sizeof(double) * 4 * (size_t)total));
  CVO_TRY(ensure(ctx, ctx->b_part_cnt, sizeof(uint32_t) * 2 * (size_t)total));
  if (trace_cap > 0) CVO_TRY(ensure(ctx, ctx->b_trace, sizeof(cvo_iter_record) * (size_t)trace_cap));
  launch_zero(ctx->b_ovf.p, sizeof(uint32_t) * nrows, st);
  launch_zero(ctx->b_state.p, sizeof(ProblemState) * 3, st);
  Problems pr;
  pr.n_probs = 3;
  pr.total_blocks = total;
  pr.max_rows = std::max(src->n, tgt->n);
  pr.states = as<ProblemState>(ctx->b_state);
  pr.trace = trace_cap > 0 ? as<cvo_iter_record>(ctx->b_trace) : nullptr;
  pr.trace_cap = trace_cap;
  pr.host.resize(3);
  const cvo_dcloud* tgts[3] = {tgt, src, tgt};
  const GridHost* grids[3] = {&gy, &gx, &gy};
  for (int q = 0; q < 3; ++q) {
    AssocProblem& P = pr.host[q];
    std::memset(&P, 0, sizeof(P));
    const bool from_tgt = q == 2;
    const cvo_dcloud* s = from_tgt ? tgt : src;
    P.src_pos = from_tgt ? as<float4>(B[0]) : as<float4>(ctx->b_src_pos);
    P.src_feat = from_tgt ? as<float>(B[1]) : as<float>(ctx->b_src_feat);
    P.src_sem = s->sem ? (from_tgt ? as<float>(B[2]) : as<float>(ctx->b_src_sem)) : nullptr;
    P.src_geo = s->geo ? (from_tgt ? as<float>(B[3]) : as<float>(ctx->b_src_geo)) : nullptr;
    P.n_src = rows[q];
    P.n_tgt = tgts[q]->n;
    P.grid = grids[q]->dev_ptr;
    P.tgt_xyz = tgts[q]->xyz;
    P.tgt_feat = tgts[q]->feat;
    P.tgt_sem = tgts[q]->sem;
    P.tgt_geo = tgts[q]->geo;
    P.st = pr.states + q;
    P.pairs = as<uint2>(ctx->b_pairs) + row0[q] * K;
    P.counts = as<uint32_t>(ctx->b_counts) + row0[q];
    P.ovf_stamp = as<uint32_t>(ctx->b_ovf) + row0[q];
    P.cap_alloc = K;
    P.block_begin = blk0[q];
    P.n_blocks = nb[q];
    P.trace_cap = q == 0 ? trace_cap : 0;
    P.part_a = as<double>(ctx->b_part_a) + (size_t)blk0[q] * 8;
    P.part_ovf = as<double>(ctx->b_part_ovf) + (size_t)q * kOvfBlocks * 10;
    P.part_s = as<double>(ctx->b_part_s) + (size_t)blk0[q] * 4;
    P.part_cnt = as<uint32_t>(ctx->b_part_cnt) + (size_t)blk0[q] * 2;
    P.trace = q == 0 ? pr.trace : nullptr;
    P.xx_st = q == 0 ? pr.states + 1 : nullptr;
    P.yy_st = q == 0 ? pr.states + 2 : nullptr;
  }
  pr.dev = as<AssocProblem>(ctx->b_probs);
  CVO_HIP(ctx, hipMemcpyAsync(pr.dev, pr.host.data(), sizeof(AssocProblem) * 3, hipMemcpyHostToDevice, st));
  CVO_TRY(ensure(ctx, ctx->b_tmp2, 64 * sizeof(float)));
  CVO_HIP(ctx, hipMemcpyAsync(ctx->b_tmp2.p, T_init, sizeof(float) * 16, hipMemcpyHostToDevice, st));
  launch_init_state(pr.dev, 1, as<float>(ctx->b_tmp2), 16, kc, p.ell_init, K, st);
  launch_adaptive_sync(pr.dev, kc, st);
  CVO_HIP(ctx, hipGetLastError());
  double dev_s = 0.0;
  if (p.MAX_ITER > 0) CVO_TRY(run_loop(ctx, pr, kc, p.MAX_ITER, true, &dev_s));
  ProblemState hs;
  CVO_HIP(ctx, hipMemcpyAsync(&hs, pr.states, sizeof(ProblemState), hipMemcpyDeviceToHost, st));
  CVO_HIP(ctx, hipStreamSynchronize(st));
  pose_to_output(hs, T_out);
  const int n_it = p.MAX_ITER <= 0 ? 0 : (hs.done ? hs.iters_out : p.MAX_ITER);
  if (iters) *iters = n_it;
  if (ret_code) *ret_code = hs.ret;
  if (seconds) *seconds = dev_s;
  if (want_trace) {
    const int n_rec = std::min(trace_cap, n_it + 1);
    ctx->trace.resize(n_rec);
    if (n_rec > 0)
      CVO_HIP(ctx, hipMemcpy(ctx->trace.data(), pr.trace, sizeof(cvo_iter_record) * n_rec, hipMemcpyDeviceToHost));
    ctx->trace_n = n_rec;
    if (ctx->profiling) account_bytes_adaptive(ctx, src, tgt, kc, ctx->trace, n_rec);
  }
  return CVO_OK;
}

extern "C" cvo_status cvo_align_adaptive(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                         const float T_init[16], float T_out[16], double* seconds, int32_t* iters,
                                         int32_t* ret_code) {
  if (!ctx || !src || !tgt || !T_init || !T_out) return CVO_ERR_INVALID_ARG;
  if (src->n == 0 || tgt->n == 0) return cvo_align_adaptive_d(ctx, nullptr, nullptr, T_init, T_out, seconds, iters,
                                                              ret_code);
  Uploaded a{ctx}, b{ctx};
  CVO_TRY(cvo_cloud_upload(ctx, src, &a.c));
  CVO_TRY(cvo_cloud_upload(ctx, tgt, &b.c));
  return cvo_align_adaptive_d(ctx, a.c, b.c, T_init, T_out, seconds, iters, ret_code);
}

extern "C" cvo_status cvo_inner_product_adaptive(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                                 const float T_s2t[16], float* out) {
  if (!ctx || !src || !tgt || !T_s2t || !out) return CVO_ERR_INVALID_ARG;
  *out = 0.0f;
  if (src->n == 0 || tgt->n == 0) return CVO_OK;  // the reference returns 0
  Uploaded a{ctx}, b{ctx};
  CVO_TRY(cvo_cloud_upload(ctx, src, &a.c));
  CVO_TRY(cvo_cloud_upload(ctx, tgt, &b.c));
  if (!a.c->feat || !b.c->feat || a.c->F != b.c->F)
    return set_err(ctx, CVO_ERR_INVALID_ARG, "the adaptive-lengthscale association needs features on both clouds");
  const cvo_params& p = ctx->params;
  if (p.is_using_semantics && (!a.c->sem || !b.c->sem || a.c->K != b.c->K))
    return set_err(ctx, CVO_ERR_INVALID_ARG, "is_using_semantics needs class vectors on both clouds");
  KConsts kc = adaptive_kconsts(p, a.c->F, p.is_using_semantics ? a.c->K : 0);
  if (!p.is_using_semantics) kc.use_sem = 0;
  kc.amp = kc.sigma2 * kc.c_sigma2 * (kc.use_sem ? kc.s_sigma2 : 1.0f);
  kc.mode = MODE_IP;
  kc.adaptive = 0;
  kc.ip_ad2 = 0;
  kc.uncapped = 1;
  // the target moves by T itself (rot y + trans); a pass moves it by the
  // inverse of its pose, so the pass gets T^-1
  float Ti[16];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) Ti[r * 4 + c] = T_s2t[c * 4 + r];
    Ti[r * 4 + 3] = (float)(-((double)T_s2t[0 * 4 + r] * T_s2t[3] + (double)T_s2t[1 * 4 + r] * T_s2t[7] +
                              (double)T_s2t[2 * 4 + r] * T_s2t[11]));
  }
  Ti[12] = Ti[13] = Ti[14] = 0.0f;
  Ti[15] = 1.0f;
  Problems pr;
  KConsts kcu;
  double ip = 0.0;
  CVO_TRY(single_pass(ctx, a.c, b.c, Ti, p.ell_init, pr, kcu, &ip, 0, &kc));
  *out = (float)(ip / (double)src->n * 1e6 / (double)tgt->n);
  return CVO_OK;
}

extern "C" cvo_status cvo_set_pose_arithmetic(cvo_ctx* ctx, int32_t mode) {
  if (!ctx) return CVO_ERR_INVALID_ARG;
  if (mode != CVO_POSE_FP64 && mode != CVO_POSE_REFERENCE)
    return set_err(ctx, CVO_ERR_INVALID_ARG, "pose arithmetic must be CVO_POSE_FP64 or CVO_POSE_REFERENCE");
  ctx->pose_arith = mode;
  return CVO_OK;
}

extern "C" cvo_status cvo_workspace_bytes(const cvo_ctx* ctx, uint64_t* out) {
  if (!ctx || !out) return CVO_ERR_INVALID_ARG;
  *out = (uint64_t)ctx->ws_bytes;
  return CVO_OK;
}

extern "C" cvo_status cvo_set_workspace_limit(cvo_ctx* ctx, uint64_t bytes) {
  if (!ctx) return CVO_ERR_INVALID_ARG;
  ctx->ws_limit = (size_t)bytes;
  return CVO_OK;
}

extern "C" cvo_status cvo_release_workspace(cvo_ctx* ctx) {
  if (!ctx) return CVO_ERR_INVALID_ARG;
  CVO_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for_each_buf(ctx, free_buf);
  return CVO_OK;
}

// ---------------------------------------------------------------------------
// Inner product without device allocation
// ---------------------------------------------------------------------------

// For the duration of one call every arena buffer of the context is swapped
// for an empty one and ensure() carves from the device's scratch array; the
// arena buffers come back untouched afterwards.  The grid of the call is
// capped at 2^21 cells (larger cells, the same pairs).
struct ScratchScope {
  cvo_ctx* ctx;
  BumpArena arena;
  std::vector<DevBuf*> bufs;
  std::vector<DevBuf> saved;
  uint32_t cells;
  explicit ScratchScope(cvo_ctx* c) : ctx(c) {
    arena.base = ctx->scratch;
    arena.size = kScratchBytes;
    arena.used = 0;
    bufs = all_bufs(ctx);
    saved.reserve(bufs.size());
    for (DevBuf* b : bufs) {
      saved.push_back(*b);
      *b = DevBuf{};
    }
    cells = ctx->grid_max_cells;
    ctx->grid_max_cells = 1u << 21;
    ctx->bump = &arena;
  }
  ~ScratchScope() {
    for (size_t i = 0; i < bufs.size(); ++i) *bufs[i] = saved[i];
    ctx->grid_max_cells = cells;
    ctx->bump = nullptr;
  }
};

extern "C" cvo_status cvo_inner_product_d_noalloc(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt,
                                                  const float T[16], float ell, double* ip) {
  if (!ctx || !T || !ip) return CVO_ERR_INVALID_ARG;
  *ip = 0.0;
  if (!src || !tgt || src->n == 0 || tgt->n == 0) return set_err(ctx, CVO_ERR_EMPTY, "empty point cloud");
  if (!(ell > 0.0f)) return set_err(ctx, CVO_ERR_INVALID_ARG, "ell must be positive");
  if (src->n > kNoAllocMaxPoints || tgt->n > kNoAllocMaxPoints)
    return set_err(ctx, CVO_ERR_LIMIT, "cvo_inner_product_d_noalloc takes at most 131072 points per cloud");
  if (!ctx->scratch) return set_err(ctx, CVO_ERR_HIP, "no scratch address on this device");
  CVO_TRY(validate_pair(ctx, src, tgt));
  ScratchScope scope(ctx);
  const cvo_params& p = ctx->params;
  KConsts kc = make_kconsts(p, src->F, src->K, MODE_IP);
  apply_tuning(ctx, kc);
  kc.n_src_f = (float)src->n;
  kc.n_tgt_f = (float)tgt->n;
  kc.max_src_norm = src->max_norm;
  kc.max_iter = 1;
  const int K0 = p.nearest_neighbors_max;
  CVO_TRY(prepare_source(ctx, src));
  const float tq = std::max(std::fabs(T[3]), std::max(std::fabs(T[7]), std::fabs(T[11])));
  GridHost gh;
  // one level: cells at least the ball radius
  CVO_TRY(build_grid(ctx, tgt, level_sizes(kc, ell, ell, src->max_norm, tq, 1.0f), gh));
  Problems pr;
  CVO_TRY(setup_problems(ctx, src, tgt, gh, 1, 0, 0, kc.max_group, pr));
  CVO_TRY(ensure(ctx, ctx->b_tmp2, 64 * sizeof(float)));
  hipStream_t st = ctx->stream;
  CVO_HIP(ctx, hipMemcpyAsync(ctx->b_tmp2.p, T, sizeof(float) * 16, hipMemcpyHostToDevice, st));
  launch_init_state(pr.dev, 1, as<float>(ctx->b_tmp2), 16, kc, ell, K0, st);
  launch_assoc(pr.dev, 1, pr.total_blocks, kc, st, pr.host.data());
  launch_overflow(pr.dev, 1, kc, st, pr.host.data());
  CVO_HIP(ctx, hipGetLastError());
  ProblemState hs;
  CVO_HIP(ctx, hipMemcpyAsync(&hs, pr.states, sizeof(ProblemState), hipMemcpyDeviceToHost, st));
  CVO_HIP(ctx, hipStreamSynchronize(st));
  *ip = hs.ip;
  return CVO_OK;
}
