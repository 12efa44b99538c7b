// Association-kernel autotuner and its JSON cache.
//
// The knobs (lanes per row and grid cell scale per lengthscale range,
// XCD-contiguous block ranges) change the launch shape and the candidate
// enumeration of the ell-ball search, never the pair set: every setting is
// timed on the caller's own clouds at four lengthscales and the policy with
// the least modelled time per registration is kept on the context.

#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "cvo_internal.h"

using namespace cvo;

namespace {

template <typename T>
T* as(DevBuf& b) {
  return reinterpret_cast<T*>(b.p);
}

bool valid_lanes(int g) { return g == 4 || g == 8 || g == 16; }

bool valid_tuning(const cvo_tuning& t) {
  auto scale_ok = [](float s) { return std::isfinite(s) && s >= 0.4f && s <= 1.0f; };
  return valid_lanes(t.lanes_per_row) && valid_lanes(t.lanes_small_ell) && std::isfinite(t.split_ell) &&
         t.split_ell >= 0.0f && (t.xcd_remap == 0 || t.xcd_remap == 1) && scale_ok(t.cell_scale) &&
         scale_ok(t.cell_scale_small_ell);
}

// ---- cache file: one entry per line inside "entries" -----------------------
//
// {
// "version": 1,
// "entries": {
// "<key>": {"lanes_per_row": 16, "lanes_small_ell": 8, "split_ell": 0.11, "xcd_remap": 1,
//           "cell_scale": 0.7, "cell_scale_small_ell": 1.0, "score_us": 21.5, "n_tried": 48},
//                                                                     (one line per entry)
// ...
// }
// }

bool read_file(const char* path, std::string& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

// Entry lines of a cache file: key -> object text "{...}".
std::map<std::string, std::string> parse_entries(const std::string& text) {
  std::map<std::string, std::string> m;
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    size_t a = line.find_first_not_of(" \t");
    if (a == std::string::npos || line[a] != '"') continue;
    size_t b = line.find('"', a + 1);
    if (b == std::string::npos) continue;
    size_t colon = line.find(':', b + 1);
    size_t lb = line.find('{', b + 1);
    if (colon == std::string::npos || lb == std::string::npos || lb < colon) continue;
    size_t rb = line.rfind('}');
    if (rb == std::string::npos || rb < lb) continue;
    m[line.substr(a + 1, b - a - 1)] = line.substr(lb, rb - lb + 1);
  }
  return m;
}

bool field_num(const std::string& obj, const char* name, double* v) {
  const std::string pat = std::string("\"") + name + "\"";
  size_t p = obj.find(pat);
  if (p == std::string::npos) return false;
  p = obj.find(':', p + pat.size());
  if (p == std::string::npos) return false;
  const char* s = obj.c_str() + p + 1;
  char* e = nullptr;
  errno = 0;
  double x = std::strtod(s, &e);
  if (e == s || errno != 0 || !std::isfinite(x)) return false;
  *v = x;
  return true;
}

std::string format_entry(const cvo_tuning& t) {
  char buf[256];
  std::snprintf(buf, sizeof(buf),
                "{\"lanes_per_row\": %d, \"lanes_small_ell\": %d, \"split_ell\": %.6g, \"xcd_remap\": %d, "
                "\"cell_scale\": %.4f, \"cell_scale_small_ell\": %.4f, \"score_us\": %.3f, \"n_tried\": %d}",
                t.lanes_per_row, t.lanes_small_ell, (double)t.split_ell, t.xcd_remap, (double)t.cell_scale,
                (double)t.cell_scale_small_ell, (double)t.score_us, t.n_tried);
  return buf;
}

bool mkdirs_for(const std::string& path) {
  size_t slash = path.rfind('/');
  if (slash == std::string::npos || slash == 0) return true;
  std::string dir = path.substr(0, slash);
  for (size_t i = 1; i <= dir.size(); ++i) {
    if (i == dir.size() || dir[i] == '/') {
      std::string part = dir.substr(0, i);
      if (::mkdir(part.c_str(), 0755) != 0 && errno != EEXIST) return false;
    }
  }
  return true;
}

bool valid_key(const char* key) {
  if (!key || !key[0]) return false;
  for (const char* c = key; *c; ++c)
    if (*c == '"' || *c == '\\' || *c == '\n' || *c == '{' || *c == '}') return false;
  return true;
}

// Half-octave size class of a point count (log2 rounded to 0.5).
double size_class(int n) { return std::round(2.0 * std::log2((double)std::max(n, 1))) / 2.0; }

}  // namespace

extern "C" cvo_status cvo_tuning_cache_lookup(const char* path, const char* key, cvo_tuning* out) {
  if (!path || !out || !valid_key(key)) return CVO_ERR_INVALID_ARG;
  std::string text;
  if (!read_file(path, text)) return CVO_ERR_IO;
  auto m = parse_entries(text);
  auto it = m.find(key);
  if (it == m.end()) return CVO_ERR_IO;
  double lanes = 0, small = 0, split = 0, xcd = 0, scale = 0, scale_small = 0, score = 0, tried = 0;
  if (!field_num(it->second, "lanes_per_row", &lanes) || !field_num(it->second, "xcd_remap", &xcd) ||
      !field_num(it->second, "cell_scale", &scale))
    return CVO_ERR_PARSE;
  if (!field_num(it->second, "lanes_small_ell", &small)) small = lanes;
  field_num(it->second, "split_ell", &split);
  if (!field_num(it->second, "cell_scale_small_ell", &scale_small)) scale_small = scale;
  field_num(it->second, "score_us", &score);
  field_num(it->second, "n_tried", &tried);
  cvo_tuning t{};
  t.lanes_per_row = (int32_t)lanes;
  t.lanes_small_ell = (int32_t)small;
  t.split_ell = (float)split;
  t.xcd_remap = (int32_t)xcd;
  t.cell_scale = (float)scale;
  t.cell_scale_small_ell = (float)scale_small;
  t.source = CVO_TUNE_CACHED;
  t.score_us = (float)score;
  t.n_tried = (int32_t)tried;
  if (!valid_tuning(t)) return CVO_ERR_PARSE;
  *out = t;
  return CVO_OK;
}

extern "C" cvo_status cvo_tuning_cache_store(const char* path, const char* key, const cvo_tuning* t) {
  if (!path || !t || !valid_key(key) || !valid_tuning(*t)) return CVO_ERR_INVALID_ARG;
  std::string text;
  std::map<std::string, std::string> m;
  if (read_file(path, text)) m = parse_entries(text);
  m[key] = format_entry(*t);
  if (!mkdirs_for(path)) return CVO_ERR_IO;
  std::string tmp = std::string(path) + ".tmp" + std::to_string((long)::getpid());
  {
    std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
    if (!f) return CVO_ERR_IO;
    f << "{\n\"version\": 1,\n\"entries\": {\n";
    size_t i = 0;
    for (const auto& kv : m) {
      f << "\"" << kv.first << "\": " << kv.second << (++i < m.size() ? ",\n" : "\n");
    }
    f << "}\n}\n";
    if (!f) return CVO_ERR_IO;
  }
  if (std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return CVO_ERR_IO;
  }
  return CVO_OK;
}

extern "C" cvo_status cvo_get_tuning(const cvo_ctx* ctx, cvo_tuning* out) {
  if (!ctx || !out) return CVO_ERR_INVALID_ARG;
  *out = ctx->tuning;
  return CVO_OK;
}

extern "C" cvo_status cvo_set_tuning(cvo_ctx* ctx, const cvo_tuning* t) {
  if (!ctx || !t) return CVO_ERR_INVALID_ARG;
  if (!valid_tuning(*t))
    return set_err(ctx, CVO_ERR_INVALID_ARG,
                   "tuning: lanes 8|16, split_ell >= 0, xcd_remap 0|1, cell scales in [0.4, 1]");
  ctx->tuning = *t;
  ctx->tuning.source = CVO_TUNE_SET;
  return CVO_OK;
}

extern "C" cvo_status cvo_tuning_key(const cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, char* buf,
                                     int32_t len) {
  if (!ctx || !src || !tgt || !buf || len <= 0) return CVO_ERR_INVALID_ARG;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) return CVO_ERR_HIP;
  std::string arch = prop.gcnArchName;
  size_t colon = arch.find(':');
  if (colon != std::string::npos) arch = arch.substr(0, colon);
  const cvo_params& p = ctx->params;
  char tmp[512];
  const int n = std::snprintf(
      tmp, sizeof(tmp), "%s/%dcu|src%.1f|tgt%.1f|F%d|K%d|geo%d|col%d|sem%d|gt%d|ell%.4g-%.4g|sp%.3g|nn%d|knn%d",
      arch.c_str(), prop.multiProcessorCount, size_class(src->n), size_class(tgt->n), src->F, src->K,
      p.is_using_geometry, p.is_using_intensity, p.is_using_semantics, src->has_geo && tgt->has_geo ? 1 : 0,
      (double)std::max(p.ell_init, p.ell_min), (double)std::min(p.ell_init, p.ell_min), (double)p.sp_thres,
      p.nearest_neighbors_max, p.is_using_kdtree);
  if (n < 0 || n >= len) return CVO_ERR_LIMIT;
  std::memcpy(buf, tmp, (size_t)n + 1);
  return CVO_OK;
}

extern "C" cvo_status cvo_autotune_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt,
                                     const float T_init[16], const char* cache_path, cvo_tuning* out) {
  if (!ctx || !src || !tgt || !T_init) return CVO_ERR_INVALID_ARG;
  if (src->n == 0 || tgt->n == 0) return set_err(ctx, CVO_ERR_EMPTY, "empty point cloud");
  CVO_TRY(validate_pair(ctx, src, tgt));
  char key[512];
  CVO_TRY(cvo_tuning_key(ctx, src, tgt, key, (int32_t)sizeof(key)));
  if (cache_path) {
    cvo_tuning c;
    if (cvo_tuning_cache_lookup(cache_path, key, &c) == CVO_OK) {
      ctx->tuning = c;
      if (out) *out = c;
      return CVO_OK;
    }
  }
  const cvo_params& p = ctx->params;
  hipStream_t st = ctx->stream;
  KConsts kc = make_kconsts(p, src->F, src->K, MODE_ALIGN);
  kc.n_src_f = (float)src->n;
  kc.n_tgt_f = (float)tgt->n;
  kc.max_src_norm = src->max_norm;
  CVO_TRY(prepare_source(ctx, src));
  const float tq = std::max(std::fabs(T_init[3]), std::max(std::fabs(T_init[7]), std::fabs(T_init[11])));
  const float e_hi = std::max(p.ell_init, p.ell_min), e_lo = std::min(p.ell_init, p.ell_min);
  // lengthscale samples (descending) and the iterations each stands for: the
  // decay schedule spends about two indicator windows per decay before
  // ell_min, which keeps the rest of MAX_ITER
  std::vector<float> ells;
  std::vector<double> w;
  const int n_samples = e_lo < e_hi ? 4 : 1;
  for (int i = 0; i < n_samples; ++i)
    ells.push_back(n_samples == 1 ? e_hi : (float)(e_hi * std::pow((double)e_lo / e_hi, (double)i / (n_samples - 1))));
  if (n_samples == 1) {
    w.push_back(1.0);
  } else {
    const double rate = (p.ell_decay_rate > 0.0f && p.ell_decay_rate < 1.0f) ? p.ell_decay_rate : 0.9;
    const double n_dec = std::ceil(std::log((double)e_lo / e_hi) / std::log(rate));
    const double anneal = n_dec * 2.0 * std::max(1, p.indicator_window_size) + std::max(0, p.ell_decay_start);
    const double rest = std::max((double)p.MAX_ITER - anneal, 2.0 * std::max(1, p.indicator_window_size));
    for (int i = 0; i < n_samples - 1; ++i) w.push_back(anneal / (n_samples - 1));
    w.push_back(rest);
  }
  CVO_TRY(ensure(ctx, ctx->b_tmp2, 64 * sizeof(float)));
  CVO_HIP(ctx, hipMemcpyAsync(ctx->b_tmp2.p, T_init, sizeof(float) * 16, hipMemcpyHostToDevice, st));
  CVO_TRY(ensure_events(ctx, 2));
  hipEvent_t e0 = ctx->ev_pool[0], e1 = ctx->ev_pool[1];
  const char* vb = std::getenv("CVO_TUNE_VERBOSE");
  const bool verbose = vb && vb[0] && vb[0] != '0';
  // settings (lanes, cell scale) timed at every sample, per block order
  constexpr int kNS = 9;
  const int set_lanes[kNS] = {16, 8, 4, 16, 8, 4, 16, 8, 4};
  const float set_scale[kNS] = {1.0f, 1.0f, 1.0f, 0.7f, 0.7f, 0.7f, 0.5f, 0.5f, 0.5f};
  constexpr int kReps = 5, kRounds = 3;
  // one grid serves every scale: its levels reach the finest scale's radius
  GridHost gh;
  CVO_TRY(build_grid(ctx, tgt, level_sizes(kc, e_hi, e_lo, src->max_norm, tq, 0.5f), gh));
  Problems pr;
  CVO_TRY(setup_problems(ctx, src, tgt, gh, 1, p.nearest_neighbors_max, 0, 16, pr));
  double wsum = 0.0;
  for (double x : w) wsum += x;
  // the best policy per block order
  cvo_tuning best_x[2] = {};
  double score_x[2] = {1e300, 1e300};
  int n_tried = 0;
  for (int xcd = 0; xcd < 2; ++xcd) {
    cvo_tuning& best = best_x[xcd];
    double& best_score = score_x[xcd];
    // t[i][j]: microseconds per association pass at sample i, setting j
    double t[4][kNS];
    for (int j = 0; j < kNS; ++j) {
      cvo_tuning one{set_lanes[j], set_lanes[j], 0.0f, xcd, set_scale[j], set_scale[j], 0, 0.0f, 0};
      KConsts kv = kc;
      apply_tuning(one, kv);
      kv.max_group = 16;  // the problems' block ranges
      for (int i = 0; i < n_samples; ++i) {
        launch_init_state(pr.dev, 1, as<float>(ctx->b_tmp2), 16, kv, ells[i], p.nearest_neighbors_max, st);
        launch_assoc(pr.dev, 1, pr.total_blocks, kv, st, pr.host.data());  // warm-up
        CVO_HIP(ctx, hipGetLastError());
        float best_ms = 1e30f;
        for (int rd = 0; rd < kRounds; ++rd) {
          CVO_HIP(ctx, hipEventRecord(e0, st));
          for (int r = 0; r < kReps; ++r) launch_assoc(pr.dev, 1, pr.total_blocks, kv, st, pr.host.data());
          CVO_HIP(ctx, hipEventRecord(e1, st));
          CVO_HIP(ctx, hipEventSynchronize(e1));
          float ms = 0.0f;
          CVO_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
          best_ms = std::min(best_ms, ms / kReps);
        }
        t[i][j] = 1e3 * best_ms;
        ++n_tried;
      }
    }
    // two-segment policy: samples [0, k) take one setting, [k, n) another;
    // k = 0 or n is a single setting
    for (int k = 0; k <= n_samples; ++k) {
      double seg[2] = {0.0, 0.0};
      int sel[2] = {0, 0};
      for (int part = 0; part < 2; ++part) {
        const int i0 = part == 0 ? 0 : k, i1 = part == 0 ? k : n_samples;
        if (i0 == i1) continue;
        seg[part] = 1e300;
        for (int j = 0; j < kNS; ++j) {
          double v = 0.0;
          for (int i = i0; i < i1; ++i) v += w[i] * t[i][j];
          if (v < seg[part]) {
            seg[part] = v;
            sel[part] = j;
          }
        }
      }
      const double score = (seg[0] + seg[1]) / wsum;
      if (score < best_score) {
        best_score = score;
        best.xcd_remap = xcd;
        const bool single = k == 0 || k == n_samples || sel[0] == sel[1];
        const int jb = (k == 0) ? sel[1] : sel[0], js = single ? jb : sel[1];
        best.lanes_per_row = set_lanes[jb];
        best.cell_scale = set_scale[jb];
        best.lanes_small_ell = set_lanes[js];
        best.cell_scale_small_ell = set_scale[js];
        best.split_ell = single ? 0.0f : std::sqrt(ells[k - 1] * ells[k]);
      }
    }
    if (verbose) {
      for (int i = 0; i < n_samples; ++i) {
        std::string line;
        for (int j = 0; j < kNS; ++j) {
          char b[64];
          std::snprintf(b, sizeof(b), " %d/%.1f %.2f", set_lanes[j], (double)set_scale[j], t[i][j]);
          line += b;
        }
        std::fprintf(stderr, "[cvo_autotune] xcd %d ell %.4g (weight %.0f) lanes/scale us:%s\n", xcd,
                     (double)ells[i], w[i], line.c_str());
      }
    }
  }
  // XCD-contiguous row blocks unless dispatch order is clearly faster: the
  // passes above run back to back on warm caches, while in a registration
  // the other kernels of each iteration evict the L2 between passes, where
  // the contiguous order keeps each XCD's targets local (60k street pair:
  // equal pass scores, 124.3 against 121.6 ms per registration)
  const int xc = score_x[0] < 0.98 * score_x[1] ? 0 : 1;
  cvo_tuning best = best_x[xc];
  const double best_score = score_x[xc];
  best.source = CVO_TUNE_MEASURED;
  best.score_us = (float)best_score;
  best.n_tried = n_tried;
  if (verbose)
    std::fprintf(stderr,
                 "[cvo_autotune] chosen: lanes %d scale %.2f, below ell %.4g lanes %d scale %.2f; xcd %d; %.2f us\n",
                 best.lanes_per_row, (double)best.cell_scale, (double)best.split_ell, best.lanes_small_ell,
                 (double)best.cell_scale_small_ell, best.xcd_remap, best_score);
  ctx->tuning = best;
  if (cache_path) {
    cvo_status s = cvo_tuning_cache_store(cache_path, key, &best);
    if (s != CVO_OK) return set_err(ctx, s, std::string("tuning cache not writable: ") + cache_path);
  }
  if (out) *out = best;
  return CVO_OK;
}

extern "C" cvo_status cvo_autotune(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                   const float T_init[16], const char* cache_path, cvo_tuning* out) {
  if (!ctx || !src || !tgt || !T_init) return CVO_ERR_INVALID_ARG;
  cvo_dcloud *a = nullptr, *b = nullptr;
  cvo_status s = cvo_cloud_upload(ctx, src, &a);
  if (s == CVO_OK) s = cvo_cloud_upload(ctx, tgt, &b);
  if (s == CVO_OK) s = cvo_autotune_d(ctx, a, b, T_init, cache_path, out);
  if (a) cvo_cloud_free(ctx, a);
  if (b) cvo_cloud_free(ctx, b);
  return s;
}
