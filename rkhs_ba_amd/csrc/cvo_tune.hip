// Association-kernel autotuner and its JSON cache.
//
// The knobs (lanes per row, XCD-contiguous block ranges, grid cell scale)
// change the launch shape and the candidate enumeration of the ell-ball
// search, never the pair set: every setting is timed on the caller's own
// clouds at three lengthscales and the fastest is kept on the context.

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

bool valid_tuning(const cvo_tuning& t) {
  return (t.lanes_per_row == 8 || t.lanes_per_row == 16) && (t.xcd_remap == 0 || t.xcd_remap == 1) &&
         std::isfinite(t.cell_scale) && t.cell_scale >= 0.4f && t.cell_scale <= 1.0f;
}

// ---- cache file: one entry per line inside "entries" -----------------------
//
// {
// "version": 1,
// "entries": {
// "<key>": {"lanes_per_row": 8, "xcd_remap": 1, "cell_scale": 0.7, "score_us": 21.5, "n_tried": 12},
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
                "{\"lanes_per_row\": %d, \"xcd_remap\": %d, \"cell_scale\": %.4f, \"score_us\": %.3f, \"n_tried\": %d}",
                t.lanes_per_row, t.xcd_remap, (double)t.cell_scale, (double)t.score_us, t.n_tried);
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
  double lanes = 0, xcd = 0, scale = 0, score = 0, tried = 0;
  if (!field_num(it->second, "lanes_per_row", &lanes) || !field_num(it->second, "xcd_remap", &xcd) ||
      !field_num(it->second, "cell_scale", &scale))
    return CVO_ERR_PARSE;
  field_num(it->second, "score_us", &score);
  field_num(it->second, "n_tried", &tried);
  cvo_tuning t{};
  t.lanes_per_row = (int32_t)lanes;
  t.xcd_remap = (int32_t)xcd;
  t.cell_scale = (float)scale;
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
    return set_err(ctx, CVO_ERR_INVALID_ARG, "tuning: lanes_per_row 8|16, xcd_remap 0|1, cell_scale in [0.4, 1]");
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
      tmp, sizeof(tmp), "%s/%dcu|src%.1f|tgt%.1f|F%d|K%d|geo%d|col%d|sem%d|gt%d|ell%.4g-%.4g|sp%.3g|nn%d",
      arch.c_str(), prop.multiProcessorCount, size_class(src->n), size_class(tgt->n), src->F, src->K,
      p.is_using_geometry, p.is_using_intensity, p.is_using_semantics, src->has_geo && tgt->has_geo ? 1 : 0,
      (double)std::max(p.ell_init, p.ell_min), (double)std::min(p.ell_init, p.ell_min), (double)p.sp_thres,
      p.nearest_neighbors_max);
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
  std::vector<float> ells = {e_hi};
  std::vector<double> w = {1.0};
  if (e_lo < e_hi) {
    ells.push_back(std::sqrt(e_hi * e_lo));
    w.push_back(1.0);
    ells.push_back(e_lo);
    w.push_back(2.0);
  }
  CVO_TRY(ensure(ctx, ctx->b_tmp2, 64 * sizeof(float)));
  CVO_HIP(ctx, hipMemcpyAsync(ctx->b_tmp2.p, T_init, sizeof(float) * 16, hipMemcpyHostToDevice, st));
  CVO_TRY(ensure_events(ctx, 2));
  hipEvent_t e0 = ctx->ev_pool[0], e1 = ctx->ev_pool[1];
  const float scales[3] = {1.0f, 0.7f, 0.5f};
  const int groups[2] = {16, 8};
  constexpr int kReps = 5, kRounds = 3;
  const char* vb = std::getenv("CVO_TUNE_VERBOSE");
  const bool verbose = vb && vb[0] && vb[0] != '0';
  cvo_tuning best{};
  double best_score = 1e300;
  int n_tried = 0;
  for (float sc : scales) {
    GridHost gh;
    CVO_TRY(build_grid(ctx, tgt, level_sizes(kc, e_hi, e_lo, src->max_norm, tq, sc), gh));
    for (int grp : groups) {
      Problems pr;
      CVO_TRY(setup_problems(ctx, src, tgt, gh, 1, p.nearest_neighbors_max, 0, grp, pr));
      for (int xcd = 0; xcd < 2; ++xcd) {
        KConsts kv = kc;
        kv.group = grp;
        kv.xcd_remap = xcd;
        kv.cell_scale = sc;
        double score = 0.0, wsum = 0.0;
        std::string line;
        for (size_t i = 0; i < ells.size(); ++i) {
          launch_init_state(pr.dev, 1, as<float>(ctx->b_tmp2), 16, kv, ells[i], p.nearest_neighbors_max, st);
          launch_assoc(pr.dev, 1, pr.total_blocks, kv, st);  // warm-up
          CVO_HIP(ctx, hipGetLastError());
          float best_ms = 1e30f;
          for (int rd = 0; rd < kRounds; ++rd) {
            CVO_HIP(ctx, hipEventRecord(e0, st));
            for (int r = 0; r < kReps; ++r) launch_assoc(pr.dev, 1, pr.total_blocks, kv, st);
            CVO_HIP(ctx, hipEventRecord(e1, st));
            CVO_HIP(ctx, hipEventSynchronize(e1));
            float ms = 0.0f;
            CVO_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
            best_ms = std::min(best_ms, ms / kReps);
          }
          score += w[i] * 1e3 * best_ms;
          wsum += w[i];
          char b[64];
          std::snprintf(b, sizeof(b), " ell %.4g: %.2f us", (double)ells[i], 1e3 * best_ms);
          line += b;
        }
        score /= wsum;
        if (verbose)
          std::fprintf(stderr, "[cvo_autotune] lanes %d xcd %d scale %.2f:%s -> %.2f us\n", grp, xcd, (double)sc,
                       line.c_str(), score);
        ++n_tried;
        if (score < best_score) {
          best_score = score;
          best.lanes_per_row = grp;
          best.xcd_remap = xcd;
          best.cell_scale = sc;
        }
      }
    }
  }
  best.source = CVO_TUNE_MEASURED;
  best.score_us = (float)best_score;
  best.n_tried = n_tried;
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
