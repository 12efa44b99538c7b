// Parameter defaults and the flat YAML reader of the CvoParams file format.
//
// Defaults follow CvoParams::CvoParams() (include/UnifiedCvo/cvo/CvoParams.hpp:80-136);
// the reader follows read_CvoParams_yaml (CvoParams.hpp:203-323): a key that is
// present overrides the current value, an absent key leaves it alone, unknown keys
// are ignored.  The files are flat "key: value" maps with an OpenCV-style
// "%YAML:1.0" directive and '#' comments; only the first document is read.
// yaml-cpp is not available, so this is a small line parser of that subset.

#include "cvo_c.h"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace {

enum class Kind { F32, F64, I32 };

struct Field {
  const char* key;
  Kind kind;
  size_t offset;
};

#define CVO_F(name, kind) {#name, Kind::kind, offsetof(cvo_params, name)}

// Keys read by read_CvoParams_yaml.  multiframe_iter_per_ell and step are not read
// by the reference (CvoParams.hpp:124 vs :298-299); the same holds here.
const Field kFields[] = {
    CVO_F(ell_init_first_frame, F32),
    CVO_F(ell_init, F32),
    CVO_F(ell_min, F32),
    CVO_F(min_ell_iter_limit, I32),
    CVO_F(ell_max, F32),
    CVO_F(dl, F64),
    CVO_F(dl_step, F64),
    CVO_F(sigma, F32),
    CVO_F(sp_thres, F32),
    CVO_F(c, F32),
    CVO_F(d, F32),
    CVO_F(c_ell, F32),
    CVO_F(c_sigma, F32),
    CVO_F(s_ell, F32),
    CVO_F(s_sigma, F32),
    CVO_F(MAX_ITER, I32),
    CVO_F(eps, F32),
    CVO_F(eps_2, F32),
    CVO_F(min_step, F32),
    CVO_F(max_step, F32),
    CVO_F(ell_decay_rate, F32),
    CVO_F(ell_decay_rate_first_frame, F32),
    CVO_F(ell_decay_start, I32),
    CVO_F(ell_decay_start_first_frame, I32),
    CVO_F(indicator_window_size, I32),
    CVO_F(indicator_stable_threshold, F32),
    CVO_F(is_pcl_visualization_on, I32),
    CVO_F(is_using_least_square, I32),
    CVO_F(is_full_ip_matrix, I32),
    CVO_F(is_using_geometry, I32),
    CVO_F(is_using_intensity, I32),
    CVO_F(is_using_semantics, I32),
    CVO_F(is_using_range_ell, I32),
    CVO_F(is_using_kdtree, I32),
    CVO_F(is_using_geometric_type, I32),
    CVO_F(is_global_angle_registration, I32),
    CVO_F(is_exporting_association, I32),
    CVO_F(nearest_neighbors_max, I32),
    CVO_F(multiframe_using_cpu, I32),
    CVO_F(multiframe_ell_init, F32),
    CVO_F(multiframe_max_iters, I32),
    CVO_F(multiframe_ell_min, F32),
    CVO_F(multiframe_ell_max, F32),
    CVO_F(multiframe_ell_decay_rate, F32),
    CVO_F(multiframe_iterations_per_ell, I32),
    CVO_F(multiframe_iterations_per_solve, I32),
    CVO_F(multiframe_downsample_voxel_size, F32),
    CVO_F(multiframe_expected_points, I32),
    CVO_F(multiframe_num_neighbors, I32),
    CVO_F(multiframe_min_nonzeros, I32),
    CVO_F(multiframe_least_squares_num_threads, I32),
    CVO_F(multiframe_is_optimizing_ell, I32),
    CVO_F(multiframe_is_sorting_inner_product, I32),
    CVO_F(multiframe_is_free_memory_each_iter, I32),
};
#undef CVO_F

std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
  while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
  return s.substr(b, e - b);
}

std::string strip_quotes(const std::string& s) {
  if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
    return s.substr(1, s.size() - 2);
  return s;
}

// YAML 1.1 booleans convert to 1/0 for integer fields.
bool parse_bool(const std::string& v, long* out) {
  static const char* t[] = {"true", "True", "TRUE", "yes", "Yes", "YES", "on", "On", "ON"};
  static const char* f[] = {"false", "False", "FALSE", "no", "No", "NO", "off", "Off", "OFF"};
  for (const char* s : t)
    if (v == s) { *out = 1; return true; }
  for (const char* s : f)
    if (v == s) { *out = 0; return true; }
  return false;
}

bool parse_double(const std::string& v, double* out) {
  if (v.empty()) return false;
  errno = 0;
  char* end = nullptr;
  double x = std::strtod(v.c_str(), &end);
  if (end == v.c_str() || *end != '\0' || errno == ERANGE) return false;
  *out = x;
  return true;
}

bool parse_int(const std::string& v, long* out) {
  if (parse_bool(v, out)) return true;
  errno = 0;
  char* end = nullptr;
  long x = std::strtol(v.c_str(), &end, 0);
  if (end != v.c_str() && *end == '\0' && errno != ERANGE) {
    *out = x;
    return true;
  }
  return false;
}

}  // namespace

extern "C" cvo_status cvo_params_default(cvo_params* p) {
  if (!p) return CVO_ERR_INVALID_ARG;
  std::memset(p, 0, sizeof(*p));
  p->ell_init_first_frame = 0.5f;
  p->ell_init = 0.5f;
  p->ell_min = 0.05f;
  p->min_ell_iter_limit = 1;
  p->ell_max = 1.2f;
  p->dl = 0.0;
  p->dl_step = 0.3;
  p->sigma = 0.1f;
  p->sp_thres = 0.0006f;
  p->c = 7.0f;
  p->d = 7.0f;
  p->c_ell = 0.15f;
  p->c_sigma = 0.6f;
  p->s_ell = 0.1f;
  p->s_sigma = 0.8f;
  p->MAX_ITER = 10000;
  p->min_step = 2e-5f;
  p->max_step = 0.01f;  // no default in the reference; documented choice
  p->step = 0.0f;
  p->eps = 0.00005f;
  p->eps_2 = 0.000012f;
  p->ell_decay_rate = 0.9f;
  p->ell_decay_rate_first_frame = 0.99f;
  p->ell_decay_start = 30;
  p->ell_decay_start_first_frame = 300;
  p->indicator_window_size = 15;
  p->indicator_stable_threshold = 0.2f;
  p->is_pcl_visualization_on = 0;
  p->is_using_least_square = 0;
  p->is_ell_adaptive = 0;
  p->is_full_ip_matrix = 0;
  p->is_using_geometry = 1;
  p->is_using_intensity = 0;
  p->is_using_semantics = 0;
  p->is_using_range_ell = 0;
  p->is_using_kdtree = 0;
  p->is_using_geometric_type = 0;
  p->is_global_angle_registration = 0;
  p->is_exporting_association = 0;
  p->multiframe_using_cpu = 1;
  p->multiframe_max_iters = 200;
  p->nearest_neighbors_max = 512;
  p->multiframe_ell_init = 0.15f;
  p->multiframe_ell_min = 0.05f;
  p->multiframe_ell_max = 5.0f;
  p->multiframe_iter_per_ell = 10;
  p->multiframe_ell_decay_rate = 0.7f;
  p->multiframe_iterations_per_ell = 50;
  p->multiframe_iterations_per_solve = 8;
  p->multiframe_downsample_voxel_size = 0.5f;
  p->multiframe_expected_points = 1000;
  p->multiframe_num_neighbors = 128;
  p->multiframe_min_nonzeros = 300;
  p->multiframe_least_squares_num_threads = 24;
  p->multiframe_is_optimizing_ell = 0;
  p->multiframe_is_sorting_inner_product = 0;
  p->multiframe_is_free_memory_each_iter = 0;
  return CVO_OK;
}

extern "C" cvo_status cvo_params_read_yaml(const char* path, cvo_params* p) {
  if (!path || !p) return CVO_ERR_INVALID_ARG;
  FILE* f = std::fopen(path, "r");
  if (!f) return CVO_ERR_IO;
  char buf[4096];
  bool in_doc = false;  // saw the first "---" or a key
  cvo_status st = CVO_OK;
  while (std::fgets(buf, sizeof(buf), f)) {
    std::string line(buf);
    // strip comments (a '#' starting the line or preceded by whitespace)
    for (size_t i = 0; i < line.size(); ++i) {
      if (line[i] == '#' && (i == 0 || std::isspace(static_cast<unsigned char>(line[i - 1])))) {
        line.resize(i);
        break;
      }
    }
    std::string t = trim(line);
    if (t.empty()) continue;
    if (t[0] == '%') continue;  // %YAML:1.0 directive
    if (t.rfind("---", 0) == 0) {
      if (in_doc) break;  // a second document: only the first is read
      in_doc = true;
      continue;
    }
    if (t == "...") break;
    size_t colon = t.find(':');
    if (colon == std::string::npos) continue;
    in_doc = true;
    std::string key = trim(t.substr(0, colon));
    std::string val = strip_quotes(trim(t.substr(colon + 1)));
    for (const Field& fd : kFields) {
      if (key != fd.key) continue;
      char* base = reinterpret_cast<char*>(p) + fd.offset;
      if (fd.kind == Kind::I32) {
        long x;
        double dx;
        if (parse_int(val, &x)) {
          *reinterpret_cast<int32_t*>(base) = static_cast<int32_t>(x);
        } else if (parse_double(val, &dx) && std::floor(dx) == dx) {
          *reinterpret_cast<int32_t*>(base) = static_cast<int32_t>(dx);
        } else {
          st = CVO_ERR_PARSE;
        }
      } else {
        double x;
        long bx;
        if (parse_double(val, &x)) {
          if (fd.kind == Kind::F32) *reinterpret_cast<float*>(base) = static_cast<float>(x);
          else *reinterpret_cast<double*>(base) = x;
        } else if (parse_bool(val, &bx)) {
          if (fd.kind == Kind::F32) *reinterpret_cast<float*>(base) = static_cast<float>(bx);
          else *reinterpret_cast<double*>(base) = static_cast<double>(bx);
        } else {
          st = CVO_ERR_PARSE;
        }
      }
      break;
    }
  }
  std::fclose(f);
  return st;
}
