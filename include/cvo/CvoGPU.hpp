// C++ facade over the C ABI (cvo_c.h) with the class and method names of the
// reference's cvo::CvoGPU (include/UnifiedCvo/cvo/CvoGPU.hpp:43-274), so code
// written against that class compiles against librkhs_ba_amd.so.  Header only;
// errors from the library are thrown as cvo::CvoError (the reference exits the
// process on CUDA errors).  Eigen overloads are added when <Eigen/Dense> is
// available; the float[16] overloads are always there (row-major 4x4).
#pragma once

#include <cstdint>
#include <algorithm>
#include <cstring>
#include <list>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../cvo_c.h"

#if __has_include(<Eigen/Dense>)
#include <Eigen/Dense>
#define CVO_FACADE_HAS_EIGEN 1
#endif

namespace cvo {

class CvoError : public std::runtime_error {
 public:
  CvoError(cvo_status s, const std::string& what) : std::runtime_error(what), status(s) {}
  cvo_status status;
};

inline void check(cvo_status s, const cvo_ctx* ctx = nullptr) {
  if (s == CVO_OK) return;
  std::string msg = cvo_status_string(s);
  if (ctx) {
    const char* e = cvo_last_error(ctx);
    if (e && *e) msg += std::string(": ") + e;
  }
  throw CvoError(s, msg);
}

// Host point cloud in the layout the library reads (structure of arrays):
// the fields of cvo::CvoPointCloud (CvoPointCloud.hpp:38-221) without PCL.
class CvoPointCloud {
 public:
  CvoPointCloud() = default;
  CvoPointCloud(int feature_dimensions, int num_classes)
      : feature_dimensions_(feature_dimensions), num_classes_(num_classes) {}

  void reserve(int num_points, int feature_dims, int num_classes) {
    feature_dimensions_ = feature_dims;
    num_classes_ = num_classes;
    xyz_.assign((size_t)num_points * 3, 0.0f);
    features_.assign((size_t)num_points * feature_dims, 0.0f);
    labels_.assign((size_t)num_points * num_classes, 0.0f);
    geometric_types_.assign((size_t)num_points * 2, 0.0f);
    num_points_ = num_points;
  }
  // add_point (src/utils/CvoPointCloud.cpp:1579-1620) into a reserved slot.
  int add_point(int index, const float xyz[3], const float* feature, const float* label,
                const float geometric_type[2]) {
    if (index < 0 || index >= num_points_) return -1;
    std::memcpy(&xyz_[3 * (size_t)index], xyz, 3 * sizeof(float));
    if (feature_dimensions_ && feature)
      std::memcpy(&features_[(size_t)index * feature_dimensions_], feature, feature_dimensions_ * sizeof(float));
    if (num_classes_ && label)
      std::memcpy(&labels_[(size_t)index * num_classes_], label, num_classes_ * sizeof(float));
    if (geometric_type) std::memcpy(&geometric_types_[2 * (size_t)index], geometric_type, 2 * sizeof(float));
    return 0;
  }
  int num_points() const { return num_points_; }
  int size() const { return num_points_; }
  int num_classes() const { return num_classes_; }
  int num_features() const { return feature_dimensions_; }
  int feature_dimensions() const { return feature_dimensions_; }
  const float* xyz_at(int i) const { return &xyz_[3 * (size_t)i]; }
  const float* feature_at(int i) const { return &features_[(size_t)i * feature_dimensions_]; }
  const float* label_at(int i) const { return &labels_[(size_t)i * num_classes_]; }
  bool has_geometric_types = false;

  // Borrowed view for one library call.
  cvo_cloud_view view() const {
    cvo_cloud_view v;
    v.n = num_points_;
    v.feat_dim = feature_dimensions_;
    v.num_classes = num_classes_;
    v.xyz = xyz_.data();
    v.feat = feature_dimensions_ ? features_.data() : nullptr;
    v.sem = num_classes_ ? labels_.data() : nullptr;
    v.geo_type = has_geometric_types ? geometric_types_.data() : nullptr;
    return v;
  }

 private:
  int num_points_ = 0, feature_dimensions_ = 0, num_classes_ = 0;
  std::vector<float> xyz_, features_, labels_, geometric_types_;
};

// Association (Association.hpp:7-14) as CSR: row_ptr over source rows,
// target ids ascending per row.
struct Association {
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> col;
  std::vector<float> val;
  std::vector<int> source_inliers() const {
    std::vector<int> out;
    for (size_t i = 0; i + 1 < row_ptr.size(); ++i)
      if (row_ptr[i + 1] > row_ptr[i]) out.push_back((int)i);
    return out;
  }
};

// CvoResultInfo (CvoGPU.hpp:34-40).
struct CvoResultInfo {
  float T_s2t[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  int num_iters = 0;
  double registration_seconds = 0.0;
  int return_code = 0;
  Association association;
};

// CvoFrame (CvoFrame.hpp:12-39): a sensor-frame cloud and its pose
// (row-major 3x4, world <- sensor), updated in place by the BA align.
struct CvoFrame {
  const CvoPointCloud* points = nullptr;
  double pose_vec[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  CvoFrame(const CvoPointCloud* pts, const double pose[12]) : points(pts) {
    if (pose) std::memcpy(pose_vec, pose, sizeof(pose_vec));
  }
  typedef std::shared_ptr<CvoFrame> Ptr;
};

class CvoGPU {
 public:
  explicit CvoGPU(const std::string& param_file) { check(cvo_create(param_file.c_str(), &ctx_)); }
  explicit CvoGPU(const cvo_params& p) { check(cvo_create_from_params(&p, &ctx_)); }
  ~CvoGPU() { cvo_destroy(ctx_); }
  CvoGPU(const CvoGPU&) = delete;
  CvoGPU& operator=(const CvoGPU&) = delete;

  cvo_params& get_params() {
    check(cvo_get_params(ctx_, &params_), ctx_);
    return params_;
  }
  void write_params(const cvo_params* p) { check(cvo_set_params(ctx_, p), ctx_); }

  // int align(source, target, T_init, T_out, association, seconds, iters)
  // (CvoGPU.cu:1516-1563).  Returns the reference's code (0, or -1 when the
  // gradient vanished); empty input returns 0 and T_out = identity.
  int align(const CvoPointCloud& source, const CvoPointCloud& target, const float T_init[16], float T_out[16],
            Association* association = nullptr, double* registration_seconds = nullptr,
            int* num_iters = nullptr) const {
    const cvo_cloud_view s = source.view(), t = target.view();
    cvo_assoc a{};
    double secs = 0.0;
    int32_t it = 0, ret = 0;
    cvo_status st = cvo_align(ctx_, &s, &t, T_init, T_out, association ? &a : nullptr, &secs, &it, &ret);
    if (st == CVO_ERR_EMPTY) return 0;
    check(st, ctx_);
    if (registration_seconds) *registration_seconds = secs;
    if (num_iters) *num_iters = it;
    if (association && a.row_ptr) take(a, *association);
    return ret;
  }

  // CvoResultInfo align(source, target, T_init_s2t, is_returning_association)
  // (CvoGPU.cu:1566-1611): starts from the inverse of its init, or with
  // is_global_angle_registration from get_nearest_init_pose's result itself;
  // return_code -1 on empty input.
  CvoResultInfo align(const CvoPointCloud& source, const CvoPointCloud& target, const float T_init_s2t[16],
                      bool is_returning_association = false) const {
    CvoResultInfo r;
    const cvo_cloud_view s = source.view(), t = target.view();
    cvo_assoc a{};
    int32_t it = 0, ret = 0;
    cvo_status st = cvo_align_s2t(ctx_, &s, &t, T_init_s2t, r.T_s2t, is_returning_association ? &a : nullptr,
                                  &r.registration_seconds, &it, &ret);
    if (st == CVO_ERR_EMPTY) {
      r.return_code = -1;
      return r;
    }
    check(st, ctx_);
    r.num_iters = it;
    r.return_code = ret;
    if (is_returning_association && a.row_ptr) take(a, r.association);
    return r;
  }

  // Matrix4f get_nearest_init_pose(source, target, ell, num_guess, times)
  // (CvoGPU.hpp:249-253, CvoGPU.cpp:367-398): the rotation of the 60-rotation
  // table with the largest approximate cos angle, as a row-major 4x4
  // source-to-target transform; *times is incremented by the search time.
  void get_nearest_init_pose(const CvoPointCloud& source, const CvoPointCloud& target, float init_guess_ell,
                             int num_guess, float T_out[16], double* times = nullptr) const {
    const cvo_cloud_view s = source.view(), t = target.view();
    check(cvo_get_nearest_init_pose(ctx_, &s, &t, init_guess_ell, num_guess, T_out, times), ctx_);
  }

  // int align(frames, hold_const, edges, seconds[, inliers, num_iters]): the
  // IRLS bundle adjustment (CvoGPU.cu:1614-1671, :1955-1999); the edges are
  // frame pairs (the BinaryState list of the reference).  Poses of the
  // frames are updated in place; inliers (optional) receives each edge's
  // association of the last association pass, in edge order.
  int align(std::vector<CvoFrame::Ptr>& frames, const std::vector<bool>& hold_const,
            const std::list<std::pair<CvoFrame::Ptr, CvoFrame::Ptr>>& edges, double* seconds,
            std::list<std::shared_ptr<Association>>* inliers = nullptr, int* num_iters = nullptr) const {
    const int n = (int)frames.size();
    std::vector<cvo_dcloud*> dev(n, nullptr);
    std::vector<double> poses((size_t)n * 12);
    std::vector<uint8_t> hc(n, 0);
    std::vector<int32_t> e;
    cvo_status st = CVO_OK;
    for (int f = 0; f < n && st == CVO_OK; ++f) {
      const cvo_cloud_view v = frames[f]->points->view();
      st = cvo_cloud_upload(ctx_, &v, &dev[f]);
      std::memcpy(&poses[12 * (size_t)f], frames[f]->pose_vec, 12 * sizeof(double));
      hc[f] = f < (int)hold_const.size() && hold_const[f] ? 1 : 0;
    }
    for (const auto& pr : edges) {
      int a = -1, b = -1;
      for (int f = 0; f < n; ++f) {
        if (frames[f] == pr.first) a = f;
        if (frames[f] == pr.second) b = f;
      }
      e.push_back(a);
      e.push_back(b);
    }
    int32_t iters = 0;
    const int32_t n_edges = (int32_t)(e.size() / 2);
    std::vector<cvo_assoc> per_edge(inliers ? std::max(n_edges, 1) : 0);
    if (st == CVO_OK)
      st = cvo_align_multiframe(ctx_, n, dev.data(), poses.data(), hc.data(), n_edges, e.data(),
                                inliers ? per_edge.data() : nullptr, seconds, &iters, nullptr, 0);
    for (cvo_dcloud* d : dev)
      if (d) cvo_cloud_free(ctx_, d);
    check(st, ctx_);
    for (int f = 0; f < n; ++f) std::memcpy(frames[f]->pose_vec, &poses[12 * (size_t)f], 12 * sizeof(double));
    if (num_iters) *num_iters = iters;
    if (inliers)
      for (int32_t k = 0; k < n_edges; ++k) {
        auto as = std::make_shared<Association>();
        if (per_edge[k].row_ptr) take(per_edge[k], *as);
        inliers->push_back(as);
      }
    return 0;
  }

  // float function_angle(source, target, T, ell, is_approximate, is_gpu)
  // (CvoGPU.cu:1768-1806); is_gpu false takes the inner_product_cpu
  // semantics (computed on the GPU as well).
  float function_angle(const CvoPointCloud& source, const CvoPointCloud& target, const float t2s[16], float ell,
                       bool is_approximate = true, bool is_gpu = true) const {
    const cvo_cloud_view s = source.view(), t = target.view();
    double c = 0.0;
    check(cvo_function_angle_ex(ctx_, &s, &t, t2s, ell, is_approximate ? 1 : 0, is_gpu ? 1 : 0, &c), ctx_);
    return (float)c;
  }

  // float inner_product_cpu(source, target, T, ell) (CvoGPU.cpp:204-231):
  // fixed lengthscale, no row cap, no gates but the radius.
  float inner_product_cpu(const CvoPointCloud& source, const CvoPointCloud& target, const float T[16],
                          float ell) const {
    const cvo_cloud_view s = source.view(), t = target.view();
    double ip = 0.0;
    check(cvo_inner_product_cpu(ctx_, &s, &t, T, ell, &ip), ctx_);
    return (float)ip;
  }

  // float inner_product_gpu(source, target, T, ell) (CvoGPU.cu:1734-1748).
  float inner_product_gpu(const CvoPointCloud& source, const CvoPointCloud& target, const float T[16],
                          float ell) const {
    const cvo_cloud_view s = source.view(), t = target.view();
    double ip = 0.0;
    cvo_status st = cvo_inner_product(ctx_, &s, &t, T, ell, &ip);
    if (st == CVO_ERR_EMPTY) return 0.0f;
    check(st, ctx_);
    return (float)ip;
  }

  // void compute_association_gpu(source, target, T, ell, association)
  // (CvoGPU.cu:1836-1871).
  void compute_association_gpu(const CvoPointCloud& source, const CvoPointCloud& target, const float T[16],
                               float ell, Association& association) const {
    const cvo_cloud_view s = source.view(), t = target.view();
    cvo_assoc a{};
    check(cvo_compute_association(ctx_, &s, &t, T, ell, &a), ctx_);
    take(a, association);
  }

  // void compute_association_gpu(source, target, T, const Eigen::Matrix3f&
  // non_isotropic_kernel, association) (CvoGPU.cu:1873-1952); kernel row-major.
  void compute_association_gpu(const CvoPointCloud& source, const CvoPointCloud& target, const float T[16],
                               const float non_isotropic_kernel[9], Association& association) const {
    const cvo_cloud_view s = source.view(), t = target.view();
    cvo_assoc a{};
    check(cvo_compute_association_nonisotropic(ctx_, &s, &t, T, non_isotropic_kernel, &a), ctx_);
    take(a, association);
  }

#ifdef CVO_FACADE_HAS_EIGEN
  int align(const CvoPointCloud& source, const CvoPointCloud& target, const Eigen::Matrix4f& T_init,
            Eigen::Ref<Eigen::Matrix4f> T_out, Association* association = nullptr,
            double* registration_seconds = nullptr, int* num_iters = nullptr) const {
    Eigen::Matrix<float, 4, 4, Eigen::RowMajor> a = T_init, b;
    const int r = align(source, target, a.data(), b.data(), association, registration_seconds, num_iters);
    T_out = b;
    return r;
  }
  Eigen::Matrix4f get_nearest_init_pose(const CvoPointCloud& source, const CvoPointCloud& target,
                                        float init_guess_ell, int num_guess, double* times) const {
    Eigen::Matrix<float, 4, 4, Eigen::RowMajor> m;
    get_nearest_init_pose(source, target, init_guess_ell, num_guess, m.data(), times);
    return m;
  }
#endif

  cvo_ctx* handle() const { return ctx_; }

 private:
  static void take(cvo_assoc& a, Association& out) {
    out.row_ptr.assign(a.row_ptr, a.row_ptr + a.rows + 1);
    out.col.assign(a.col, a.col + a.nnz);
    out.val.assign(a.val, a.val + a.nnz);
    cvo_assoc_free(&a);
  }
  cvo_ctx* ctx_ = nullptr;
  cvo_params params_{};
};

}  // namespace cvo
