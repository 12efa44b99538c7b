/*
 * cvo_oracle.h -- CPU restatement of the reference CVO / RKHS-BA hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker of the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library (rkhs_ba_amd/librkhs_ba_amd.so) never links or calls it.
 *
 * Two precisions:
 *   ORC_MIRROR (0): the reference's own rounding points (FP64 exp rounded to
 *                   float, float products, float pose state, float Exp_SEK3 with
 *                   its small-angle branch, float indicator sums).
 *   ORC_FP64   (1): the same algorithm in FP64 throughout, with the exact
 *                   small-angle series of the SE(3) exponential.
 *
 * Every function cites the reference lines it follows (paths relative to the
 * UnifiedCvo tree).  Pins (what checks each function independently) are listed
 * in the header comment of cvo_oracle.cpp.
  *   ORC_POLICY (2): association, flow and step terms at the reference's
 *              rounding points (as ORC_MIRROR) with the pose, the SE(3)
 *              exponential and the step in FP64 -- the GPU library's default
 *              precision policy (align only).
 */
#ifndef CVO_ORACLE_H
#define CVO_ORACLE_H

#include <stdint.h>
#include "../include/cvo_c.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_MIRROR = 0, ORC_FP64 = 1, ORC_POLICY = 2 };

/* Pair counters of one association pass. */
typedef struct {
  int64_t geo;     /* pairs inside the geometric gate                  */
  int64_t color;   /* ... and the color gate                           */
  int64_t sem;     /* ... and the semantic gate                        */
  int64_t passed;  /* a > sp_thres (before the row cap)                */
  int64_t stored;  /* after the row cap                                */
  int32_t max_row; /* max stored pairs in one row                      */
} orc_counters;

/* fill_in_A_mat_gpu (src/cvo/CvoGPU.cu:951-1099) at pose T (target mapped by
 * T^-1, :98-130 and CvoGPU_impl.cu:32-83), lengthscale ell and row cap K.
 * counts[N]; cols/vals [N*K] row-major, ascending target index within a row
 * (the reference's scan order).  brute != 0 scans every target like the
 * reference; brute == 0 uses a uniform grid and gives the same set. */
int orc_assoc(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
              const float T[16], float ell, int32_t K, int32_t precision, int32_t brute,
              int32_t* counts, int32_t* cols, float* vals, orc_counters* cnt);

/* The K-nearest association of CvoGPU::align with is_using_kdtree
 * (se_kernel_kdtree, src/cvo/CvoGPU.cu:1104-1154; fill_in_A_mat_cukdtree
 * :832-949): per row the K nearest targets (ties by index), nearest first,
 * that pass the gates at the fixed lengthscale ell.  counts[N]; cols/vals
 * [N*K] in ascending distance.  brute != 0 scans every target. */
int orc_assoc_knn(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                  const float T[16], float ell, int32_t K, int32_t precision, int32_t brute,
                  int32_t* counts, int32_t* cols, float* vals, orc_counters* cnt);

/* AdaptiveCvoGPU::align (src/cvo/AdaptiveCvoGPU.cu:1077-1248): the
 * lengthscale follows its gradient from the A, Axx, Ayy associations.
 * precision MIRROR (float pose and row sums) or POLICY (the library's FP64
 * pose and sums). */
int orc_align_adaptive(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                       const float T_init[16], int32_t precision, float T_out[16], int32_t* iters,
                       int32_t* ret_code, cvo_iter_record* trace, int32_t trace_cap);
/* Its association (AdaptiveCvoGPU.cu:263-408) at pose T: counts[N], cols /
 * vals [N*K] in ascending target index, ad2[N] = per-row sum of a |y - x|^2. */
int orc_assoc_adaptive(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                       const float T[16], float ell, int32_t K, int32_t precision, int32_t* counts,
                       int32_t* cols, float* vals, double* ad2);
/* AdaptiveCvoGPU::inner_product (AdaptiveCvoGPU.cu:1251-1371): the target
 * moved by T_s2t, every pair passing the adaptive gates at ell_init (no row
 * cap), sum(a) / N * 1e6 / M. */
double orc_inner_product_adaptive(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                  const float T_s2t[16]);
/* Its step rule (AdaptiveCvoGPU.cu:1039-1074) on B, C, D, E. */
float orc_select_step_adaptive(const cvo_params* p, const double bcde[4]);

/* For the near-cutoff exemption of the parity tests: the reference's product
 * kernel value a (before the sp_thres test) of one pair, or -1 when a gate
 * rejects it, plus the relative margin of the closest gate. */
int orc_pair_value(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                   const float T[16], float ell, int32_t i, int32_t j,
                   double* a_out, double* gate_margin);

/* Transformed target y' = Ainv y + binv in float (CvoGPU_impl.cu:32-83). */
int orc_transform_target(const float T[16], int32_t m, const float* xyz, float* out);

/* compute_flow_gpu_no_eigen + compute_flow (CvoGPU.cu:2001-2120): omega, v
 * sums (before normalisation) in out6, and the normalised float 6-vector in
 * unit6.  yprime: transformed target xyz. */
int orc_flow(const cvo_params* p, int32_t n, const float* src_xyz, const float* yprime,
             int32_t K, const int32_t* counts, const int32_t* cols, const float* vals,
             int32_t precision, double out6[6], double unit6[6]);

/* compute_step_size_xi + compute_step_size_poly_coeff (CvoGPU.cu:333-462):
 * B, C, D, E.  flow6 = normalised [omega; v]. */
int orc_step_coeffs(const cvo_params* p, float ell, int32_t n, const float* src_xyz,
                    const float* yprime, int32_t K, const int32_t* counts, const int32_t* cols,
                    const float* vals, const double flow6[6], int32_t precision, double bcde[4]);

/* poly_solver_order3 (LieGroup.cpp:310-326): the three complex roots of
 * c[0] s^3 + c[1] s^2 + c[2] s + c[3]. */
int orc_cubic_roots(const double c[4], double re[3], double im[3]);

/* compute_step_size root selection and clamp (CvoGPU.cu:504-538). */
double orc_select_step(const cvo_params* p, const double bcde[4]);

/* Exp_SEK3 float form (LieGroup.cpp:244-275), output row-major 3x4. */
int orc_exp_sek3_f32(const float xi[6], float dt, float out12[12]);
/* FP64 exp of dt*[omega; v] with the exact small-angle series. */
int orc_exp_se3_f64(const double xi[6], double dt, double out12[12]);
/* ||log(T)|| of a rigid transform (row-major 3x4), the Sophus SE3d log norm
 * used by dist_se3_cpu (src/cvo/CvoGPU.cpp:28-33). */
double orc_log_se3_norm(const double T12[12]);
/* Tangent 6-vector (upsilon; omega) of log(T). */
int orc_log_se3(const double T12[12], double out6[6]);

/* A_sparsity_indicator_ell_update (CvoGPU.cu:547-665) run over a scripted
 * indicator sequence: decisions[k] = 1 when the lengthscale would decrease. */
int orc_indicator_sequence(int32_t window, float threshold, int32_t n, const float* ind,
                           int32_t* decisions);

/* align_impl (CvoGPU.cu:1239-1477) on the int overload (:1516-1563):
 * T_init is the pose (R,T); T_out = [R^T, -R^T T].  trace (optional, cap
 * records): per-iteration ell, step, nnz, ip. */
int orc_align(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
              const float T_init[16], int32_t precision, int32_t brute, float T_out[16],
              int32_t* iters, int32_t* ret_code, cvo_iter_record* trace, int32_t trace_cap);

/* inner_product_impl (CvoGPU.cu:1673-1732): sum of stored a at cap
 * nearest_neighbors_max, accumulated in double. */
double orc_inner_product(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                         const float T[16], float ell, int32_t precision);

/* Timing helper for the CPU baseline: runs iterations of the mirror algorithm
 * at fixed lengthscale on threads OpenMP threads, returns seconds per
 * iteration. */
double orc_time_iterations(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                           const float T[16], float ell, int32_t K, int32_t n_iter, int32_t threads);

/* ---- IRLS bundle adjustment (src/cvo/IRLS.cpp, IRLS_State_GPU.*) -------- */

/* Residual and reference Jacobians of PairwiseAnalyticalDiffFunctor
 * (include/UnifiedCvo/cvo/IRLS_Cost_CPU.hpp:305-357): r = a*||T1 p1 - T2 p2||^2,
 * J1 (1x12), J2 (1x12). */
int orc_ba_pair_residual(const double T1[12], const double T2[12], const double p1[3],
                         const double p2[3], double a, double* r, double J1[12], double J2[12]);
/* LocalParameterizationSE3::Plus / ComputeJacobian
 * (local_parameterization_se3.hpp:21-83). */
int orc_se3_plus(const double T[12], const double delta[6], double out[12]);
int orc_se3_plus_jacobian(const double T[12], double J[72]);

/* Whole IRLS window solve with the oracle's own LM on the reference's (r, J).
 * poses in/out n*12. */
int orc_align_multiframe(const cvo_params* p, int32_t n_frames, const cvo_cloud_view* frames,
                         double* poses, const uint8_t* hold_const, int32_t n_edges,
                         const int32_t* edges, int32_t threads, int32_t* iters,
                         double* per_iter_seconds, int32_t max_records);

/* Normal equations of one LM step over fixed pairs (for parity of the device
 * assembly): H (6n x 6n, row-major), g (6n), cost. */
int orc_ba_normal_equations(int32_t n_frames, const double* poses, int32_t n_pairs,
                            const int32_t* pair_frames /* 2 per pair */,
                            const float* p1, const float* p2, const float* a,
                            double* H, double* g, double* cost);

/* BinaryStateCPU semantics (src/cvo/IRLS_State_CPU.cpp:48-328): the same IRLS
 * window with the reference's CPU edge states -- frame 2 mapped into frame 1
 * and radius-searched over frame 1 at a fixed lengthscale (no range-ell, no
 * row cap), weight ck k (no semantic or geometric-type factor), every pair a
 * residual, lengthscale decayed every outer iteration while above the
 * minimum. */
int orc_align_multiframe_cpu_states(const cvo_params* p, int32_t n_frames, const cvo_cloud_view* frames,
                                    double* poses, const uint8_t* hold_const, int32_t n_edges,
                                    const int32_t* edges, int32_t threads, int32_t* iters,
                                    double* per_iter_seconds, int32_t max_records);
/* One BinaryStateCPU association: pairs (frame-1 row, frame-2 column, weight),
 * in frame-2 order then frame-1 order; the first cap are written. */
int orc_cpu_state_pairs(const cvo_params* p, const cvo_cloud_view* c1, const cvo_cloud_view* c2,
                        const double X1[12], const double X2[12], float ell, int32_t cap, int32_t* rows,
                        int32_t* cols, double* weights, int64_t* nnz);

/* ---- the reference's CPU two-view CVO (src/cvo/Cvo.cpp:349-1381) --------- */

/* The 27-number parameter file of cvo::cvo (Cvo.cpp:89-122) reduced to the
 * fields align() reads. */
typedef struct {
  float ell_init, ell_min, ell_max;
  float sigma, sp_thres, c, d, c_ell, c_sigma, s_ell, s_sigma;
  int32_t MAX_ITER;
  float min_step, eps, eps_2;
  float ell_reduced_1;   /* lengthscale factor of a decrease (1/factor on an increase) */
  float ell_reduced_2;   /* indicator window length */
  float ell_reduced_3;   /* indicator stability threshold */
  int32_t use_semantics; /* IS_USING_SEMANTICS (not defined in the reference's CPU library) */
} orc_legacy_params;

typedef struct {
  float ell, step, dist, indicator;  /* indicator = A.sum() of the iteration */
  int32_t nnz;
  float omega[3], v[3];              /* the (unnormalised) flow */
  float bcde[4];
} orc_legacy_record;

/* cvo::cvo::set_pcd + align: T_init is the init guess (R, T), T_out the final
 * `transform` [R^T, -R^T T]; iters as the reference counts them. */
int orc_cvo_legacy_align(const orc_legacy_params* lp, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                         const float T_init[16], int32_t threads, float T_out[16], int32_t* iters, int32_t* ret,
                         double* seconds, orc_legacy_record* trace, int32_t trace_cap);
/* compute_indicator over a scripted sequence of A.sum() values: the
 * lengthscale after each call. */
int orc_legacy_indicator_sequence(const orc_legacy_params* lp, int32_t n, const float* ind, float* ell_out);

/* compute_association_gpu with a non-isotropic kernel (CvoGPU.cu:719-830,
 * :1873-1952): counts n, cols / vals n*K (first K passing targets by index).
 * Returns 2 for a singular kernel. */
int orc_assoc_nonisotropic(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                           const float T[16], const float kernel[9], int32_t K, int32_t* counts, int32_t* cols,
                           float* vals);

/* orc_pair_value over n (i, j) pairs (a = -1 for a rejected pair). */
int orc_pair_values(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt, const float T[16],
                    float ell, int64_t n, const int32_t* i, const int32_t* j, double* a_out, double* margin_out);
/* CvoGPU::inner_product_cpu (CvoGPU.cpp:114-231): fixed lengthscale, radius
 * gate only, no row cap; FP64 sum of the float pair values. */
double orc_inner_product_cpu(const cvo_params* p, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                             const float T[16], float ell, int32_t brute, int64_t* pairs_out);

// Compiler and flags this library was built with (for benchmark reports).
const char* orc_build_info(void);

#ifdef __cplusplus
}
#endif
#endif
