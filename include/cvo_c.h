/*
 * cvo_c.h -- C ABI of the MI355X (gfx950) CVO / RKHS-BA registration library.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns a
 * cvo_status and never exits the process.  Each declaration names the
 * reference interface it replaces (paths relative to the UnifiedCvo tree).
 *
 * Conventions shared with the reference:
 *   - A pose argument "T" of cvo_align / cvo_inner_product / cvo_function_angle /
 *     cvo_compute_association is the reference's init_guess_transform: the
 *     target cloud is mapped into the source frame by T^-1
 *     (src/cvo/CvoGPU.cu:98-130 update_tf, :1306-1307).
 *   - cvo_align's T_out is the reference's `transform` output, i.e. the map that
 *     takes target coordinates into the source frame ([R^T, -R^T t]).
 *   - 4x4 matrices are row-major float[16]; BA poses are row-major 3x4 double[12]
 *     exactly like CvoFrame::pose_vec (include/UnifiedCvo/cvo/CvoFrame.hpp:26).
 */
#ifndef CVO_C_H
#define CVO_C_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  CVO_OK = 0,
  CVO_ERR_INVALID_ARG = 1,   /* null pointer, negative size, bad dimension    */
  CVO_ERR_EMPTY = 2,         /* an input cloud has zero points                 */
  CVO_ERR_IO = 3,            /* file could not be opened                       */
  CVO_ERR_PARSE = 4,         /* YAML value could not be converted              */
  CVO_ERR_HIP = 5,           /* a HIP runtime call failed                      */
  CVO_ERR_NO_DEVICE = 6,     /* no HIP device visible                          */
  CVO_ERR_OOM = 7,           /* device allocation failed                       */
  CVO_ERR_UNSUPPORTED = 8,   /* option not implemented (e.g. is_using_kdtree)  */
  CVO_ERR_LIMIT = 9          /* size beyond a documented limit                 */
} cvo_status;

/*
 * cvo_params -- field-for-field the reference's CvoParams
 * (include/UnifiedCvo/cvo/CvoParams.hpp:12-79), same names and types.  Every
 * field has a default (cvo_params_default); the reference leaves max_step and
 * step uninitialised (:80-136), here they default to 0.01 and 0.
 */
typedef struct {
  float ell_init_first_frame;
  float ell_init;
  float ell_min;
  int32_t min_ell_iter_limit;
  float ell_max;
  double dl;
  double dl_step;
  float sigma;
  float sp_thres;
  float c;
  float d;
  float c_ell;
  float c_sigma;
  float s_ell;
  float s_sigma;
  int32_t MAX_ITER;
  float eps;
  float eps_2;
  float min_step;
  float max_step;
  float step;
  int32_t nearest_neighbors_max;
  float ell_decay_rate;
  float ell_decay_rate_first_frame;
  int32_t ell_decay_start;
  int32_t ell_decay_start_first_frame;
  int32_t indicator_window_size;
  float indicator_stable_threshold;
  int32_t is_pcl_visualization_on;
  int32_t is_using_least_square;
  int32_t is_ell_adaptive;
  int32_t is_full_ip_matrix;
  int32_t is_using_geometry;
  int32_t is_using_intensity;
  int32_t is_using_semantics;
  int32_t is_using_range_ell;
  int32_t is_using_kdtree;
  int32_t is_exporting_association;
  int32_t is_using_geometric_type;
  int32_t is_global_angle_registration;
  int32_t multiframe_using_cpu;
  int32_t multiframe_max_iters;
  float multiframe_ell_init;
  float multiframe_ell_min;
  float multiframe_ell_max;
  int32_t multiframe_iter_per_ell;
  float multiframe_ell_decay_rate;
  int32_t multiframe_iterations_per_ell;
  int32_t multiframe_iterations_per_solve;
  int32_t multiframe_expected_points;
  float multiframe_downsample_voxel_size;
  int32_t multiframe_num_neighbors;
  int32_t multiframe_least_squares_num_threads;
  int32_t multiframe_min_nonzeros;
  int32_t multiframe_is_optimizing_ell;
  int32_t multiframe_is_sorting_inner_product;
  int32_t multiframe_is_free_memory_each_iter;
} cvo_params;

/* A borrowed host view of one point cloud (CvoPointCloud, include/UnifiedCvo/
 * utils/CvoPointCloud.hpp:38-221, in structure-of-arrays form).
 * xyz: n*3; feat: n*feat_dim or NULL; sem: n*num_classes or NULL;
 * geo_type: n*2 or NULL.  Read only during the call. */
typedef struct {
  int32_t n;
  int32_t feat_dim;
  int32_t num_classes;
  const float* xyz;
  const float* feat;
  const float* sem;
  const float* geo_type;
} cvo_cloud_view;

/* Association in CSR form (replaces cvo::Association, include/UnifiedCvo/cvo/
 * Association.hpp:7-14).  Rows are source points in input order; within a row
 * the target ids ascend.  Arrays are owned by the library; free with
 * cvo_assoc_free. */
typedef struct {
  int32_t rows;
  int32_t cols;            /* number of target points                         */
  int64_t nnz;
  int64_t* row_ptr;        /* rows+1                                          */
  int32_t* col;            /* nnz target ids                                  */
  float* val;              /* nnz kernel values a_ij                          */
} cvo_assoc;

typedef struct cvo_ctx cvo_ctx;       /* params + HIP stream + workspace     */
typedef struct cvo_dcloud cvo_dcloud; /* a cloud resident in device memory   */

/* Per-iteration record of an align call (optional trace). */
typedef struct {
  float ell;
  float step;
  float ip;               /* indicator nnz / sqrt(N*M)                       */
  int32_t nnz;
  int32_t max_row;
  int32_t num_neighbors;  /* row cap K used in this iteration                */
  float dist;             /* ||log(dT)|| of the step                         */
  float flow_norm;        /* ||[omega; v]|| before normalisation             */
} cvo_iter_record;

/* Device time and algorithmic traffic of the library's kernels, accumulated
 * while cvo_set_profiling(ctx, 1) is on.  Index: CVO_KSTAT_*. */
enum {
  CVO_KSTAT_ASSOC = 0,     /* association + flow                            */
  CVO_KSTAT_OVERFLOW = 1,  /* capped-row exact path                         */
  CVO_KSTAT_STEP = 2,      /* step terms + controller                       */
  CVO_KSTAT_GRID = 3,      /* grid build                                    */
  CVO_KSTAT_BA_PAIR = 4,   /* BA pair weights / normal equations            */
  CVO_KSTAT_BA_SOLVE = 5,  /* BA FP64 solve                                 */
  CVO_KSTAT_MLP = 6,       /* learned proposer forward (bytes[] holds flops) */
  CVO_KSTAT_ANISO = 7,     /* non-isotropic association                     */
  CVO_KSTAT_COUNT = 8
};
typedef struct {
  double ms[CVO_KSTAT_COUNT];        /* summed device milliseconds          */
  int64_t launches[CVO_KSTAT_COUNT];
  double bytes[CVO_KSTAT_COUNT];     /* summed algorithmic bytes            */
  double pairs_stored;               /* sum over iterations of nnz          */
  double candidates;                 /* grid candidates visited (overhead)  */
} cvo_kernel_stats;

/* ---- parameters ---------------------------------------------------------- */

/* Defaults of CvoParams::CvoParams() (CvoParams.hpp:80-136). */
cvo_status cvo_params_default(cvo_params* out);
/* read_CvoParams_yaml (CvoParams.hpp:203-323): a key present in the file
 * overrides the value already in *p; absent keys keep it; unknown keys are
 * ignored.  Accepts the %YAML:1.0 header and '#' comments. */
cvo_status cvo_params_read_yaml(const char* path, cvo_params* p);

/* ---- context ------------------------------------------------------------- */

/* CvoGPU::CvoGPU(const std::string& yaml) (src/cvo/CvoGPU.cu:65-74). */
cvo_status cvo_create(const char* yaml_path, cvo_ctx** out);
cvo_status cvo_create_from_params(const cvo_params* p, cvo_ctx** out);
/* CvoGPU::get_params / write_params (CvoGPU.hpp, CvoGPU.cu:76-80). */
cvo_status cvo_get_params(const cvo_ctx* ctx, cvo_params* out);
cvo_status cvo_set_params(cvo_ctx* ctx, const cvo_params* p);
void cvo_destroy(cvo_ctx* ctx);
const char* cvo_status_string(cvo_status s);
/* Last error text of this context (HIP error string or argument check). */
const char* cvo_last_error(const cvo_ctx* ctx);

/* ---- device-resident clouds --------------------------------------------- */

/* CvoPointCloud_to_gpu (src/cvo/CvoGPU_impl.cu:220-289): copies the view
 * into device memory once, so repeated calls skip the upload. */
cvo_status cvo_cloud_upload(cvo_ctx* ctx, const cvo_cloud_view* v, cvo_dcloud** out);
void cvo_cloud_free(cvo_ctx* ctx, cvo_dcloud* c);
int32_t cvo_cloud_size(const cvo_dcloud* c);

/* ---- two-view registration ---------------------------------------------- */

/* CvoGPU::align(source, target, init_guess_T, transform, association,
 * registration_seconds, num_iters) (src/cvo/CvoGPU.cu:1516-1563,
 * align_impl :1239-1477).  ret_code: 0, or -1 when the gradient vanished.
 * Empty input: returns CVO_ERR_EMPTY with ret_code 0 and T_out = T_init^-1
 * semantics of the reference's early return (it leaves transform untouched;
 * here T_out is set to the identity).  seconds = device time of the
 * iteration loop, like the reference's cudaEvent timer.  With
 * is_global_angle_registration > 0 the registration starts from the inverse
 * of cvo_get_nearest_init_pose(source, target, ell_init_first_frame) instead
 * of T_init, and seconds includes the search (CvoGPU.cu:1531-1543,
 * :1555-1556). */
cvo_status cvo_align(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                     const float T_init[16], float T_out[16], cvo_assoc* assoc_out,
                     double* seconds, int32_t* iters, int32_t* ret_code);
cvo_status cvo_align_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt,
                       const float T_init[16], float T_out[16], cvo_assoc* assoc_out,
                       double* seconds, int32_t* iters, int32_t* ret_code);

/* CvoResultInfo CvoGPU::align(source, target, init_guess_T_s2t,
 * is_returning_association) (include/UnifiedCvo/cvo/CvoGPU.hpp:75-80,
 * src/cvo/CvoGPU.cu:1566-1611): the registration starts from the inverse of
 * T_init_s2t (a general 4x4 inverse) or, with is_global_angle_registration > 0,
 * from cvo_get_nearest_init_pose's result itself -- not its inverse: the
 * reference's two overloads differ there (:1539 against :1587).  T_out,
 * assoc_out, seconds and iters as cvo_align.  Empty input returns
 * CVO_ERR_EMPTY with ret_code -1 (the reference's return_code). */
cvo_status cvo_align_s2t(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                         const float T_init_s2t[16], float T_out[16], cvo_assoc* assoc_out, double* seconds,
                         int32_t* iters, int32_t* ret_code);
cvo_status cvo_align_s2t_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, const float T_init_s2t[16],
                           float T_out[16], cvo_assoc* assoc_out, double* seconds, int32_t* iters,
                           int32_t* ret_code);

/* AdaptiveCvoGPU::align (AdaptiveCvoGPU.hpp:171-175, AdaptiveCvoGPU.cu:1077-1248):
 * the lengthscale follows its gradient instead of the indicator schedule.
 * Each iteration associates source-target at the fixed lengthscale plus the
 * two self associations (source-source, moved target-target), steps along the
 * unnormalised flow (omega / c, v / d; step = smallest positive real root of
 * the cubic, min_step if none, at most 0.8), then
 * ell <- ell - dl_step * dl with
 *   dl = (2/l^3 sum_A -a|y-x|^2 + 1/l^3 sum_Axx a|x'-x|^2 + 1/l^3 sum_Ayy a|y'-y|^2)
 *        / (nnz(Axx) + nnz(Ayy) - 2 nnz(A)),
 * shrinking ell and ell_max to 0.7 ell_max past ell_max, at least ell_min.
 * Association: colour gate always on (features required on both clouds),
 * class factor without a gate when is_using_semantics, no geometric type, the
 * first min(512, nearest_neighbors_max) passing targets by index.  Pose
 * conventions and stop tests as cvo_align; the trace records each iteration's
 * lengthscale. */
cvo_status cvo_align_adaptive(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                              const float T_init[16], float T_out[16], double* seconds, int32_t* iters,
                              int32_t* ret_code);
cvo_status cvo_align_adaptive_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt,
                                const float T_init[16], float T_out[16], double* seconds, int32_t* iters,
                                int32_t* ret_code);
/* AdaptiveCvoGPU::inner_product (AdaptiveCvoGPU.hpp:178-180,
 * AdaptiveCvoGPU.cu:1338-1371, se_kernel_init_ell_cpu :1251-1336): the target
 * moved by T_s2t itself (rot y + trans), every pair that passes the adaptive
 * association's gates at ell_init (no row cap), sum(a) / N * 1e6 / M.  The
 * reference computes it on the CPU; here it is one uncapped association
 * pass.  0 on empty input. */
cvo_status cvo_inner_product_adaptive(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                      const float T_s2t[16], float* out);

/* Optional per-iteration trace of the next align call(s): up to cap records. */
cvo_status cvo_set_trace(cvo_ctx* ctx, int32_t cap);
cvo_status cvo_get_trace(cvo_ctx* ctx, cvo_iter_record* out, int32_t cap, int32_t* n_out);

/* ---- inner product / cos angle / association ---------------------------- */

/* CvoGPU::inner_product_gpu (src/cvo/CvoGPU.cu:1673-1748): sum of all stored
 * a_ij at pose T and lengthscale ell, row cap nearest_neighbors_max. */
cvo_status cvo_inner_product(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                             const float T[16], float ell, double* ip);
cvo_status cvo_inner_product_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt,
                               const float T[16], float ell, double* ip);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): cvo_inner_product_d computed with the same kernels in the
 * library's fixed per-GPU scratch array (64 MiB, a file-scope device array
 * that each GPU the library runs on holds until the process ends) instead of
 * the context's arena.  Calls no hipMalloc, hipFree or allocating library
 * routine; the arena is left as it was.  At most 131072 points per cloud, and
 * the pass must fit the scratch array (its grid is capped at 2^21 cells):
 * CVO_ERR_LIMIT otherwise.  Every call on one GPU shares that GPU's array and
 * keeps nothing from earlier calls: calls that overlap on one GPU (several
 * host threads or contexts) race. */
cvo_status cvo_inner_product_d_noalloc(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt,
                                       const float T[16], float ell, double* ip);
/* CvoGPU::function_angle (src/cvo/CvoGPU.cu:1768-1806): approx != 0 gives
 * ip / (sqrt(N) sqrt(M)), otherwise ip / sqrt(ip_xx ip_yy). */
cvo_status cvo_function_angle(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                              const float T[16], float ell, int32_t approx, double* cos_out);
/* CvoGPU::inner_product_cpu (include/UnifiedCvo/cvo/CvoGPU.hpp:230-233,
 * src/cvo/CvoGPU.cpp:204-231, se_kernel_init_ell_cpu :114-186), computed on
 * the GPU with that function's semantics: the target moved into the source
 * frame by T^-1, every pair within the squared radius
 * -2 ell^2 log(sp_thres / sigma^2) at the fixed lengthscale (no range
 * scaling), no row cap and no colour, semantic or geometric-type gate; a pair
 * adds a = ck k sk (k only with is_using_geometry, ck with is_using_intensity,
 * sk with is_using_semantics) when a > sp_thres.  0 on empty input. */
cvo_status cvo_inner_product_cpu(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                 const float T[16], float ell, double* ip);
/* CvoGPU::function_angle with its is_gpu argument (CvoGPU.cu:1768-1806):
 * is_gpu != 0 uses cvo_inner_product's association, is_gpu == 0 the
 * inner_product_cpu semantics of cvo_inner_product_cpu (:1780-1783, :1797-1800). */
cvo_status cvo_function_angle_ex(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                 const float T[16], float ell, int32_t approx, int32_t is_gpu, double* cos_out);
/* CvoGPU::compute_association_gpu (src/cvo/CvoGPU.cu:1836-1871). */
cvo_status cvo_compute_association(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                   const float T[16], float ell, cvo_assoc* out);
/* CvoGPU::compute_association_gpu with a non-isotropic kernel
 * (CvoGPU.hpp:241-246, CvoGPU.cu:1873-1952, fill_in_A_mat_gpu_dense_mat_kernel
 * :719-830): kernel is the row-major 3x3 K; per source point x,
 * M = U K^-1 U^T with U = [u, z x u, z], z = x/|x|, u = (1, 1, -(z0+z1)/z2),
 * k = sigma^2 exp(-(x-y)^T M (x-y) / 2) (no lengthscale, no distance gate),
 * colour and semantic gates as the isotropic kernel, geometric type off; a row
 * keeps the first nearest_neighbors_max passing targets by index.  Rows in
 * source order, target ids ascending.  CVO_ERR_INVALID_ARG for a non-finite or
 * singular kernel. */
cvo_status cvo_compute_association_nonisotropic(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                                const float T[16], const float kernel[9], cvo_assoc* out);
/* One K-nearest association pass, the one CvoGPU::align runs each iteration
 * with is_using_kdtree (se_kernel_kdtree, CvoGPU.cu:1104-1154;
 * fill_in_A_mat_cukdtree :832-949), at pose T, lengthscale ell and row cap K
 * (1..512): per source point the K nearest targets (ties by ascending target
 * index) that pass the gates at the fixed lengthscale (no range scaling),
 * geometric-type floor 0.001.  The reference keeps this pass inside align
 * (no entry point of its own); it is exposed here for inspection.  Rows in
 * source order, target ids ascending.  CVO_ERR_UNSUPPORTED without
 * is_using_geometry. */
cvo_status cvo_compute_association_knn(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt,
                                       const float T[16], float ell, int32_t K, cvo_assoc* out);
void cvo_assoc_free(cvo_assoc* a);

/* ---- global initialisation ------------------------------------------------ */

/* The 60 rotations of gen_icosahedron_init_rots (src/cvo/CvoGPU.cpp:279-342),
 * row-major 3x3 each, in that table's order.  Generated (the icosahedral group
 * in the table's orientation, FP64, rounded to float once); they agree with
 * the reference's float literals to 1e-6.  Needs no device. */
cvo_status cvo_icosahedral_rotations(float out[540]);

/* CvoGPU::get_nearest_init_pose (include/UnifiedCvo/cvo/CvoGPU.hpp:249-253,
 * src/cvo/CvoGPU.cpp:367-398): each rotation R of the table is scored by the
 * approximate cos angle function_angle(source, target, [R^T 0; 0 1], ell) --
 * all 60 in one batched association pass -- and, scanning in table order, a
 * score replaces the best only when strictly larger, starting from 0; T_out =
 * [R 0; 0 1] of the winner (source to target frame), the identity when no
 * score is positive.  num_guess is accepted and unused, as in the reference.
 * seconds (optional) is incremented by the search's wall time (the reference
 * adds its whole seconds). */
cvo_status cvo_get_nearest_init_pose(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt, float ell,
                                     int32_t num_guess, float T_out[16], double* seconds);
cvo_status cvo_get_nearest_init_pose_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, float ell,
                                       int32_t num_guess, float T_out[16], double* seconds);

/* ---- batched multi-start global registration ----------------------------- */

/* Replaces CvoGPU::get_nearest_init_pose + align (src/cvo/CvoGPU.cpp:367-398,
 * src/cvo/CvoGPU.cu:1533-1543).  T_starts: n_starts initial poses (row-major
 * 4x4, the align init convention).  mode 0: every start runs the full align
 * loop, all starts batched into one launch per kernel per iteration with
 * per-start lengthscale annealing; each final pose is scored by the cos angle
 * (function_angle, approximate form) at score_ell and the largest wins.
 * mode 1: the reference's global initialisation -- score the initial poses
 * only, keep the first maximum, then one align from it.  T_best: the winning
 * registration's transform (target -> source, like cvo_align's T_out).
 * per_start_T (n*16), per_start_cos (n), per_start_iters (n) are optional. */
cvo_status cvo_align_multistart(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, int32_t n_starts,
                                const float* T_starts, float score_ell, int32_t mode, float T_best[16],
                                int32_t* best_index, float* per_start_T, double* per_start_cos,
                                int32_t* per_start_iters, double* seconds);

/* ---- multi-frame IRLS bundle adjustment ---------------------------------- */

/* Replaces CvoGPU::align(frames, hold_const, edges, seconds) and
 * CvoGPU::align(frames, hold_const, edge_states, seconds, inliers, num_iters)
 * (include/UnifiedCvo/cvo/CvoGPU.hpp:146-164, src/cvo/CvoGPU.cu:1614-1671,
 * :1955-1999) together with CvoBatchIRLS::solve (src/cvo/IRLS.cpp:148-367)
 * and BinaryStateGPU (src/cvo/IRLS_State_GPU.cu:144-351,
 * src/cvo/IRLS_State_GPU.cpp:10-146).  frames: n_frames device clouds (sensor
 * frame).  poses: n_frames * 12 doubles, row-major 3x4 world <- sensor
 * (CvoFrame::pose_vec), updated in place.  hold_const (nullable): nonzero
 * keeps a frame fixed.  edges: 2 * n_edges frame indices (f1, f2); each edge
 * associates f1's points (rows) with f2's.  per_edge (nullable, n_edges
 * entries): each edge's association of the last outer iteration that ran
 * one, in CSR form (rows: f1's points, cols: f2's ids ascending) -- the
 * `inliers` list of BinaryState::export_association (IRLS_State_GPU.cu:
 * 345-351); free each with cvo_assoc_free.  Limits: 64 frames, 20 free frames,
 * 4096 edges.  per_iter_seconds (optional, capacity max_iter_records)
 * receives the wall time of each IRLS outer iteration; seconds the whole
 * call; iters the outer iterations run. */
cvo_status cvo_align_multiframe(cvo_ctx* ctx, int32_t n_frames, const cvo_dcloud* const* frames, double* poses,
                                const uint8_t* hold_const, int32_t n_edges, const int32_t* edges,
                                cvo_assoc* per_edge, double* seconds, int32_t* iters, double* per_iter_seconds,
                                int32_t max_iter_records);

/* ---- learned pose proposer (no counterpart in the reference) -------------- */

/* For global registration callers (the multi-start path): a point-set
 * encoder shared by both clouds (8 -> 64 -> 128 -> 256, max-pool) and a head
 * (512 -> 256 -> 8 x 7 + 8) proposing 8 rigid transforms with confidences.
 * BF16 weights and activations, FP32 accumulation on MFMA, FP32 master weights
 * trained with Adam by the library's own forward and backward kernels.
 * Clouds are passed as n_pairs x npts points of 8 floats (xyz in metres, 5
 * features); npts is a multiple of 64.  Parameters are one flat FP32 vector
 * (W1 8x64, b1, W2 64x128, b2, W3 128x256, b3, W4 512x256, b4, W5 256x64, b5;
 * row-major, input index first). */
typedef struct cvo_model cvo_model;
cvo_status cvo_model_create(cvo_ctx* ctx, int32_t k, uint64_t seed, cvo_model** out); /* k must be 8 */
void cvo_model_destroy(cvo_model* m);
int64_t cvo_model_num_params(void);
cvo_status cvo_model_get_weights(const cvo_model* m, float* out, int64_t cap, int64_t* n_out);
cvo_status cvo_model_set_weights(cvo_model* m, const float* w, int64_t n); /* also resets Adam state */
/* targets: n_pairs x 7 (unit quaternion w,x,y,z and translation of the
 * source-to-target transform, the align-init convention).  Mean loss over
 * the batch; grad_out (optional, num_params floats). */
cvo_status cvo_model_gradient(cvo_ctx* ctx, cvo_model* m, int32_t n_pairs, int32_t npts, const float* a,
                              const float* b, const float* targets, float* grad_out, float* loss_out);
/* One Adam step (beta 0.9 / 0.999) with learning rate lr; loss_out optional. */
cvo_status cvo_model_train_step(cvo_ctx* ctx, cvo_model* m, int32_t n_pairs, int32_t npts, const float* a,
                                const float* b, const float* targets, float lr, float* loss_out);
/* Batched inference: quat_t n_pairs x 8 x 7 (unit quaternion with w >= 0,
 * translation in metres; source-to-target), conf n_pairs x 8 (softmax);
 * raw_out (optional) n_pairs x 64 head outputs. */
cvo_status cvo_propose_poses(cvo_ctx* ctx, const cvo_model* m, int32_t n_pairs, int32_t npts, const float* a,
                             const float* b, float* quat_t, float* conf, float* raw_out);
/* cvo_propose_poses with the two point batches already in device memory
 * (a_dev, b_dev: device pointers, n_pairs x npts x 8 floats each, written
 * before the call on any stream the caller has synchronised); outputs to
 * host arrays as above. */
cvo_status cvo_propose_poses_d(cvo_ctx* ctx, const cvo_model* m, int32_t n_pairs, int32_t npts, const float* a_dev,
                               const float* b_dev, float* quat_t, float* conf, float* raw_out);
/* The BF16 MFMA GEMM on host arrays, C (M x N) = A (M x K) B (K x N), inputs
 * rounded to BF16, FP32 result: the tile kernel the proposer runs on. */
cvo_status cvo_gemm_bf16(cvo_ctx* ctx, int32_t M, int32_t N, int32_t K, const float* A, const float* B, float* C);

/* ---- pose arithmetic ------------------------------------------------------ */

/* The align loop's pose update.  CVO_POSE_FP64 (default): FP64 pose, SE(3)
 * exponential and step.  CVO_POSE_REFERENCE: the reference's float pose with
 * its FP32 Exp_SEK3 (LieGroup.cpp:244-275, small-angle branch included) and a
 * float step (CvoGPU.cu:1289-1341), for bit-level comparisons with it. */
enum { CVO_POSE_FP64 = 0, CVO_POSE_REFERENCE = 1 };
cvo_status cvo_set_pose_arithmetic(cvo_ctx* ctx, int32_t mode);

/* ---- association-kernel tuning (no counterpart in the reference) ----------
 *
 * The reference's association kernel has one fixed shape (one thread per
 * source row over every target, CvoGPU.cu:951-1099); here the grid search has
 * knobs, none of which changes the pair set, the pair values or the row cap
 * semantics (only the order a row's pairs are summed in, at FP64 rounding):
 *   lanes_per_row    4, 8 or 16 lanes of a wave share one source row, at
 *                    lengthscales >= split_ell (every lengthscale when
 *                    split_ell <= 0); launches over several problems
 *                    (multi-start, adaptive, BA) and the K-nearest pass use
 *                    at least 8;
 *   cell_scale       grid cells are chosen with edge >= cell_scale x the
 *                    ell-ball radius, in [0.4, 1] (smaller: more cells per
 *                    query box, fewer candidates outside the ball), at
 *                    lengthscales >= split_ell;
 *   lanes_small_ell, cell_scale_small_ell
 *                    the same below split_ell (the device controller applies
 *                    them per iteration, so they follow the annealing inside
 *                    the captured iteration graph);
 *   xcd_remap        1: each of the eight XCDs takes one contiguous range of
 *                    the spatially sorted row blocks (its own L2 serves the
 *                    overlapping target neighbourhoods), 0: dispatch order.
 * The tuning of a context applies to every later align / multi-start / BA /
 * inner-product call on it. */
typedef struct {
  int32_t lanes_per_row;
  int32_t lanes_small_ell;
  float split_ell;
  int32_t xcd_remap;
  float cell_scale;
  float cell_scale_small_ell;
  int32_t source;   /* out: CVO_TUNE_DEFAULT / _SET / _MEASURED / _CACHED */
  float score_us;   /* out: modelled association time per iteration of the chosen setting */
  int32_t n_tried;  /* out: (setting, lengthscale) pairs timed */
} cvo_tuning;
enum { CVO_TUNE_DEFAULT = 0, CVO_TUNE_SET = 1, CVO_TUNE_MEASURED = 2, CVO_TUNE_CACHED = 3 };

cvo_status cvo_get_tuning(const cvo_ctx* ctx, cvo_tuning* out);
/* CVO_ERR_INVALID_ARG for lane counts not in {8, 16}, a negative or
 * non-finite split_ell, xcd_remap not in {0, 1} or a cell scale outside
 * [0.4, 1]. */
cvo_status cvo_set_tuning(cvo_ctx* ctx, const cvo_tuning* t);
/* Times the association pass of this source / target pair at T_init for
 * every lane count (8, 16), block order (2) and cell scale (1, 0.7, 0.5) at
 * four lengthscales log-spaced from ell_init to ell_min, then keeps the
 * block order and the two-segment (lanes, cell scale) policy with the least
 * modelled time per registration: the three annealing samples share the
 * iterations of the decay schedule (decays x 2 x indicator window), ell_min
 * takes the rest of MAX_ITER.  With cache_path non-null the result is looked up first under
 * the workload key (cvo_tuning_key) and stored after a measurement (a JSON
 * file, replaced atomically).  The measurement takes well under a second at
 * 60k points. */
cvo_status cvo_autotune(cvo_ctx* ctx, const cvo_cloud_view* src, const cvo_cloud_view* tgt, const float T_init[16],
                        const char* cache_path, cvo_tuning* out);
cvo_status cvo_autotune_d(cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, const float T_init[16],
                          const char* cache_path, cvo_tuning* out);
/* The cache key of a workload on this context's device: architecture, CU
 * count, half-octave size classes of both clouds, feature / class dimensions,
 * enabled gates, lengthscale range, sparsity threshold and row cap. */
cvo_status cvo_tuning_key(const cvo_ctx* ctx, const cvo_dcloud* src, const cvo_dcloud* tgt, char* buf, int32_t len);
/* Cache file access without a device: lookup returns CVO_ERR_IO when the
 * file or the key is absent; store creates the file (and its directory). */
cvo_status cvo_tuning_cache_lookup(const char* path, const char* key, cvo_tuning* out);
cvo_status cvo_tuning_cache_store(const char* path, const char* key, const cvo_tuning* t);

/* ---- workspace ------------------------------------------------------------
 * Replaces CvoState (include/UnifiedCvo/cvo/CvoState.cuh:22-82), which the
 * reference allocates on every align call and whose N x K pair matrix it
 * clears every iteration (src/cvo/CvoState.cu:22-157).  A context keeps one
 * device arena, grown on demand (25% headroom) and reused by later calls; the
 * row counts make a clear unnecessary. */
/* Device bytes the context's arena holds now. */
cvo_status cvo_workspace_bytes(const cvo_ctx* ctx, uint64_t* out);
/* Caps the arena (0: no cap).  A call that would grow it beyond the cap fails
 * with CVO_ERR_LIMIT before allocating, and the context stays usable. */
cvo_status cvo_set_workspace_limit(cvo_ctx* ctx, uint64_t bytes);
/* Frees the arena (the next call grows it again). */
cvo_status cvo_release_workspace(cvo_ctx* ctx);

/* ---- profiling ----------------------------------------------------------- */

cvo_status cvo_set_profiling(cvo_ctx* ctx, int32_t on);
cvo_status cvo_get_kernel_stats(cvo_ctx* ctx, cvo_kernel_stats* out);
cvo_status cvo_reset_kernel_stats(cvo_ctx* ctx);

/* Library version string. */
const char* cvo_version(void);

/* No counterpart in the reference: runs the device's controller math on the
 * given inputs so that tests can hold it against published values (numpy
 * roots, scipy expm).  Cubic i: roots of coef[4i] s^3 + coef[4i+1] s^2 +
 * coef[4i+2] s + coef[4i+3] into re[3i..3i+2] / im[3i..3i+2] (the closed form
 * the step selection uses, CvoGPU.cu:495-536).  Twist i: T64[12i..] = the FP64
 * SE(3) exponential of dt[i] * xi[6i..6i+5] ([omega; v], row-major 3x4),
 * log_norm[i] = the norm of its logarithm (dist_se3, CvoGPU.cpp:28-33), T32 =
 * the reference's float Exp_SEK3 (LieGroup.cpp:244-275) of the same twist.
 * Uses the current device, allocates and synchronises. */
cvo_status cvo_device_math_check(int32_t n_cubic, const double* coef, double* re, double* im, int32_t n_exp,
                                 const double* xi, const double* dt, double* T64, double* log_norm, float* T32);

#ifdef __cplusplus
}
#endif
#endif /* CVO_C_H */
