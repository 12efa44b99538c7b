// Internal device/host data structures of the gfx950 CVO library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "cvo_c.h"

namespace cvo {

constexpr int kMaxLevels = 12;         // grid levels per target
constexpr int kMaxWindow = 64;         // indicator window capacity
constexpr int kBlock = 256;            // threads per block of the row kernels
constexpr int kOvfBlocks = 96;         // capped-row kernel blocks per problem
constexpr uint32_t kMaxCells = 1u << 25;
constexpr int kKnnMaxK = 512;          // K-nearest row cap (KDTREE_K_SIZE, cukdtree.h:12)

// One grid level over a target cloud (uniform dense grid, x-fastest cells).
struct GridLevel {
  float h, inv_h;
  float ox, oy, oz;
  int32_t nx, ny, nz;
  const uint32_t* cell_start;  // nx*ny*nz + 1
  const float4* pos;           // sorted: x, y, z, original id (bits)
  const float* feat;           // sorted, n*F
  const float* sem;            // sorted, n*K
  const float* geo;            // sorted, n*2
  const uint32_t* inv_perm;    // original id -> sorted position
};

struct GridDev {
  int32_t n_levels;
  int32_t n;
  GridLevel lv[kMaxLevels];
};

// Per-problem iteration state, resident in device memory.  One per
// registration (two-view), per start (multi-start) or per edge (BA).
struct alignas(16) ProblemState {
  // pose (FP64), R row-major: the reference's (R, T) of align_impl
  double R[9];
  double T[3];
  // per-iteration totals of the association pass (raw flow sums, before the
  // 1/c, 1/d scaling), written by the capped-row kernel's last block
  double flow[6];
  double ip;
  double bcde[4];
  double step;
  double dist;
  uint64_t nnz;
  // float transforms used by the kernels
  float A[9], b[3];     // target map y' = A y + b
  float Rq[9], Tq[3];   // query map x~ = Rq x + Tq into the target frame
  float As[9], bs[3];   // source map (identity for two-view)
  float ell;
  float r_need;         // query radius bound for level selection
  int32_t has_src_xf;
  int32_t K;            // row cap
  int32_t level;
  int32_t iter;
  int32_t done;
  int32_t ret;
  int32_t iters_out;
  uint32_t max_row;
  uint32_t n_overflow;
  uint32_t blocks_done_assoc;
  uint32_t blocks_done_ovf;
  uint32_t blocks_done_step;
  // indicator windows (A_sparsity_indicator_ell_update, float sums)
  float qs[kMaxWindow], qe[kMaxWindow];
  int32_t qs_head, qs_size, qe_head, qe_size;
  float ss, se;
  float flow_norm;
  int32_t level_assoc;  // grid level of the last association pass
  int32_t group;        // lanes per source row of the next association pass
  // adaptive lengthscale (AdaptiveCvoGPU): its current bound and the last
  // gradient step dl
  float ell_max;
  double dl;
};

// Read-only description of one association problem.
struct AssocProblem {
  const float4* src_pos;   // sorted source: x, y, z, original row id (bits)
  const float* src_feat;
  const float* src_sem;
  const float* src_geo;
  int32_t n_src;
  int32_t n_tgt;
  const GridDev* grid;
  const float* tgt_xyz;    // original order (capped-row path)
  const float* tgt_feat;
  const float* tgt_sem;
  const float* tgt_geo;
  ProblemState* st;
  uint2* pairs;            // n_src * cap_alloc: (sorted target position, a bits)
  float4* pairs4;          // bundle adjustment: n_src * cap_alloc (target xyz in its frame, a)
  uint32_t* pair_ids;      // bundle adjustment, optional: the sorted target position of each pairs4 slot
  uint32_t* counts;        // n_src
  uint32_t* ovf_stamp;     // n_src
  int32_t cap_alloc;
  int32_t block_begin;     // first assoc/step block of this problem
  int32_t n_blocks;        // assoc/step blocks of this problem
  int32_t trace_cap;
  double* part_a;          // per block: flow6, ip, -> 8 doubles
  double* part_ovf;        // per capped-row block: flow6, ip, nnz, max row, - -> 10 doubles
  double* part_s;          // per block: 4 doubles
  uint32_t* part_cnt;      // per block: nnz and max row
  cvo_iter_record* trace;  // per iteration (may be null)
  // adaptive lengthscale: the source and target self-association problems
  // whose totals the controller reads (null otherwise)
  ProblemState* xx_st;
  ProblemState* yy_st;
};

// Association launch shape as a function of the lengthscale: lanes per row
// and grid cell scale above / below split_ell (one setting when split_ell <= 0).
struct EllPolicy {
  float split_ell;
  int32_t group_big, group_small;  // 4, 8 or 16 (8 or 16 for the K-nearest pass)
  float scale_big, scale_small;    // level choice: cell edge >= scale * radius
};

__host__ __device__ inline bool policy_big(const EllPolicy& pol, float ell) {
  return !(pol.split_ell > 0.0f) || ell >= pol.split_ell;
}

// Scalar constants of the kernel (fill_in_A_mat_gpu, CvoGPU.cu:966-987).
struct KConsts {
  float sigma2, c_sigma2, s_sigma2;
  float c2, s_ell;
  float lg;                   // logf(sp / sigma2)
  float d2_c_thres, d2_s_thres;
  float q_c, q_s;             // 1 / (2 c2), 1 / (2 s_ell^2)
  float sp;
  float amp;                  // sigma2 * c_sigma2 * s_sigma2 (enabled factors)
  int32_t use_geo, use_col, use_sem, use_gtype;
  int32_t F, Ks;              // feature and class dimensions
  float c, d;                 // flow scales
  // controller constants
  float eps, eps_2, min_step, max_step;
  float ell_min, ell_decay_rate;
  int32_t ell_decay_start, max_iter, nn_max, window;
  float ind_thr;
  int32_t range_ell_in_step;
  float n_src_f, n_tgt_f;     // for the indicator (two-view: N, M)
  float rq_coef;              // sqrt(-2 lg) * 1.001
  float max_src_norm;
  int32_t mode;               // 0 align, 1 inner product, 2 BA
  int32_t mirror_pose;        // 1: the reference's FP32 Exp_SEK3 and float pose
  // association launch shape (none of these changes the pair set)
  EllPolicy pol;
  int32_t max_group;          // largest lane count of the policy (sizes the grid)
  int32_t xcd_remap;          // 1: contiguous row-block ranges per XCD
  float min_scale;            // finest cell scale of the policy (sizes the levels)
  int32_t knn;                // is_using_kdtree: K-nearest association, the K multiplier (0 off)
  // adaptive lengthscale (AdaptiveCvoGPU): fixed-lengthscale ball, the ip
  // column carries sum(a d2), unnormalised flow, the dl-gradient controller
  int32_t adaptive;
  double dl_step;
  float ell_max0;             // initial lengthscale bound (ell_max)
  int32_t fixed_ell;          // the association uses ell as is (no range scaling)
  int32_t ip_ad2;             // the ip column sums a |y - x|^2 instead of a
  int32_t uncapped;           // sum every passing pair (no row cap, no capped rows)
  int32_t no_geo_factor;      // the distance gate applies, the geometric factor k does not
};

// Rows per association block for a lane count per row (8 or 16).
__host__ __device__ constexpr int rows_per_block(int group) { return kBlock / group; }
constexpr int kMaxRowsPerBlock = kBlock / 4;  // 4, 8 or 16 lanes per row
constexpr int kKnnRowsPerBlock = kBlock / 8;  // the K-nearest pass: 8 or 16 lanes per row

enum Mode { MODE_ALIGN = 0, MODE_IP = 1, MODE_BA = 2 };

}  // namespace cvo

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

// A bump region of device memory that ensure() carves from instead of
// allocating (the no-allocation entry points): base / size of the region,
// used bytes so far.
struct BumpArena {
  char* base = nullptr;
  size_t size = 0, used = 0;
};

struct cvo_dcloud {
  int32_t n = 0, F = 0, K = 0;
  bool has_geo = false;
  float* xyz = nullptr;   // device, n*3
  float* feat = nullptr;  // device, n*F
  float* sem = nullptr;
  float* geo = nullptr;
  float bmin[3], bmax[3];  // bounding box
  float max_norm = 0;
  std::vector<float> host_xyz;  // kept for the bounding box / sort keys
};

struct cvo_ctx {
  cvo_params params;
  hipStream_t stream = nullptr;
  int device = 0;
  std::string last_error;
  int32_t trace_cap = 0;
  std::vector<cvo_iter_record> trace;
  int32_t trace_n = 0;
  bool profiling = false;
  cvo_kernel_stats stats{};
  // workspace arena, grown on demand; ws_bytes its total, ws_limit a cap (0 none)
  size_t ws_bytes = 0, ws_limit = 0;
  DevBuf b_src_pos, b_src_feat, b_src_sem, b_src_geo, b_src_key, b_src_perm;
  DevBuf b_grid_store;    // all levels' arrays
  DevBuf b_pairs, b_counts, b_ovf;
  DevBuf b_state, b_probs, b_gridhdr;
  DevBuf b_part_a, b_part_ovf, b_part_s, b_part_cnt;
  DevBuf b_trace;
  DevBuf b_tmp0, b_tmp1, b_tmp2, b_tmp3, b_cub;
  DevBuf b_export;
  void* pinned = nullptr;  // small pinned host buffer for flags
  size_t pinned_cap = 0;
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  std::vector<hipEvent_t> ev_pool;
  // BA scratch
  DevBuf b_ba[20];
  std::vector<DevBuf> b_frames;  // per frame: sorted source (4) and grid (2)
  int32_t ba_last_accepted = 0, ba_last_evals = 0;
  int32_t pose_arith = 0;  // CVO_POSE_FP64 or CVO_POSE_REFERENCE
  cvo_tuning tuning{16, 16, 0.0f, 1, 1.0f, 1.0f, CVO_TUNE_DEFAULT, 0.0f, 0};
  // model scratch
  DevBuf b_mlp[16];
  // adaptive-lengthscale scratch: sorted target rows (4), source grid (2)
  DevBuf b_acvo[8];
  // while set, ensure() carves from this region and never allocates
  BumpArena* bump = nullptr;
  // grid cell count limit of build_grid_into
  uint32_t grid_max_cells = 1u << 25;
  // device address of this device's scratch array (no-allocation calls)
  char* scratch = nullptr;
};

// Internal helpers shared by the translation units.
namespace cvo {
struct GridHost {
  GridDev dev;  // host copy of the device header
  GridDev* dev_ptr = nullptr;
};
cvo_status set_err(cvo_ctx* ctx, cvo_status s, const std::string& msg);
cvo_status hip_check(cvo_ctx* ctx, hipError_t e, const char* what);
cvo_status ensure(cvo_ctx* ctx, DevBuf& b, size_t bytes);
KConsts make_kconsts(const cvo_params& p, int F, int Ks, int mode);
cvo_status validate_pair(cvo_ctx* ctx, const cvo_dcloud* s, const cvo_dcloud* t);
float radius_for(const KConsts& kc, float ell, float max_norm, float tq);
std::vector<float> level_sizes(const KConsts& kc, float ell_hi, float ell_lo, float max_norm, float tq,
                               float cell_scale);
void apply_tuning(const cvo_ctx* ctx, KConsts& kc, bool batched = false);
void apply_tuning(const cvo_tuning& t, KConsts& kc, bool batched = false);
cvo_status ensure_events(cvo_ctx* ctx, size_t n);
cvo_status pinned(cvo_ctx* ctx, size_t bytes);
cvo_status prepare_source_into(cvo_ctx* ctx, const cvo_dcloud* s, DevBuf& b_pos, DevBuf& b_feat, DevBuf& b_sem,
                               DevBuf& b_geo);
cvo_status build_grid_into(cvo_ctx* ctx, const cvo_dcloud* t, const std::vector<float>& hs, GridHost& gh,
                           DevBuf& b_store, DevBuf& b_hdr);
// host_one: the host copy of a single problem's descriptor, passed by value;
// e0 / e1 (optional): events recorded at the kernel's own start and end
void launch_assoc(const AssocProblem* probs, int n_probs, int total_blocks, const KConsts& kc, hipStream_t s,
                  const AssocProblem* host_one = nullptr, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_overflow(const AssocProblem* probs, int n_probs, const KConsts& kc, hipStream_t s,
                     const AssocProblem* host_one = nullptr, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_init_state(const AssocProblem* probs, int n_probs, const float* T16, int stride16, const KConsts& kc,
                       float ell0, int K0, hipStream_t s);
cvo_status prepare_source(cvo_ctx* ctx, const cvo_dcloud* s);
cvo_status export_problem(cvo_ctx* ctx, const AssocProblem* P_dev, int n_src, int n_tgt, cvo_assoc* out);
cvo_status build_grid(cvo_ctx* ctx, const cvo_dcloud* t, const std::vector<float>& hs, GridHost& gh);

// Descriptors of n_probs association problems sharing one source and target.
struct Problems {
  int n_probs = 0;
  int total_blocks = 0;
  int max_rows = 0;  // largest source row count over the problems
  AssocProblem* dev = nullptr;
  std::vector<AssocProblem> host;
  ProblemState* states = nullptr;
  cvo_iter_record* trace = nullptr;
  int trace_cap = 0;
};
// max_group: the largest lane count any level uses (fewest rows per block).
cvo_status setup_problems(cvo_ctx* ctx, const cvo_dcloud* s, const cvo_dcloud* t, const GridHost& gh, int n_probs,
                          int cap_alloc, int trace_cap, int max_group, Problems& pr);
}  // namespace cvo

#define CVO_HIP(ctx, call)                                              \
  do {                                                                  \
    hipError_t _e = (call);                                             \
    if (_e != hipSuccess) return cvo::hip_check((ctx), _e, #call);      \
  } while (0)

#define CVO_TRY(expr)                      \
  do {                                     \
    cvo_status _s = (expr);                \
    if (_s != CVO_OK) return _s;           \
  } while (0)
