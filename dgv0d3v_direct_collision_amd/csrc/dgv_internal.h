// dgv_internal.h -- internal types of the MI355X direct DGV collision library.
//
// Device layout of the A-array ("pair-block" layout)
// ---------------------------------------------------
// The reference stores, per canonical basis function l, a list of records
// (A, xi, xi1) (commvar.f90:157-167) and the evaluator re-reads the whole list
// once per node (collision_mod.f90:186-227).  Every record of every l belongs to
// an unordered node pair (xi, xi1); the geometric prefilter that decides which
// pairs exist (dgvtools_mod.f90:493-530) does not depend on l.  On the device the
// records are therefore regrouped by pair:
//
//   pair   = (T[0..D-1], code)     T[l] = A of basis l (0 when l has no record),
//                                   code = p << 8 | q, p / q = local index of xi / xi1
//                                   (low word of the slot; the high word holds the
//                                   same pair as LDS byte offsets, pair_code_bits);
//   block  = all pairs whose xi lies in cell (cc + a) and xi1 in cell (cc + b),
//            sorted by p = local index of xi; seg[p] = first pair with index >= p;
//   group  = all blocks sharing the same box of output cells c for which both
//            c + a and c + b stay on the grid (the evaluator's bounds tests,
//            collision_mod.f90:196-211, hold for exactly that box).
//
// One Q evaluation then computes, for every output cell c and basis l,
//   Q[c,l] = (2/w) sum_blocks sum_pairs T[l] * f[c+a][p] * f[c+b][q],
// sharing the product f*f of a pair between its D basis functions.
#pragma once
#include <hip/hip_runtime.h>
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dgv.h"

namespace dgv {

constexpr int kMaxS = 8;
// Zero pairs after every pair-record array on the device: the evaluators' load
// rings read up to 4 * 8 + 3 pairs past a block without clamping.
constexpr int kPairPad = 64;

// The 64-bit code slot of a pair record.  Low word: p << 8 | q.  High word: the
// byte offsets of rows p and q in the FP64 evaluators' LDS staging (64 doubles
// per row), p * 512 in bits 0..15 and q * 512 in bits 16..31, so that a kernel
// forms both operand addresses with one add each.
__host__ __device__ inline int64_t pair_code_bits(int p, int q) {
  const uint64_t lo = (uint32_t)((p << 8) | q);
  const uint64_t hi = (uint32_t)(p * 512) | ((uint32_t)(q * 512) << 16);
  return (int64_t)((hi << 32) | lo);
}

// The evaluators' coefficient rings read a block's pair records through a
// buffer resource based at the block: the step is a scalar offset and the
// lane's part a fixed 32-bit offset, so a load costs no 64-bit address
// arithmetic on the VALU.  The resource does no range check (maximum record
// count); the kPairPad zero records after every pair array keep the rings'
// reads, at most 4 * 8 + 3 pairs past a block, inside the allocation.  The
// base must be the same in every lane of the wave (a block of the workgroup's
// item); it is read from the first lane, so that the resource lives in scalar
// registers even where the compiler cannot prove the base uniform.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t ring_rsrc(const void* block_base) {
  const uint64_t a = reinterpret_cast<uint64_t>(block_base);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a >> 32));
  void* base = reinterpret_cast<void*>(((uint64_t)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, -1, 0x00020000);
}
__device__ __forceinline__ double ring_f64(__amdgpu_buffer_rsrc_t r, uint32_t lane_off, uint32_t step_off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, lane_off, step_off, 0));
}
__device__ __forceinline__ float ring_f32(__amdgpu_buffer_rsrc_t r, uint32_t lane_off, uint32_t step_off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, lane_off, step_off, 0));
}
__device__ __forceinline__ int ring_i32(__amdgpu_buffer_rsrc_t r, uint32_t lane_off, uint32_t step_off) {
  return (int)__builtin_amdgcn_raw_buffer_load_b32(r, lane_off, step_off, 0);
}
__device__ __forceinline__ void ring_i32x2(__amdgpu_buffer_rsrc_t r, uint32_t lane_off, uint32_t step_off, int& a,
                                           int& b) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, lane_off, step_off, 0);
  a = (int)v[0];
  b = (int)v[1];
}

// Blocks for a grid-stride loop over n items. HIP limits one grid dimension to
// fewer than 2^32 work-items, and the coefficient arrays of a 33^3 grid already
// hold more elements than that, so element-wise kernels stride over a capped grid.
inline unsigned stride_blocks(int64_t n, int block) {
  int64_t b = (n + block - 1) / block;
  if (b > (1 << 20)) b = 1 << 20;
  return (unsigned)(b < 1 ? 1 : b);
}

struct Grid {
  int Mu = 0, Mv = 0, Mw = 0, su = 0, sv = 0, sw = 0, D = 0;
  double box[6] = {0, 0, 0, 0, 0, 0};
  int64_t N = 0, ncells = 0;
  int ccell = 0;  // 0-based canonical cell
  int ccu = 0, ccv = 0, ccw = 0;
  double gn[3][kMaxS] = {}, gw[3][kMaxS] = {};  // Gauss nodes/weights per dimension
  std::vector<double> u, v, w, gwts;            // node coordinates and weights
  std::vector<double> lu, ru, lv, rv, lw, rw;   // cell bounds
};

// Block header, read with scalar loads by the evaluator (uniform per wave).
struct alignas(16) BlockHdr {
  int64_t pbeg;        // first pair of the block
  int8_t a[3], b[3];   // cell offsets of xi and xi1 from the canonical cell
  int16_t npairs;
  int32_t seg[65];     // seg[p] relative to pbeg, p = 0..D (D <= 64)
};

struct Group {
  int32_t lo[3], ext[3];  // box of output cells
  int32_t ncells;
  int32_t bbeg, bend;     // range in the group's block-index list
  int64_t npairs;
};

struct WorkItem {
  int32_t group, chunk;   // 64-cell chunk of the group's box (chunk < 0: empty slot)
  int32_t bbeg, bend;     // blocks (indices into group_blocks)
  int64_t prow;           // first partial row (64 rows per item)
};

// Host-side description of the pair-block layout (built once per A-array).
struct PairBlock {
  int D = 0;
  int64_t npairs = 0;
  std::vector<BlockHdr> blocks;       // sorted by group
  std::vector<Group> groups;
  std::vector<int32_t> group_blocks;  // block indices grouped by box
  // device copies
  double* d_pairs = nullptr;          // npairs * (D+1) doubles: T[0..D-1], code bits
  BlockHdr* d_blocks = nullptr;
  int32_t* d_group_blocks = nullptr;
  Group* d_groups = nullptr;
  int64_t nrec_ref = 0;               // record count in the reference layout
  std::vector<int64_t> nrec_basis;    // per canonical basis function
};

// Work decomposition of one evaluation (depends on tile size).
struct Plan {
  int32_t pairs_per_item = 0;
  std::vector<WorkItem> items;
  int64_t nrows = 0;                  // partial rows = 64 * items
  WorkItem* d_items = nullptr;
  double* d_partial = nullptr;        // nrows * D doubles (FP64 partial sums)
  int64_t* d_cell_ptr = nullptr;      // CSR over output cells of partial rows
  int32_t* d_cell_rows = nullptr;
  int64_t ncell_rows = 0;
};

struct FmaItem;

// Decomposition / linearised regimes (PrepareFMA_DGV, EvalCollisionLinear,
// CheckSolutionMode_DGV).  Built on first use, dropped with the A-array.
struct LinearState {
  double* d_fma = nullptr;        // N x N, fmA(r, i) at r * N + i
  int16_t* d_pidx = nullptr;      // per block: pair index of (p, q) or -1
  FmaItem* d_items = nullptr;
  int32_t* d_l1 = nullptr;        // block lists by a-offset / b-offset
  int32_t* d_l2 = nullptr;
  double* d_part = nullptr;       // GEMV partial sums
  int64_t nitems = 0;
  bool fma_valid = false;
  double* d_vec[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // fm, new Maxwellian, Df, scratch, reduction
  int32_t run_mode = 0;
  double decomp_lev = 0.0, linear_lev = 0.0;
  double last_l1 = 0.0, last_l1_fm = 0.0;
  int64_t nchecks = 0, nprepare = 0;
};

}  // namespace dgv

struct dgv_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  dgv::Grid g;
  // device geometry
  double* d_gwts = nullptr;
  double* d_u = nullptr;
  double* d_v = nullptr;
  double* d_w = nullptr;
  // A-array.  a_gen changes with every change of the A-array (new, read,
  // truncated, |A|): batches and surrogates built on the context keep the
  // value they were built with and refuse to run once it differs.
  bool has_a = false;
  uint64_t a_gen = 0;
  dgv::PairBlock pb;
  dgv::Plan plan;
  int32_t tile = 0;          // pairs per work item (0 = default)
  int kernel = DGV_KERNEL_AUTO;
  int project = 0;           // project every Q onto the complement of the invariants
  // reference layout on the device for the record kernel (built lazily)
  bool has_rec = false;
  int64_t* d_rec_off = nullptr;   // D+1 offsets of the canonical basis functions
  double* d_rec_A = nullptr;
  int32_t* d_rec_xi = nullptr;    // 0-based node ids
  int32_t* d_rec_xi1 = nullptr;
  // scratch
  double* d_f = nullptr;
  double* d_q = nullptr;
  double* d_tmp[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double* d_red = nullptr;        // moment reduction scratch
  double* d_hi = nullptr;         // high-moment reduction scratch (on first use)
  hipGraphExec_t rk_graph = nullptr;
  int rk_graph_key[2] = {0, 0};
  double rk_graph_dt = 0, rk_graph_coef = 0;
  double* rk_graph_f = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  dgv::LinearState lin;
  // host copy of the reference layout, built on first export
  bool exp_valid = false;
  std::vector<std::vector<std::array<int64_t, 2>>> exp_idx;
  std::vector<std::vector<double>> exp_val;
  std::string err;
};

namespace dgv {

int set_error(dgv_ctx* ctx, int code, const std::string& msg);
// gives the context's A-array a new generation number (unique in the process)
void bump_aarray_generation(dgv_ctx* ctx);
void set_global_error(const std::string& msg);

#define DGV_HIP(ctx, call)                                                        \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::dgv::set_error((ctx), DGV_ERR_HIP,                                 \
                              std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// geometry (dgv_host.cpp)
int build_grid(const dgv_grid_desc& d, Grid& g, std::string& err);
int gauss_legendre(int n, double* x, double* w);

// layout (dgv_host.cpp)
int build_pairblock_from_records(dgv_ctx* ctx, const int32_t* capphi, int64_t ncap,
                                 const double* A, const int32_t* xi, const int32_t* xi1,
                                 int64_t nrec, int index_base);
int upload_pairblock(dgv_ctx* ctx, std::vector<double>& pairs);
int build_plan(dgv_ctx* ctx, int32_t pairs_per_item);
void free_pairblock(dgv_ctx* ctx);
void free_plan(dgv_ctx* ctx);
int ensure_record_layout(dgv_ctx* ctx);

// kernels (dgv_kernels.hip)
int launch_eval_pairblock(dgv_ctx* ctx, const double* d_f, double* d_q, double scale, bool mfma);
int launch_eval_record(dgv_ctx* ctx, const double* d_f, double* d_q, double scale);
int eval_q_dev(dgv_ctx* ctx, const double* d_f, double* d_q, double scale);
int eval_mixed_dev(dgv_ctx* ctx, const double* d_f, const double* d_fm, double* d_q);
int launch_copy16(dgv_ctx* ctx, void* dst, const void* src, size_t bytes);
int launch_axpy(dgv_ctx* ctx, int64_t n, double* out, const double* x0, int nterms,
                const double* const* xs, const double* coefs);
int launch_scale(dgv_ctx* ctx, int64_t n, double* x, double s);
int device_moments(dgv_ctx* ctx, const double* d_f, double out[5]);
int device_collision_moments(dgv_ctx* ctx, const double* d_q, double out[5]);
int device_moments_high(dgv_ctx* ctx, const double* d_f, double out[20]);
int device_project(dgv_ctx* ctx, double* d_q, double* scratch = nullptr);
// the file-scope device scratch of dgv_eval_q_dev_noalloc, looked up once per
// device at context creation (dgv_kernels.hip)
int static_scratch_setup(int device);
// Cholesky factor (row-major, lower) of the Gram matrix of psi = 1,u,v,w,|u|^2
// in the w-weighted inner product (dgv_capi.cpp)
int gram_cholesky(const Grid& g, double L[25]);

// decomposition / linearised regimes (dgv_linear.hip)
void free_linear(dgv_ctx* ctx);
void free_linear_vectors(dgv_ctx* ctx);
int prepare_fma_dev(dgv_ctx* ctx, const double* d_fm);
int eval_linear_dev(dgv_ctx* ctx, const double* d_f, int64_t ldf, double* d_fc, int64_t ldq, int nvec);
int check_solution_mode(dgv_ctx* ctx, const double* d_f);
int spatial_oper_device(dgv_ctx* ctx, const double* d_f, double* d_out, double scale);

// true when a grid's A-array changed after the batch was created (dgv_batch.hip)
bool batch_stale(const dgv_batch* b);

// generator (dgv_agen.hip)
int generate_pairblock(dgv_ctx* ctx, const dgv_agen_params& p);

}  // namespace dgv
