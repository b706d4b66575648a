// dgv_linear.hip -- the decomposition and linearised regimes on the device.
//
//   PrepareFMA_DGV        collision_mod.f90:377-464   dense fmA = (1/w_i) sum A fm  (N x N)
//   EvalCollisionLinear   collision_mod.f90:346-359   fc = 2 f . fmA (GEMV, HBM bound)
//   CheckSolutionMode_DGV spat_oper_mod.f90:195-272   regime switch on |f - f_M|_1
//   SpatialOperSH_DGV     spat_oper_mod.f90:28-68     run_mode 0 / 1 / 2
//
// fmA lives in HBM row-major over its first index: element fmA(r, i) (row r =
// the node the record's f_M factor does NOT multiply, column i = the output
// node) is at d_fma[r * N + i].  Column i of the reference's array is the
// stride-N sequence; the GEMV reads rows (contiguous in i) and the build writes
// each lane's 27 columns of one cell contiguously.
//
// Build, deterministic and without atomics: the entries of the tile (column
// cell c, row cell c + e) receive
//   term 1: sum over blocks with a = e, pairs (s, q):  T[l] fm[c+b][q]
//   term 2: sum over blocks with b = e, pairs (p, s):  T[l] fm[c+a][p]
// (the reference's two scatter-adds per record, lines 447-448), then the
// division by w_i (line 458).  A work item owns (e, s, 64 column cells) and
// writes its entries once.  A per-block table idx[p*D+q] (pair index or -1)
// lets term 2 walk a block by q as term 1 walks it by p.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "dgv_internal.h"

namespace dgv {

struct FmaItem {
  int32_t e[3];
  int32_t s;
  int32_t lo[3], ext[3];
  int32_t ncells, chunk;
  int32_t l1b, l1e, l2b, l2e;  // ranges in the block lists with a = e and b = e
};

__global__ void k_pair_index(int D, const BlockHdr* __restrict__ blocks, const double* __restrict__ pairs,
                             int16_t* __restrict__ idx) {
  __shared__ int16_t t[64 * 64];
  const BlockHdr* B = blocks + blockIdx.x;
  const int DD = D * D;
  for (int i = threadIdx.x; i < DD; i += blockDim.x) t[i] = -1;
  __syncthreads();
  const int n = B->seg[D];
  const double* P = pairs + B->pbeg * (D + 1);
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int code = (int)(__double_as_longlong(P[(int64_t)k * (D + 1) + D]) & 0xffff);
    t[(code >> 8) * D + (code & 255)] = (int16_t)k;
  }
  __syncthreads();
  int16_t* out = idx + (int64_t)blockIdx.x * DD;
  for (int i = threadIdx.x; i < DD; i += blockDim.x) out[i] = t[i];
}

// One wave per item; lane = column cell.  The coefficients and the index
// table entries are wave-uniform (scalar loads); fm of the partner cell is the
// lane's own (0 where that cell is off the grid: the record is then skipped,
// as the reference's bounds tests skip it).
template <int D>
__global__ __launch_bounds__(64) void k_fma_build(const FmaItem* __restrict__ items, const int32_t* __restrict__ l1,
                                                  const int32_t* __restrict__ l2, const BlockHdr* __restrict__ blocks,
                                                  const double* __restrict__ pairs, const int16_t* __restrict__ pidx,
                                                  const double* __restrict__ fm, const double* __restrict__ gwts,
                                                  int Mu, int Mv, int Mw, int64_t N, double* __restrict__ out) {
  const FmaItem it = items[blockIdx.x];
  const int lane = threadIdx.x;
  const int cl = it.chunk * 64 + lane;
  const bool active = cl < it.ncells;
  int r = active ? cl : 0;
  const int cw = it.lo[2] + r % it.ext[2];
  r /= it.ext[2];
  const int cv = it.lo[1] + r % it.ext[1];
  r /= it.ext[1];
  const int cu = it.lo[0] + r;
  const int s = it.s;
  double acc[D];
#pragma unroll
  for (int l = 0; l < D; ++l) acc[l] = 0.0;

  for (int k = it.l1b; k < it.l1e; ++k) {
    const int bi = l1[k];
    const BlockHdr* B = blocks + bi;
    const int yu = cu + B->b[0], yv = cv + B->b[1], yw = cw + B->b[2];
    const bool ok = yu >= 0 && yu < Mu && yv >= 0 && yv < Mv && yw >= 0 && yw < Mw;
    const double* fy = fm + (int64_t)(ok ? (yu * Mv + yv) * Mw + yw : 0) * D;
    const int16_t* ix = pidx + (int64_t)bi * D * D + s * D;
    const double* P = pairs + B->pbeg * (D + 1);
    for (int q = 0; q < D; ++q) {
      const int j = ix[q];
      if (j < 0) continue;
      const double x = ok ? fy[q] : 0.0;
      const double* T = P + (int64_t)j * (D + 1);
#pragma unroll
      for (int l = 0; l < D; ++l) acc[l] = __fma_rn(T[l], x, acc[l]);
    }
  }
  for (int k = it.l2b; k < it.l2e; ++k) {
    const int bi = l2[k];
    const BlockHdr* B = blocks + bi;
    const int xu = cu + B->a[0], xv = cv + B->a[1], xw = cw + B->a[2];
    const bool ok = xu >= 0 && xu < Mu && xv >= 0 && xv < Mv && xw >= 0 && xw < Mw;
    const double* fx = fm + (int64_t)(ok ? (xu * Mv + xv) * Mw + xw : 0) * D;
    const int16_t* ix = pidx + (int64_t)bi * D * D + s;
    const double* P = pairs + B->pbeg * (D + 1);
    for (int p = 0; p < D; ++p) {
      const int j = ix[p * D];
      if (j < 0) continue;
      const double x = ok ? fx[p] : 0.0;
      const double* T = P + (int64_t)j * (D + 1);
#pragma unroll
      for (int l = 0; l < D; ++l) acc[l] = __fma_rn(T[l], x, acc[l]);
    }
  }
  if (active) {
    const int64_t row = (int64_t)(((cu + it.e[0]) * Mv + (cv + it.e[1])) * Mw + (cw + it.e[2])) * D + s;
    const int64_t col = (int64_t)((cu * Mv + cv) * Mw + cw) * D;
    double* o = out + row * N + col;
#pragma unroll
    for (int l = 0; l < D; ++l) o[l] = acc[l] / gwts[col + l];
  }
}

// fc = 2 f . fmA for nv vectors: split over row chunks of kGemvRows, FP64
// partial sums per chunk, then a fixed-order sum of the chunks.
constexpr int kGemvRows = 256;
constexpr int kGemvMaxVec = 8;

template <int NV>
__global__ __launch_bounds__(256) void k_fma_gemv(int64_t N, const double* __restrict__ fma,
                                                  const double* __restrict__ f, int64_t ldf,
                                                  double* __restrict__ part) {
  constexpr int kU = 16;  // row loads in flight per thread
  __shared__ double fs[NV][kGemvRows];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.y * kGemvRows;
  const int nr = (int)(r0 + kGemvRows < N ? kGemvRows : N - r0);
  for (int t = threadIdx.x; t < NV * kGemvRows; t += 256) {
    const int v = t / kGemvRows, rr = t % kGemvRows;
    fs[v][rr] = rr < nr ? f[v * ldf + r0 + rr] : 0.0;
  }
  __syncthreads();
  const double* col = fma + r0 * N + (i < N ? i : N - 1);
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  int r = 0;
  for (; r + kU <= nr; r += kU) {
    double m[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) m[u] = col[(int64_t)(r + u) * N];
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[v] = __fma_rn(fs[v][r + u], m[u], acc[v]);
  }
  for (; r < nr; ++r) {
    const double m = col[(int64_t)r * N];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = __fma_rn(fs[v][r], m, acc[v]);
  }
  if (i < N) {
#pragma unroll
    for (int v = 0; v < NV; ++v) part[((int64_t)v * gridDim.y + blockIdx.y) * N + i] = acc[v];
  }
}

__global__ void k_fma_gemv_sum(int64_t N, int nchunk, int nv, const double* __restrict__ part,
                               double* __restrict__ fc, int64_t ldq) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int v = 0; v < nv; ++v) {
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[((int64_t)v * nchunk + c) * N + i];
    fc[v * ldq + i] = 2.0 * s;
  }
}

// maxwelveldist (distributions_mod.f90:83-94) at every node
__global__ void k_maxwell(int64_t n, double T, double u0, double v0, double w0, double dens,
                          const double* __restrict__ u, const double* __restrict__ v,
                          const double* __restrict__ w, double* __restrict__ out) {
  const double pi = 3.14159265358979323846264338327950288;
  const double beta = sqrt(pi * T) * (pi * T);
  const double tt = T > 0.0000001 ? T : 0.0000001;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
    const double du = u[i] - u0, dv = v[i] - v0, dw = w[i] - w0;
    out[i] = dens * exp(-(du * du + dv * dv + dw * dw) / tt) / beta;
  }
}

// sum |a - b| w, fixed-order two-level reduction (256 blocks, then one thread)
constexpr int kL1Blocks = 256;

__global__ __launch_bounds__(256) void k_l1_part(int64_t n, const double* __restrict__ a,
                                                 const double* __restrict__ b, const double* __restrict__ w,
                                                 double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)kL1Blocks * 256)
    s += fabs(a[i] - b[i]) * w[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_l1_final(const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < kL1Blocks; ++b) s += part[b];
  out[0] = s;
}

__global__ void k_sub(int64_t n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) out[i] = a[i] - b[i];
}

// out = (x + y) * scale, or x * scale when y is null
__global__ void k_sum_scale(int64_t n, const double* __restrict__ x, const double* __restrict__ y, double scale,
                            double* __restrict__ out) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step)
    out[i] = (y ? x[i] + y[i] : x[i]) * scale;
}

// ---------------------------------------------------------------------------

void free_linear(dgv_ctx* ctx) {
  LinearState& L = ctx->lin;
  hipFree(L.d_fma); hipFree(L.d_pidx); hipFree(L.d_items); hipFree(L.d_l1); hipFree(L.d_l2);
  hipFree(L.d_part);
  L.d_fma = nullptr; L.d_pidx = nullptr; L.d_items = nullptr; L.d_l1 = nullptr; L.d_l2 = nullptr;
  L.d_part = nullptr;
  L.nitems = 0;
  L.fma_valid = false;
  L.run_mode = 0;
}

void free_linear_vectors(dgv_ctx* ctx) {
  LinearState& L = ctx->lin;
  for (int k = 0; k < 5; ++k) { hipFree(L.d_vec[k]); L.d_vec[k] = nullptr; }
}

static int ensure_vectors(dgv_ctx* ctx) {
  LinearState& L = ctx->lin;
  if (L.d_vec[0]) return DGV_OK;
  for (int k = 0; k < 5; ++k) {
    const size_t nb = sizeof(double) * (k == 4 ? 2 * kL1Blocks : std::max<int64_t>(ctx->g.N, 1));
    if (hipMalloc(&L.d_vec[k], nb) != hipSuccess) {
      free_linear_vectors(ctx);
      return set_error(ctx, DGV_ERR_NOMEM, "hipMalloc(linearised-regime vectors) failed");
    }
  }
  return DGV_OK;
}

// Block index tables, block lists by offset and the work items (once per A-array).
static int ensure_fma_plan(dgv_ctx* ctx) {
  LinearState& L = ctx->lin;
  if (L.d_items) return DGV_OK;
  const Grid& g = ctx->g;
  const PairBlock& pb = ctx->pb;
  const int D = g.D;
  if (D > 64) return set_error(ctx, DGV_ERR_UNSUPPORTED, "PrepareFMA: more than 64 nodes per cell");
  const int64_t nb = (int64_t)pb.blocks.size();
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  if (hipMalloc(&L.d_pidx, sizeof(int16_t) * (size_t)std::max<int64_t>(nb, 1) * D * D) != hipSuccess)
    return set_error(ctx, DGV_ERR_NOMEM, "hipMalloc(pair index tables) failed");
  if (nb > 0)
    hipLaunchKernelGGL(k_pair_index, dim3((unsigned)nb), dim3(256), 0, ctx->stream, D, pb.d_blocks, pb.d_pairs,
                       L.d_pidx);
  DGV_HIP(ctx, hipGetLastError());
  // block lists keyed by the offset (a for term 1, b for term 2)
  auto key = [](const int8_t* e) { return ((e[0] + 128) << 16) | ((e[1] + 128) << 8) | (e[2] + 128); };
  std::map<int, std::pair<std::vector<int32_t>, std::vector<int32_t>>> lists;
  for (int64_t bi = 0; bi < nb; ++bi) {
    const BlockHdr& B = pb.blocks[bi];
    if (B.seg[D] == 0) continue;
    lists[key(B.a)].first.push_back((int32_t)bi);
    lists[key(B.b)].second.push_back((int32_t)bi);
  }
  std::vector<int32_t> l1, l2;
  std::vector<FmaItem> items;
  const int M[3] = {g.Mu, g.Mv, g.Mw};
  for (auto& kv : lists) {
    const int e[3] = {((kv.first >> 16) & 255) - 128, ((kv.first >> 8) & 255) - 128, (kv.first & 255) - 128};
    FmaItem it{};
    int ncells = 1;
    for (int d = 0; d < 3; ++d) {
      const int lo = std::max(0, -e[d]), hi = std::min(M[d], M[d] - e[d]);
      it.e[d] = e[d];
      it.lo[d] = lo;
      it.ext[d] = std::max(0, hi - lo);
      ncells *= it.ext[d];
    }
    if (ncells == 0) continue;
    it.ncells = ncells;
    it.l1b = (int32_t)l1.size();
    l1.insert(l1.end(), kv.second.first.begin(), kv.second.first.end());
    it.l1e = (int32_t)l1.size();
    it.l2b = (int32_t)l2.size();
    l2.insert(l2.end(), kv.second.second.begin(), kv.second.second.end());
    it.l2e = (int32_t)l2.size();
    for (int s = 0; s < D; ++s)
      for (int c = 0; c * 64 < ncells; ++c) {
        it.s = s;
        it.chunk = c;
        items.push_back(it);
      }
  }
  L.nitems = (int64_t)items.size();
  if (hipMalloc(&L.d_items, sizeof(FmaItem) * std::max<size_t>(items.size(), 1)) != hipSuccess ||
      hipMalloc(&L.d_l1, sizeof(int32_t) * std::max<size_t>(l1.size(), 1)) != hipSuccess ||
      hipMalloc(&L.d_l2, sizeof(int32_t) * std::max<size_t>(l2.size(), 1)) != hipSuccess)
    return set_error(ctx, DGV_ERR_NOMEM, "hipMalloc(fmA plan) failed");
  DGV_HIP(ctx, hipMemcpyAsync(L.d_items, items.data(), sizeof(FmaItem) * items.size(), hipMemcpyHostToDevice,
                              ctx->stream));
  DGV_HIP(ctx, hipMemcpyAsync(L.d_l1, l1.data(), sizeof(int32_t) * l1.size(), hipMemcpyHostToDevice, ctx->stream));
  DGV_HIP(ctx, hipMemcpyAsync(L.d_l2, l2.data(), sizeof(int32_t) * l2.size(), hipMemcpyHostToDevice, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

int prepare_fma_dev(dgv_ctx* ctx, const double* d_fm) {
  if (!ctx->has_a) return set_error(ctx, DGV_ERR_STATE, "no A-array set");
  LinearState& L = ctx->lin;
  const Grid& g = ctx->g;
  const int D = g.D;
  if (D != 8 && D != 27) return set_error(ctx, DGV_ERR_UNSUPPORTED, "PrepareFMA: 2 or 3 nodes per dimension only");
  int rc = ensure_fma_plan(ctx);
  if (rc) return rc;
  rc = ensure_vectors(ctx);
  if (rc) return rc;
  const int64_t N = g.N;
  if (!L.d_fma && hipMalloc(&L.d_fma, sizeof(double) * (size_t)N * N) != hipSuccess) {
    L.d_fma = nullptr;
    return set_error(ctx, DGV_ERR_NOMEM, "hipMalloc(fmA, N x N doubles) failed");
  }
  if (d_fm != L.d_vec[0])
    DGV_HIP(ctx, hipMemcpyAsync(L.d_vec[0], d_fm, sizeof(double) * N, hipMemcpyDeviceToDevice, ctx->stream));
  DGV_HIP(ctx, hipMemsetAsync(L.d_fma, 0, sizeof(double) * (size_t)N * N, ctx->stream));
  if (L.nitems > 0) {
    if (D == 27)
      hipLaunchKernelGGL(k_fma_build<27>, dim3((unsigned)L.nitems), dim3(64), 0, ctx->stream, L.d_items, L.d_l1,
                         L.d_l2, ctx->pb.d_blocks, ctx->pb.d_pairs, L.d_pidx, L.d_vec[0], ctx->d_gwts, g.Mu, g.Mv,
                         g.Mw, N, L.d_fma);
    else
      hipLaunchKernelGGL(k_fma_build<8>, dim3((unsigned)L.nitems), dim3(64), 0, ctx->stream, L.d_items, L.d_l1,
                         L.d_l2, ctx->pb.d_blocks, ctx->pb.d_pairs, L.d_pidx, L.d_vec[0], ctx->d_gwts, g.Mu, g.Mv,
                         g.Mw, N, L.d_fma);
  }
  DGV_HIP(ctx, hipGetLastError());
  L.fma_valid = true;
  L.nprepare++;
  return DGV_OK;
}

int eval_linear_dev(dgv_ctx* ctx, const double* d_f, int64_t ldf, double* d_fc, int64_t ldq, int nvec) {
  LinearState& L = ctx->lin;
  if (!L.fma_valid) return set_error(ctx, DGV_ERR_STATE, "EvalCollisionLinear: no linearised operator (PrepareFMA)");
  const int64_t N = ctx->g.N;
  const int nchunk = (int)((N + kGemvRows - 1) / kGemvRows);
  if (!L.d_part &&
      hipMalloc(&L.d_part, sizeof(double) * (size_t)kGemvMaxVec * nchunk * std::max<int64_t>(N, 1)) != hipSuccess) {
    L.d_part = nullptr;
    return set_error(ctx, DGV_ERR_NOMEM, "hipMalloc(GEMV partials) failed");
  }
  const dim3 grid((unsigned)((N + 255) / 256), (unsigned)nchunk);
  for (int v0 = 0; v0 < nvec; v0 += kGemvMaxVec) {
    const int nv = std::min(kGemvMaxVec, nvec - v0);
    const double* f = d_f + v0 * ldf;
    switch (nv) {
#define DGV_GEMV_CASE(K) \
  case K: hipLaunchKernelGGL(k_fma_gemv<K>, grid, dim3(256), 0, ctx->stream, N, L.d_fma, f, ldf, L.d_part); break;
      DGV_GEMV_CASE(1) DGV_GEMV_CASE(2) DGV_GEMV_CASE(3) DGV_GEMV_CASE(4)
      DGV_GEMV_CASE(5) DGV_GEMV_CASE(6) DGV_GEMV_CASE(7) DGV_GEMV_CASE(8)
#undef DGV_GEMV_CASE
    }
    hipLaunchKernelGGL(k_fma_gemv_sum, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream, N, nchunk, nv,
                       L.d_part, d_fc + v0 * ldq, ldq);
  }
  DGV_HIP(ctx, hipGetLastError());
  return DGV_OK;
}

static int l1_weighted(dgv_ctx* ctx, const double* a, const double* b, double* out) {
  LinearState& L = ctx->lin;
  hipLaunchKernelGGL(k_l1_part, dim3(kL1Blocks), dim3(256), 0, ctx->stream, ctx->g.N, a, b, ctx->d_gwts, L.d_vec[4]);
  hipLaunchKernelGGL(k_l1_final, dim3(1), dim3(1), 0, ctx->stream, L.d_vec[4], L.d_vec[4] + kL1Blocks);
  DGV_HIP(ctx, hipGetLastError());
  DGV_HIP(ctx, hipMemcpyAsync(out, L.d_vec[4] + kL1Blocks, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

// CheckSolutionMode_DGV (spat_oper_mod.f90:195-272).  Vectors: d_vec[0] = fm
// (the Maxwellian fmA was built at), d_vec[1] = the new Maxwellian, d_vec[2] =
// Df = f - fm, d_vec[3] = scratch of the spatial operator.
int check_solution_mode(dgv_ctx* ctx, const double* d_f) {
  LinearState& L = ctx->lin;
  int rc = ensure_vectors(ctx);
  if (rc) return rc;
  const Grid& g = ctx->g;
  double mom[5];
  rc = device_moments(ctx, d_f, mom);  // MassCheckRec, line 236
  if (rc) return rc;
  const double dens = mom[0];
  hipLaunchKernelGGL(k_maxwell, dim3(stride_blocks(g.N, 256)), dim3(256), 0, ctx->stream, g.N, mom[4], mom[1], mom[2],
                     mom[3], dens, ctx->d_u, ctx->d_v, ctx->d_w, L.d_vec[1]);
  DGV_HIP(ctx, hipGetLastError());
  double s = 0.0;
  rc = l1_weighted(ctx, d_f, L.d_vec[1], &s);
  if (rc) return rc;
  const double l1 = s / dens;  // line 238
  L.last_l1 = l1;
  L.nchecks++;
  auto refresh_fm = [&]() -> int {  // fm = fMaxwellNew; PrepareFMA_DGV
    DGV_HIP(ctx, hipMemcpyAsync(L.d_vec[0], L.d_vec[1], sizeof(double) * g.N, hipMemcpyDeviceToDevice, ctx->stream));
    return prepare_fma_dev(ctx, L.d_vec[0]);
  };
  auto l1_fm = [&](double* v) -> int {
    double t = 0.0;
    int r = l1_weighted(ctx, L.d_vec[0], L.d_vec[1], &t);
    *v = t / dens;
    return r;
  };
  if (l1 < L.decomp_lev) {
    switch (L.run_mode) {
      case 0:
        if ((rc = refresh_fm())) return rc;
        L.run_mode = l1 < L.linear_lev ? 2 : 1;
        break;
      case 1: {
        double e = 0.0;
        if ((rc = l1_fm(&e))) return rc;
        L.last_l1_fm = e;
        if (e > L.linear_lev / 2.0 && (rc = refresh_fm())) return rc;
        if (l1 < L.linear_lev) L.run_mode = 2;
        break;
      }
      case 2: {
        double e = 0.0;
        if ((rc = l1_fm(&e))) return rc;
        L.last_l1_fm = e;
        if (e > L.linear_lev / 2.0 && (rc = refresh_fm())) return rc;
        if (l1 > L.linear_lev) L.run_mode = 1;
        break;
      }
      default:
        return set_error(ctx, DGV_ERR_STATE, "CheckSolutionMode_DGV: cannot process the value of run_mode");
    }
    hipLaunchKernelGGL(k_sub, dim3(stride_blocks(g.N, 256)), dim3(256), 0, ctx->stream, g.N, d_f, L.d_vec[0],
                       L.d_vec[2]);  // Df = f - fm
    DGV_HIP(ctx, hipGetLastError());
  } else {
    L.run_mode = 0;
  }
  return DGV_OK;
}

// SpatialOperSH_DGV (spat_oper_mod.f90:28-68): the regime check (when switching
// is enabled), then run_mode 0: Q(f); 1: L(Df) + Q(Df); 2: L(Df); times scale.
int spatial_oper_device(dgv_ctx* ctx, const double* d_f, double* d_out, double scale) {
  LinearState& L = ctx->lin;
  if (!(L.decomp_lev > 0.0)) {
    L.run_mode = 0;
    return eval_q_dev(ctx, d_f, d_out, scale);
  }
  int rc = check_solution_mode(ctx, d_f);
  if (rc) return rc;
  const int64_t N = ctx->g.N;
  const unsigned nblk = stride_blocks(N, 256);
  switch (L.run_mode) {
    case 0:
      return eval_q_dev(ctx, d_f, d_out, scale);
    case 1: {
      const int save = ctx->project;
      ctx->project = 0;
      rc = eval_linear_dev(ctx, L.d_vec[2], N, L.d_vec[3], N, 1);
      if (!rc) rc = eval_q_dev(ctx, L.d_vec[2], d_out, 1.0);
      ctx->project = save;
      if (rc) return rc;
      hipLaunchKernelGGL(k_sum_scale, dim3(nblk), dim3(256), 0, ctx->stream, N, d_out, L.d_vec[3], scale, d_out);
      break;
    }
    case 2:
      rc = eval_linear_dev(ctx, L.d_vec[2], N, d_out, N, 1);
      if (rc) return rc;
      hipLaunchKernelGGL(k_sum_scale, dim3(nblk), dim3(256), 0, ctx->stream, N, d_out, (const double*)nullptr, scale,
                         d_out);
      break;
    default:
      return set_error(ctx, DGV_ERR_STATE, "SpatialOperSH_DGV: cannot process the value of run_mode");
  }
  DGV_HIP(ctx, hipGetLastError());
  if (ctx->project) return device_project(ctx, d_out);
  return DGV_OK;
}

}  // namespace dgv

using namespace dgv;

#define LIN_CHECK(ctx, n)                                                            \
  do {                                                                               \
    if (!(ctx)) return DGV_ERR_ARG;                                                  \
    if ((n) != (ctx)->g.N) return set_error((ctx), DGV_ERR_ARG, "array size mismatch"); \
  } while (0)

extern "C" {

int dgv_prepare_fma_dev(dgv_ctx* ctx, const double* d_fm, int64_t n) {
  LIN_CHECK(ctx, n);
  if (!d_fm) return set_error(ctx, DGV_ERR_ARG, "null array");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  int rc = ensure_vectors(ctx);
  if (rc) return rc;
  return prepare_fma_dev(ctx, d_fm);
}

int dgv_prepare_fma(dgv_ctx* ctx, const double* fm, int64_t n) {
  LIN_CHECK(ctx, n);
  if (!fm) return set_error(ctx, DGV_ERR_ARG, "null array");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  int rc = ensure_vectors(ctx);
  if (rc) return rc;
  DGV_HIP(ctx, hipMemcpyAsync(ctx->lin.d_vec[0], fm, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  rc = prepare_fma_dev(ctx, ctx->lin.d_vec[0]);
  if (rc) return rc;
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

int dgv_eval_linear_dev(dgv_ctx* ctx, const double* d_f, int64_t ldf, double* d_fc, int64_t ldq, int64_t n,
                        int32_t nvec) {
  LIN_CHECK(ctx, n);
  if (!d_f || !d_fc || nvec < 1 || ldf < n || ldq < n) return set_error(ctx, DGV_ERR_ARG, "bad arguments");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  return eval_linear_dev(ctx, d_f, ldf, d_fc, ldq, nvec);
}

int dgv_eval_linear(dgv_ctx* ctx, const double* f, double* fc, int64_t n) {
  LIN_CHECK(ctx, n);
  if (!f || !fc) return set_error(ctx, DGV_ERR_ARG, "null array");
  const size_t nb = sizeof(double) * n;
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_f, f, nb, hipMemcpyHostToDevice, ctx->stream));
  int rc = eval_linear_dev(ctx, ctx->d_f, n, ctx->d_q, n, 1);
  if (rc) return rc;
  DGV_HIP(ctx, hipMemcpyAsync(fc, ctx->d_q, nb, hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

int dgv_ctx_get_fma(dgv_ctx* ctx, double* fmA, int64_t n) {
  LIN_CHECK(ctx, n);
  if (!fmA) return set_error(ctx, DGV_ERR_ARG, "null array");
  if (!ctx->lin.fma_valid) return set_error(ctx, DGV_ERR_STATE, "no linearised operator (PrepareFMA)");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  // device rows r (contiguous in i) -> the reference's column-major fmA(r, i)
  std::vector<double> rows;
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)(64 << 20) / (8 * n)));
  rows.resize((size_t)chunk * n);
  for (int64_t r0 = 0; r0 < n; r0 += chunk) {
    const int64_t nr = std::min(chunk, n - r0);
    DGV_HIP(ctx, hipMemcpyAsync(rows.data(), ctx->lin.d_fma + r0 * n, sizeof(double) * nr * n, hipMemcpyDeviceToHost,
                                ctx->stream));
    DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t r = 0; r < nr; ++r)
      for (int64_t i = 0; i < n; ++i) fmA[(r0 + r) + n * i] = rows[r * n + i];
  }
  return DGV_OK;
}

int dgv_ctx_set_regime(dgv_ctx* ctx, int32_t run_mode, double decomp_lev, double linear_lev) {
  if (!ctx) return DGV_ERR_ARG;
  if (run_mode < 0 || run_mode > 2) return set_error(ctx, DGV_ERR_ARG, "run_mode must be 0, 1 or 2");
  // spat_oper_mod.f90:331-334: the decomposition threshold must lie above the
  // linearisation threshold (decomp_lev = 0 switches the regime check off)
  if (decomp_lev > 0 && decomp_lev <= linear_lev)
    return set_error(ctx, DGV_ERR_ARG,
                     "Decomposition treshold is lower than linearization treshhold (decomp_lev <= linear_lev)");
  if ((run_mode != 0) && !ctx->lin.fma_valid)
    return set_error(ctx, DGV_ERR_STATE, "run_mode 1 or 2 needs a linearised operator (PrepareFMA)");
  ctx->lin.run_mode = run_mode;
  ctx->lin.decomp_lev = decomp_lev;
  ctx->lin.linear_lev = linear_lev;
  if (ctx->rk_graph) { hipGraphExecDestroy(ctx->rk_graph); ctx->rk_graph = nullptr; }
  return DGV_OK;
}

int dgv_ctx_get_regime(dgv_ctx* ctx, int32_t* run_mode, double* l1_err, double* l1_err_fm, int64_t counts[2]) {
  if (!ctx) return DGV_ERR_ARG;
  if (run_mode) *run_mode = ctx->lin.run_mode;
  if (l1_err) *l1_err = ctx->lin.last_l1;
  if (l1_err_fm) *l1_err_fm = ctx->lin.last_l1_fm;
  if (counts) { counts[0] = ctx->lin.nchecks; counts[1] = ctx->lin.nprepare; }
  return DGV_OK;
}

// The Iphi integrals are the plain sums over the records: the conservation
// projection that the context may apply to every Q is held off while they run.
struct ProjectionOff {
  dgv_ctx* ctx;
  int save;
  explicit ProjectionOff(dgv_ctx* c) : ctx(c), save(c->project) { c->project = 0; }
  ~ProjectionOff() { ctx->project = save; }
};

// IphiDGV (dgvtools_mod.f90:847-877): for every node i, one basis function per
// node, Iphi(i) = 2 sum over its records of A f(xi) f(xi1) / w_i, times coef.
// On the shifted A-array these are the evaluator's Q of f, scaled.
int dgv_iphi(dgv_ctx* ctx, const double* f, double* iphi, int64_t n, double coef) {
  LIN_CHECK(ctx, n);
  if (!f || !iphi) return set_error(ctx, DGV_ERR_ARG, "null array");
  const size_t nb = sizeof(double) * n;
  const ProjectionOff plain(ctx);
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_f, f, nb, hipMemcpyHostToDevice, ctx->stream));
  int rc = eval_q_dev(ctx, ctx->d_f, ctx->d_q, coef);
  if (rc) return rc;
  DGV_HIP(ctx, hipMemcpyAsync(iphi, ctx->d_q, nb, hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

// IphiDGV_decomp (dgvtools_mod.f90:896-927): the same integrals with f split at
// its local Maxwellian fM (MassCheckRec, maxwelveldist; lines 912-914): per record
// A (fM(xi) Df(xi1) + Df(xi) fM(xi1) + Df(xi) Df(xi1)), Df = f - fM, i.e. the
// evaluator's mixed term of (Df, fM) plus its Q of Df, times coef.
int dgv_iphi_decomp(dgv_ctx* ctx, const double* f, double* iphi, int64_t n, double coef) {
  LIN_CHECK(ctx, n);
  if (!f || !iphi) return set_error(ctx, DGV_ERR_ARG, "null array");
  const size_t nb = sizeof(double) * n;
  const Grid& g = ctx->g;
  const ProjectionOff plain(ctx);
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_f, f, nb, hipMemcpyHostToDevice, ctx->stream));
  double mom[5];
  int rc = device_moments(ctx, ctx->d_f, mom);
  if (rc) return rc;
  double* fM = ctx->d_tmp[5];
  double* df = ctx->d_tmp[0];
  double* qd = ctx->d_tmp[1];
  const unsigned nblk = stride_blocks(g.N, 256);
  hipLaunchKernelGGL(k_maxwell, dim3(nblk), dim3(256), 0, ctx->stream, g.N, mom[4], mom[1], mom[2], mom[3], mom[0],
                     ctx->d_u, ctx->d_v, ctx->d_w, fM);
  hipLaunchKernelGGL(k_sub, dim3(nblk), dim3(256), 0, ctx->stream, g.N, ctx->d_f, fM, df);
  DGV_HIP(ctx, hipGetLastError());
  if ((rc = eval_mixed_dev(ctx, df, fM, ctx->d_q))) return rc;
  if ((rc = eval_q_dev(ctx, df, qd, 1.0))) return rc;
  hipLaunchKernelGGL(k_sum_scale, dim3(nblk), dim3(256), 0, ctx->stream, g.N, ctx->d_q, qd, coef, ctx->d_q);
  DGV_HIP(ctx, hipGetLastError());
  DGV_HIP(ctx, hipMemcpyAsync(iphi, ctx->d_q, nb, hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

int dgv_check_solution_mode(dgv_ctx* ctx, const double* f, double* df, int64_t n) {
  LIN_CHECK(ctx, n);
  if (!f) return set_error(ctx, DGV_ERR_ARG, "null array");
  const size_t nb = sizeof(double) * n;
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_f, f, nb, hipMemcpyHostToDevice, ctx->stream));
  int rc = check_solution_mode(ctx, ctx->d_f);
  if (rc) return rc;
  if (df) {
    if (ctx->lin.run_mode != 0)
      DGV_HIP(ctx, hipMemcpyAsync(df, ctx->lin.d_vec[2], nb, hipMemcpyDeviceToHost, ctx->stream));
    else
      std::memset(df, 0, nb);  // Df is left unset by the reference in run_mode 0
  }
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

int dgv_ctx_get_fm(dgv_ctx* ctx, double* fm, int64_t n) {
  LIN_CHECK(ctx, n);
  if (!fm) return set_error(ctx, DGV_ERR_ARG, "null array");
  if (!ctx->lin.d_vec[0]) return set_error(ctx, DGV_ERR_STATE, "no Maxwellian set");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMemcpyAsync(fm, ctx->lin.d_vec[0], sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

}  // extern "C"
