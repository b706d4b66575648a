// dgv_trunc.hip -- truncation of the A-array in the pair-block layout.
//
// TruncAarrsTrhls_DGV (dgvtools_mod.f90:1313-1362) keeps the records with
// |A| > trhld; TruncAarrsRadius_DGV (dgvtools_mod.f90:1375-1429) keeps the
// records whose nodes xi, xi1 are at most rad apart (squared distance from the
// node coordinates, compared with rad**2).  Both keep the order of the
// surviving records, so in the pair-block layout a record is a coefficient
// T[l] set to zero (threshold) or a whole pair dropped (radius, which depends on
// the pair only).  Pairs left without any coefficient are removed on the device,
// the block headers are rebuilt on the host and the work plan is rebuilt.
#include <algorithm>
#include <numeric>
#include <vector>

#include "dgv_internal.h"

namespace dgv {

struct TruncArgs {
  double trhld, radsq;
  int use_t, use_r;
  int Mv, Mw, ccu, ccv, ccw, D;
};

// One workgroup per block: apply the truncation to its pairs and flag the
// pairs that keep at least one coefficient.
__global__ __launch_bounds__(256) void k_truncate(const BlockHdr* __restrict__ blocks, double* __restrict__ pairs,
                                                  const double* __restrict__ u, const double* __restrict__ v,
                                                  const double* __restrict__ w, uint8_t* __restrict__ keep,
                                                  TruncArgs t) {
  const BlockHdr B = blocks[blockIdx.x];
  const int D = t.D;
  const int64_t xc = ((int64_t)(t.ccu + B.a[0]) * t.Mv + (t.ccv + B.a[1])) * t.Mw + (t.ccw + B.a[2]);
  const int64_t yc = ((int64_t)(t.ccu + B.b[0]) * t.Mv + (t.ccv + B.b[1])) * t.Mw + (t.ccw + B.b[2]);
  for (int k = threadIdx.x; k < B.npairs; k += blockDim.x) {
    const int64_t pi = B.pbeg + k;
    double* T = pairs + pi * (D + 1);
    const int code = (int)__double_as_longlong(T[D]);
    bool drop = false;
    if (t.use_r) {
      const int64_t i = xc * D + (code >> 8), j = yc * D + (code & 255);
      const double du = u[i] - u[j], dv = v[i] - v[j], dw = w[i] - w[j];
      // (du)**2 + (dv)**2 + (dw)**2 evaluated left to right, no contraction
      const double dsq = __dadd_rn(__dadd_rn(__dmul_rn(du, du), __dmul_rn(dv, dv)), __dmul_rn(dw, dw));
      drop = !(dsq <= t.radsq);
    }
    bool any = false;
    for (int l = 0; l < D; ++l) {
      double a = T[l];
      if (drop || (t.use_t && !(fabs(a) > t.trhld))) a = 0.0;
      T[l] = a;
      any = any || a != 0.0;
    }
    keep[pi] = any ? 1 : 0;
  }
}

__global__ void k_gather_pairs(int64_t npairs, int D, const double* __restrict__ src,
                               const int32_t* __restrict__ newpos, double* __restrict__ dst) {
  const int64_t n = npairs * (D + 1), step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
    const int64_t p = i / (D + 1), l = i % (D + 1);
    const int32_t np = newpos[p];
    if (np >= 0) dst[(int64_t)np * (D + 1) + l] = src[i];
  }
}

__global__ void k_count_nonzero(int64_t npairs, int D, const double* __restrict__ pairs,
                                unsigned long long* __restrict__ cnt) {
  __shared__ unsigned long long c[64];
  if (threadIdx.x < 64) c[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += (int64_t)gridDim.x * blockDim.x)
    for (int l = 0; l < D; ++l)
      if (pairs[p * (D + 1) + l] != 0.0) atomicAdd(&c[l], 1ull);
  __syncthreads();
  if (threadIdx.x < D) atomicAdd(&cnt[threadIdx.x], c[threadIdx.x]);
}

static int truncate_aarray(dgv_ctx* ctx, const TruncArgs& ta) {
  if (!ctx->has_a) return set_error(ctx, DGV_ERR_STATE, "no A-array");
  PairBlock& pb = ctx->pb;
  const int D = ctx->g.D;
  bump_aarray_generation(ctx);
  free_linear(ctx);  // fmA and its tables belong to the old A-array
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int64_t np = pb.npairs;
  uint8_t* d_keep = nullptr;
  DGV_HIP(ctx, hipMalloc(&d_keep, std::max<int64_t>(np, 1)));
  if (!pb.blocks.empty())
    hipLaunchKernelGGL(k_truncate, dim3((unsigned)pb.blocks.size()), dim3(256), 0, ctx->stream, pb.d_blocks,
                       pb.d_pairs, ctx->d_u, ctx->d_v, ctx->d_w, d_keep, ta);
  std::vector<uint8_t> keep(np);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(keep.data(), d_keep, np, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d_keep);
  if (e != hipSuccess) return set_error(ctx, DGV_ERR_HIP, std::string("truncation: ") + hipGetErrorString(e));
  // new positions, block headers (pairs stay sorted by p inside a block)
  std::vector<int32_t> newpos(np, -1);
  int64_t next = 0;
  for (BlockHdr& B : pb.blocks) {
    const int64_t old = B.pbeg;
    int32_t seg[65];
    seg[0] = 0;
    int64_t cnt = 0;
    for (int p = 0; p < D; ++p) {
      for (int64_t k = B.seg[p]; k < B.seg[p + 1]; ++k)
        if (keep[old + k]) newpos[old + k] = (int32_t)(next + cnt++);
      seg[p + 1] = (int32_t)cnt;
    }
    B.pbeg = next;
    B.npairs = (int16_t)cnt;
    for (int p = 0; p <= D; ++p) B.seg[p] = seg[p];
    next += cnt;
  }
  for (Group& G : pb.groups) {
    G.npairs = 0;
    for (int k = G.bbeg; k < G.bend; ++k) G.npairs += pb.blocks[pb.group_blocks[k]].npairs;
  }
  double* d_new = nullptr;
  int32_t* d_pos = nullptr;
  if (hipMalloc(&d_new, sizeof(double) * (next + kPairPad) * (D + 1)) != hipSuccess)
    return set_error(ctx, DGV_ERR_NOMEM, "hipMalloc(truncated pairs) failed");
  if (hipMemsetAsync(d_new + next * (D + 1), 0, sizeof(double) * kPairPad * (D + 1), ctx->stream) != hipSuccess) {
    hipFree(d_new);
    return set_error(ctx, DGV_ERR_HIP, "hipMemsetAsync(truncated pairs padding) failed");
  }
  if (hipMalloc(&d_pos, sizeof(int32_t) * std::max<int64_t>(np, 1)) != hipSuccess) {
    hipFree(d_new);
    return set_error(ctx, DGV_ERR_NOMEM, "hipMalloc(truncation map) failed");
  }
  e = hipMemcpyAsync(d_pos, newpos.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice, ctx->stream);
  const int64_t tot = np * (D + 1);
  if (e == hipSuccess && tot > 0)
    hipLaunchKernelGGL(k_gather_pairs, dim3(stride_blocks(tot, 256)), dim3(256), 0, ctx->stream, np, D,
                       pb.d_pairs, d_pos, d_new);
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d_pos);
  if (e != hipSuccess) {
    hipFree(d_new);
    return set_error(ctx, DGV_ERR_HIP, std::string("truncation: ") + hipGetErrorString(e));
  }
  hipFree(pb.d_pairs);
  pb.d_pairs = d_new;
  pb.npairs = next;
  DGV_HIP(ctx, hipMemcpyAsync(pb.d_blocks, pb.blocks.data(), sizeof(BlockHdr) * pb.blocks.size(),
                              hipMemcpyHostToDevice, ctx->stream));
  DGV_HIP(ctx, hipMemcpyAsync(pb.d_groups, pb.groups.data(), sizeof(Group) * pb.groups.size(), hipMemcpyHostToDevice,
                              ctx->stream));
  // record counts per basis function
  unsigned long long* d_cnt = nullptr;
  DGV_HIP(ctx, hipMalloc(&d_cnt, sizeof(unsigned long long) * 64));
  DGV_HIP(ctx, hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long) * 64, ctx->stream));
  if (next > 0)
    hipLaunchKernelGGL(k_count_nonzero, dim3(1024), dim3(256), 0, ctx->stream, next, D, pb.d_pairs, d_cnt);
  std::vector<unsigned long long> cnt(64, 0);
  e = hipGetLastError();
  // the counts come back on the context's stream (it does not wait for the null stream)
  if (e == hipSuccess)
    e = hipMemcpyAsync(cnt.data(), d_cnt, sizeof(unsigned long long) * 64, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d_cnt);
  if (e != hipSuccess) return set_error(ctx, DGV_ERR_HIP, std::string("truncation: ") + hipGetErrorString(e));
  pb.nrec_basis.assign(D, 0);
  for (int l = 0; l < D; ++l) pb.nrec_basis[l] = (int64_t)cnt[l];
  pb.nrec_ref = std::accumulate(pb.nrec_basis.begin(), pb.nrec_basis.end(), (int64_t)0);
  // derived state: reference-layout copies and the work plan
  hipFree(ctx->d_rec_off); hipFree(ctx->d_rec_A); hipFree(ctx->d_rec_xi); hipFree(ctx->d_rec_xi1);
  ctx->d_rec_off = nullptr; ctx->d_rec_A = nullptr; ctx->d_rec_xi = nullptr; ctx->d_rec_xi1 = nullptr;
  ctx->has_rec = false;
  ctx->exp_valid = false;
  ctx->exp_idx.clear();
  ctx->exp_val.clear();
  return build_plan(ctx, ctx->tile);
}

__global__ void k_abs_coeff(int64_t npairs, int D, double* __restrict__ pairs) {
  const int64_t n = npairs * (D + 1), step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step)
    if (i % (D + 1) < D) pairs[i] = fabs(pairs[i]);
}

}  // namespace dgv

using namespace dgv;

extern "C" {

int dgv_ctx_aarray_abs(dgv_ctx* ctx) {
  if (!ctx) { set_global_error("null context"); return DGV_ERR_ARG; }
  if (!ctx->has_a) return set_error(ctx, DGV_ERR_STATE, "no A-array");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  bump_aarray_generation(ctx);
  const int64_t tot = ctx->pb.npairs * (ctx->g.D + 1);
  if (tot > 0)
    hipLaunchKernelGGL(k_abs_coeff, dim3(stride_blocks(tot, 256)), dim3(256), 0, ctx->stream, ctx->pb.npairs,
                       ctx->g.D, ctx->pb.d_pairs);
  DGV_HIP(ctx, hipGetLastError());
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  hipFree(ctx->d_rec_off); hipFree(ctx->d_rec_A); hipFree(ctx->d_rec_xi); hipFree(ctx->d_rec_xi1);
  ctx->d_rec_off = nullptr; ctx->d_rec_A = nullptr; ctx->d_rec_xi = nullptr; ctx->d_rec_xi1 = nullptr;
  ctx->has_rec = false;
  ctx->exp_valid = false;
  ctx->exp_idx.clear();
  ctx->exp_val.clear();
  return DGV_OK;
}


int dgv_ctx_trunc_aarray_threshold(dgv_ctx* ctx, double trhld) {
  if (!ctx) { set_global_error("null context"); return DGV_ERR_ARG; }
  if (!(trhld >= 0)) return set_error(ctx, DGV_ERR_ARG, "threshold must be >= 0");
  const Grid& g = ctx->g;
  TruncArgs t{trhld, 0.0, 1, 0, g.Mv, g.Mw, g.ccu, g.ccv, g.ccw, g.D};
  return truncate_aarray(ctx, t);
}

int dgv_ctx_trunc_aarray_radius(dgv_ctx* ctx, double rad) {
  if (!ctx) { set_global_error("null context"); return DGV_ERR_ARG; }
  if (!(rad >= 0)) return set_error(ctx, DGV_ERR_ARG, "radius must be >= 0");
  const Grid& g = ctx->g;
  TruncArgs t{0.0, rad * rad, 0, 1, g.Mv, g.Mw, g.ccu, g.ccv, g.ccw, g.D};
  return truncate_aarray(ctx, t);
}

}  // extern "C"
