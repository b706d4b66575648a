// dgv_agen.hip -- GPU generator of the A-array (SetA_DGV, dgvtools_mod.f90:377-681)
// for all canonical basis functions at once, written directly in the
// pair-block layout.
//
// Pipeline (cp = ordered cell pair X <= Y in the reference's cell order):
//   1. k_count   : per cp, pairs (xi in X, xi1 in Y, xi < xi1) passing the
//                  prefilter of lines 493-530                          -> cnt[cp]
//   2. host scan -> offsets; k_list writes the flat list of (cp, p, q)
//   3. k_integrate: one thread per pair, the D values of A (dgv_agen_math.h),
//                  min_sens test, weights and |g|/8                      -> pairs
//   4. k_keep_scan + k_compact: drop pairs whose D values are all zero and
//                  emit per-cp segment tables (the block headers)
#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <numeric>

#include "dgv_agen_math.h"
#include "dgv_internal.h"

namespace dgv {

static void lagrange_monomials(const double* kap, int n, double mono[kAgenMaxS][kAgenMaxS]) {
  for (int i = 0; i < kAgenMaxS; ++i)
    for (int m = 0; m < kAgenMaxS; ++m) mono[i][m] = 0.0;
  for (int i = 0; i < n; ++i) {
    double c[kAgenMaxS + 1] = {1.0};
    int deg = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      // multiply by (kap[j] - y) / (kap[j] - kap[i])
      const double den = kap[j] - kap[i];
      double nc[kAgenMaxS + 1] = {0};
      for (int m = 0; m <= deg; ++m) {
        nc[m] += c[m] * kap[j] / den;
        nc[m + 1] += -c[m] / den;
      }
      ++deg;
      for (int m = 0; m <= deg; ++m) c[m] = nc[m];
    }
    for (int m = 0; m < n; ++m) mono[i][m] = c[m];
  }
}

static int make_consts(const Grid& g, const dgv_agen_params& p, AgenConsts& C, std::string& err) {
  std::memset(&C, 0, sizeof(C));
  if (g.su > kAgenMaxS || g.sv > kAgenMaxS || g.sw > kAgenMaxS) {
    err = "A-array generator supports at most 4 nodes per cell per dimension";
    return DGV_ERR_UNSUPPORTED;
  }
  C.su = g.su; C.sv = g.sv; C.sw = g.sw;
  const int cc = g.ccell;
  C.u0 = g.lu[cc]; C.u1 = g.ru[cc]; C.v0 = g.lv[cc]; C.v1 = g.rv[cc]; C.w0 = g.lw[cc]; C.w1 = g.rw[cc];
  const int s[3] = {g.su, g.sv, g.sw};
  for (int d = 0; d < 3; ++d) {
    for (int i = 0; i < s[d]; ++i) C.kap[d][i] = g.gn[d][i];
    lagrange_monomials(C.kap[d], s[d], C.mono[d]);
  }
  int nq = p.nquad > 0 ? p.nquad : 24;
  if (nq > kAgenMaxQ) { err = "nquad too large (max 48)"; return DGV_ERR_ARG; }
  C.nq = nq;
  std::vector<double> x(nq), w(nq);
  if (gauss_legendre(nq, x.data(), w.data())) { err = "Gauss-Legendre failed"; return DGV_ERR_ARG; }
  const double pi = 3.141592653589793238462643383279502884;
  for (int i = 0; i < nq; ++i) {
    C.qs[i] = 0.5 * pi * (x[i] + 1.0);
    C.qw[i] = 0.5 * pi * w[i];
  }
  C.Trad = p.Trad > 0 ? p.Trad : 100.0;
  C.moment_mode = p.moment_mode;
  if (p.moment_mode < 0 || p.moment_mode > 5) { err = "moment_mode must be 0..5"; return DGV_ERR_ARG; }
  return DGV_OK;
}

template <int SU, int SV, int SW>
DGV_HD inline int pair_raw_dispatch(const AgenConsts& C, const double x[3], const double y[3], double* raw,
                                    double* dsph) {
  return pair_raw<SU, SV, SW>(C, x, y, raw, dsph);
}

static int pair_raw_host(const AgenConsts& C, const double x[3], const double y[3], double* raw, double* dsph) {
  if (C.su == 3 && C.sv == 3 && C.sw == 3) return pair_raw<3, 3, 3>(C, x, y, raw, dsph);
  if (C.su == 2 && C.sv == 2 && C.sw == 2) return pair_raw<2, 2, 2>(C, x, y, raw, dsph);
  if (C.su == 1 && C.sv == 1 && C.sw == 1) return pair_raw<1, 1, 1>(C, x, y, raw, dsph);
  if (C.su == 4 && C.sv == 4 && C.sw == 4) return pair_raw<4, 4, 4>(C, x, y, raw, dsph);
  return -2;
}

// ---------------------------------------------------------------------------
struct NodeGeo {
  const double *u, *v, *w, *gwts;
};

__device__ __forceinline__ bool prefilter_dev(const AgenConsts& C, const double* x, const double* y) {
  const double d0 = x[0] - y[0], d1 = x[1] - y[1], d2 = x[2] - y[2];
  const double dsph = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  const double s0 = x[0] + y[0] - C.u0 - C.u1, s1 = x[1] + y[1] - C.v0 - C.v1, s2 = x[2] + y[2] - C.w0 - C.w1;
  const double dist2 = sqrt(s0 * s0 + s1 * s1 + s2 * s2);
  const double e0 = C.u1 - C.u0, e1 = C.v1 - C.v0, e2 = C.w1 - C.w0;
  const double dphi = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
  return !((dist2 > dsph + dphi) || (dist2 + dphi < dsph) || (dsph > C.Trad));
}

template <int D>
__global__ void k_count(AgenConsts C, int64_t ncp, const int2* __restrict__ cps, NodeGeo G,
                        uint32_t* __restrict__ cnt) {
  const int64_t cp = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cp >= ncp) return;
  const int2 XY = cps[cp];
  uint32_t n = 0;
  for (int p = 0; p < D; ++p) {
    const int64_t i = (int64_t)XY.x * D + p;
    const double x[3] = {G.u[i], G.v[i], G.w[i]};
    for (int q = (XY.x == XY.y ? p + 1 : 0); q < D; ++q) {
      const int64_t j = (int64_t)XY.y * D + q;
      const double y[3] = {G.u[j], G.v[j], G.w[j]};
      n += prefilter_dev(C, x, y) ? 1u : 0u;
    }
  }
  cnt[cp] = n;
}

template <int D>
__global__ void k_list(AgenConsts C, int64_t ncp, const int2* __restrict__ cps, NodeGeo G,
                       const int64_t* __restrict__ off, uint64_t* __restrict__ list) {
  const int64_t cp = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cp >= ncp) return;
  const int2 XY = cps[cp];
  int64_t o = off[cp];
  for (int p = 0; p < D; ++p) {
    const int64_t i = (int64_t)XY.x * D + p;
    const double x[3] = {G.u[i], G.v[i], G.w[i]};
    for (int q = (XY.x == XY.y ? p + 1 : 0); q < D; ++q) {
      const int64_t j = (int64_t)XY.y * D + q;
      const double y[3] = {G.u[j], G.v[j], G.w[j]};
      if (prefilter_dev(C, x, y)) list[o++] = ((uint64_t)cp << 16) | ((uint64_t)p << 8) | (uint64_t)q;
    }
  }
}

template <int SU, int SV, int SW>
__global__ __launch_bounds__(128) void k_integrate(AgenConsts C, int64_t npairs, const int2* __restrict__ cps,
                                                   NodeGeo G, const uint64_t* __restrict__ list, double min_sens,
                                                   double* __restrict__ pairs, uint8_t* __restrict__ keep) {
  constexpr int D = SU * SV * SW;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npairs) return;
  const uint64_t e = list[k];
  const int64_t cp = (int64_t)(e >> 16);
  const int p = (int)((e >> 8) & 255), q = (int)(e & 255);
  const int2 XY = cps[cp];
  const int64_t i = (int64_t)XY.x * D + p, j = (int64_t)XY.y * D + q;
  const double x[3] = {G.u[i], G.v[i], G.w[i]};
  const double y[3] = {G.u[j], G.v[j], G.w[j]};
  double raw[D];
  double dsph;
  pair_raw<SU, SV, SW>(C, x, y, raw, &dsph);
  const double wi = G.gwts[i], wj = G.gwts[j];
  double* out = pairs + k * (D + 1);
  bool any = false;
  for (int l = 0; l < D; ++l) {
    double a = 0.0;
    if (fabs(raw[l]) > min_sens) {  // line 641, then lines 654-659
      a = raw[l] * wj;
      a = a * wi;
      a = a * (dsph / 8.0);
      any = true;
    }
    out[l] = a;
  }
  out[D] = __longlong_as_double((long long)pair_code_bits(p, q));
  keep[k] = any ? 1 : 0;
}

// Per cp: number of kept pairs and, for each p, the first kept pair with local
// index of xi >= p.
template <int D>
__global__ void k_segments(int64_t ncp, const int64_t* __restrict__ off, const uint64_t* __restrict__ list,
                           const uint8_t* __restrict__ keep, int32_t* __restrict__ kcnt, uint16_t* __restrict__ seg) {
  const int64_t cp = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cp >= ncp) return;
  uint16_t* s = seg + cp * (D + 1);
  int n = 0, p_next = 0;
  for (int64_t k = off[cp]; k < off[cp + 1]; ++k) {
    if (!keep[k]) continue;
    const int p = (int)((list[k] >> 8) & 255);
    while (p_next <= p) s[p_next++] = (uint16_t)n;
    ++n;
  }
  while (p_next <= D) s[p_next++] = (uint16_t)n;
  kcnt[cp] = n;
}

template <int D>
__global__ void k_compact(int64_t npairs, const uint8_t* __restrict__ keep, const int64_t* __restrict__ dst,
                          const double* __restrict__ src, double* __restrict__ out) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < npairs; k += step) {
    if (!keep[k]) continue;
    const double* a = src + k * (D + 1);
    double* b = out + dst[k] * (D + 1);
    for (int l = 0; l <= D; ++l) b[l] = a[l];
  }
}

// per-basis record counts (integer atomics: exact and order independent)
template <int D>
__global__ void k_count_basis(int64_t npairs, const double* __restrict__ pairs, unsigned long long* __restrict__ cnt) {
  __shared__ unsigned long long loc[D];
  for (int l = threadIdx.x; l < D; l += blockDim.x) loc[l] = 0;
  __syncthreads();
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < npairs; k += (int64_t)gridDim.x * blockDim.x)
    for (int l = 0; l < D; ++l)
      if (pairs[k * (D + 1) + l] != 0.0) atomicAdd(&loc[l], 1ull);
  __syncthreads();
  for (int l = threadIdx.x; l < D; l += blockDim.x) atomicAdd(&cnt[l], loc[l]);
}

template <int SU, int SV, int SW>
static int generate_t(dgv_ctx* ctx, const AgenConsts& C, double min_sens) {
  constexpr int D = SU * SV * SW;
  const Grid& g = ctx->g;
  const int64_t nc = g.ncells;
  // ordered cell pairs X <= Y (node ids of X precede those of Y)
  std::vector<int2> cps;
  cps.reserve((size_t)(nc * (nc + 1) / 2));
  for (int64_t X = 0; X < nc; ++X)
    for (int64_t Y = X; Y < nc; ++Y) cps.push_back(make_int2((int)X, (int)Y));
  const int64_t ncp = (int64_t)cps.size();
  int2* d_cps = nullptr;
  uint32_t* d_cnt = nullptr;
  int64_t* d_off = nullptr;
  uint64_t* d_list = nullptr;
  double *d_raw = nullptr, *d_pairs = nullptr;
  uint8_t* d_keep = nullptr;
  int64_t* d_dst = nullptr;
  int32_t* d_kcnt = nullptr;
  uint16_t* d_seg = nullptr;
  unsigned long long* d_bcnt = nullptr;
  int rc = DGV_OK;
  auto cleanup = [&]() {
    hipFree(d_cps); hipFree(d_cnt); hipFree(d_off); hipFree(d_list); hipFree(d_raw); hipFree(d_keep);
    hipFree(d_dst); hipFree(d_kcnt); hipFree(d_seg); hipFree(d_bcnt);
  };
#define AG_HIP(call)                                                                           \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) {                                                                    \
      cleanup();                                                                               \
      if (d_pairs) hipFree(d_pairs);                                                           \
      return set_error(ctx, e_ == hipErrorOutOfMemory ? DGV_ERR_NOMEM : DGV_ERR_HIP,           \
                       std::string("A-array generator: ") + #call + ": " + hipGetErrorString(e_)); \
    }                                                                                          \
  } while (0)
  NodeGeo G{ctx->d_u, ctx->d_v, ctx->d_w, ctx->d_gwts};
  hipStream_t st = ctx->stream;
  AG_HIP(hipMalloc(&d_cps, sizeof(int2) * ncp));
  AG_HIP(hipMemcpy(d_cps, cps.data(), sizeof(int2) * ncp, hipMemcpyHostToDevice));
  AG_HIP(hipMalloc(&d_cnt, sizeof(uint32_t) * ncp));
  const unsigned gcp = (unsigned)((ncp + 127) / 128);
  hipLaunchKernelGGL(k_count<D>, dim3(gcp), dim3(128), 0, st, C, ncp, d_cps, G, d_cnt);
  AG_HIP(hipGetLastError());
  std::vector<uint32_t> cnt(ncp);
  AG_HIP(hipMemcpyAsync(cnt.data(), d_cnt, sizeof(uint32_t) * ncp, hipMemcpyDeviceToHost, st));
  AG_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> off(ncp + 1, 0);
  for (int64_t i = 0; i < ncp; ++i) off[i + 1] = off[i] + cnt[i];
  const int64_t np = off[ncp];
  AG_HIP(hipMalloc(&d_off, sizeof(int64_t) * (ncp + 1)));
  AG_HIP(hipMemcpy(d_off, off.data(), sizeof(int64_t) * (ncp + 1), hipMemcpyHostToDevice));
  AG_HIP(hipMalloc(&d_list, sizeof(uint64_t) * std::max<int64_t>(np, 1)));
  hipLaunchKernelGGL(k_list<D>, dim3(gcp), dim3(128), 0, st, C, ncp, d_cps, G, d_off, d_list);
  AG_HIP(hipGetLastError());
  AG_HIP(hipMalloc(&d_raw, sizeof(double) * (D + 1) * std::max<int64_t>(np, 1)));
  AG_HIP(hipMalloc(&d_keep, std::max<int64_t>(np, 1)));
  // integrate in slices so that each launch stays short
  const int64_t slice = 1 << 22;
  for (int64_t b = 0; b < np; b += slice) {
    const int64_t n = std::min(slice, np - b);
    k_integrate<SU, SV, SW><<<dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st>>>(
        C, n, d_cps, G, d_list + b, min_sens, d_raw + b * (D + 1), d_keep + b);
    AG_HIP(hipGetLastError());
  }
  // compaction
  std::vector<uint8_t> keep(np);
  AG_HIP(hipMemcpyAsync(keep.data(), d_keep, np, hipMemcpyDeviceToHost, st));
  AG_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> dst(np);
  int64_t nk = 0;
  for (int64_t k = 0; k < np; ++k) { dst[k] = nk; nk += keep[k]; }
  AG_HIP(hipMalloc(&d_dst, sizeof(int64_t) * std::max<int64_t>(np, 1)));
  AG_HIP(hipMemcpy(d_dst, dst.data(), sizeof(int64_t) * np, hipMemcpyHostToDevice));
  AG_HIP(hipMalloc(&d_pairs, sizeof(double) * (D + 1) * (nk + kPairPad)));
  AG_HIP(hipMemsetAsync(d_pairs + (D + 1) * nk, 0, sizeof(double) * (D + 1) * kPairPad, st));
  if (np > 0)
    hipLaunchKernelGGL(k_compact<D>, dim3(stride_blocks(np, 256)), dim3(256), 0, st, np, d_keep, d_dst, d_raw,
                       d_pairs);
  AG_HIP(hipMalloc(&d_kcnt, sizeof(int32_t) * ncp));
  AG_HIP(hipMalloc(&d_seg, sizeof(uint16_t) * (D + 1) * ncp));
  hipLaunchKernelGGL(k_segments<D>, dim3(gcp), dim3(128), 0, st, ncp, d_off, d_list, d_keep, d_kcnt, d_seg);
  AG_HIP(hipMalloc(&d_bcnt, sizeof(unsigned long long) * D));
  AG_HIP(hipMemsetAsync(d_bcnt, 0, sizeof(unsigned long long) * D, st));
  if (nk > 0) hipLaunchKernelGGL(k_count_basis<D>, dim3(1024), dim3(256), 0, st, nk, d_pairs, d_bcnt);
  AG_HIP(hipGetLastError());
  std::vector<int32_t> kcnt(ncp);
  std::vector<uint16_t> seg((size_t)ncp * (D + 1));
  std::vector<unsigned long long> bcnt(D);
  AG_HIP(hipMemcpyAsync(kcnt.data(), d_kcnt, sizeof(int32_t) * ncp, hipMemcpyDeviceToHost, st));
  AG_HIP(hipMemcpyAsync(seg.data(), d_seg, sizeof(uint16_t) * (D + 1) * ncp, hipMemcpyDeviceToHost, st));
  AG_HIP(hipMemcpyAsync(bcnt.data(), d_bcnt, sizeof(unsigned long long) * D, hipMemcpyDeviceToHost, st));
  AG_HIP(hipStreamSynchronize(st));
  cleanup();
  d_cps = nullptr; d_cnt = nullptr; d_off = nullptr; d_list = nullptr; d_raw = nullptr; d_keep = nullptr;
  d_dst = nullptr; d_kcnt = nullptr; d_seg = nullptr; d_bcnt = nullptr;
  // blocks (one per cell pair with kept pairs) and groups
  free_plan(ctx);
  free_pairblock(ctx);
  PairBlock& pb = ctx->pb;
  pb.D = D;
  pb.npairs = nk;
  pb.d_pairs = d_pairs;
  int64_t pbeg = 0;
  const int cc[3] = {g.ccu, g.ccv, g.ccw};
  std::map<std::array<int, 6>, std::vector<int32_t>> groups;
  for (int64_t cp = 0; cp < ncp; ++cp) {
    if (kcnt[cp] == 0) continue;
    BlockHdr h;
    std::memset(&h, 0, sizeof(h));
    h.pbeg = pbeg;
    const int X = cps[cp].x, Y = cps[cp].y;
    const int xc[3] = {(int)(X / ((int64_t)g.Mv * g.Mw)), (int)((X / g.Mw) % g.Mv), (int)(X % g.Mw)};
    const int yc[3] = {(int)(Y / ((int64_t)g.Mv * g.Mw)), (int)((Y / g.Mw) % g.Mv), (int)(Y % g.Mw)};
    int lo[3], ext[3];
    const int M[3] = {g.Mu, g.Mv, g.Mw};
    for (int d = 0; d < 3; ++d) {
      h.a[d] = (int8_t)(xc[d] - cc[d]);
      h.b[d] = (int8_t)(yc[d] - cc[d]);
      lo[d] = std::max(0, std::max(-(int)h.a[d], -(int)h.b[d]));
      ext[d] = std::min(M[d], std::min(M[d] - h.a[d], M[d] - h.b[d])) - lo[d];
    }
    h.npairs = (int16_t)kcnt[cp];
    for (int p = 0; p <= D; ++p) h.seg[p] = seg[(size_t)cp * (D + 1) + p];
    groups[{lo[0], lo[1], lo[2], ext[0], ext[1], ext[2]}].push_back((int32_t)pb.blocks.size());
    pb.blocks.push_back(h);
    pbeg += kcnt[cp];
  }
  for (auto& kv : groups) {
    Group Gr;
    for (int d = 0; d < 3; ++d) { Gr.lo[d] = kv.first[d]; Gr.ext[d] = kv.first[3 + d]; }
    Gr.ncells = Gr.ext[0] * Gr.ext[1] * Gr.ext[2];
    Gr.bbeg = (int32_t)pb.group_blocks.size();
    Gr.npairs = 0;
    for (int32_t b : kv.second) { pb.group_blocks.push_back(b); Gr.npairs += pb.blocks[b].npairs; }
    Gr.bend = (int32_t)pb.group_blocks.size();
    pb.groups.push_back(Gr);
  }
  pb.nrec_basis.assign(D, 0);
  for (int l = 0; l < D; ++l) pb.nrec_basis[l] = (int64_t)bcnt[l];
  pb.nrec_ref = std::accumulate(pb.nrec_basis.begin(), pb.nrec_basis.end(), (int64_t)0);
  DGV_HIP(ctx, hipMalloc(&pb.d_blocks, sizeof(BlockHdr) * std::max<size_t>(pb.blocks.size(), 1)));
  DGV_HIP(ctx, hipMemcpy(pb.d_blocks, pb.blocks.data(), sizeof(BlockHdr) * pb.blocks.size(), hipMemcpyHostToDevice));
  DGV_HIP(ctx, hipMalloc(&pb.d_group_blocks, sizeof(int32_t) * std::max<size_t>(pb.group_blocks.size(), 1)));
  DGV_HIP(ctx, hipMemcpy(pb.d_group_blocks, pb.group_blocks.data(), sizeof(int32_t) * pb.group_blocks.size(),
                         hipMemcpyHostToDevice));
  DGV_HIP(ctx, hipMalloc(&pb.d_groups, sizeof(Group) * std::max<size_t>(pb.groups.size(), 1)));
  DGV_HIP(ctx, hipMemcpy(pb.d_groups, pb.groups.data(), sizeof(Group) * pb.groups.size(), hipMemcpyHostToDevice));
  ctx->has_a = true;
  rc = build_plan(ctx, ctx->tile);
#undef AG_HIP
  return rc;
}

int generate_pairblock(dgv_ctx* ctx, const dgv_agen_params& p) {
  AgenConsts C;
  std::string err;
  int rc = make_consts(ctx->g, p, C, err);
  if (rc) return set_error(ctx, rc, err);
  if (p.min_sens < 0) return set_error(ctx, DGV_ERR_ARG, "min_sens must be >= 0");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  const Grid& g = ctx->g;
  if (g.su == 3 && g.sv == 3 && g.sw == 3) return generate_t<3, 3, 3>(ctx, C, p.min_sens);
  if (g.su == 2 && g.sv == 2 && g.sw == 2) return generate_t<2, 2, 2>(ctx, C, p.min_sens);
  if (g.su == 1 && g.sv == 1 && g.sw == 1) return generate_t<1, 1, 1>(ctx, C, p.min_sens);
  return set_error(ctx, DGV_ERR_UNSUPPORTED, "A-array generator supports 1, 2 or 3 nodes per cell per dimension");
}

}  // namespace dgv

extern "C" int dgv_agen_pair(const dgv_grid_desc* grid, int64_t i2, int64_t i3, const dgv_agen_params* params,
                             double* raw, double* A) {
  using namespace dgv;
  if (!grid || !raw) { set_global_error("null argument"); return DGV_ERR_ARG; }
  Grid g;
  std::string err;
  int rc = build_grid(*grid, g, err);
  if (rc) { set_global_error(err); return rc; }
  dgv_agen_params p;
  p.Trad = 100.0; p.min_sens = 0.0; p.nquad = 0; p.moment_mode = 0;
  if (params) p = *params;
  AgenConsts C;
  rc = make_consts(g, p, C, err);
  if (rc) { set_global_error(err); return rc; }
  if (i2 < 1 || i3 < 1 || i2 > g.N || i3 > g.N) { set_global_error("node id out of range"); return DGV_ERR_INDEX; }
  const double x[3] = {g.u[i2 - 1], g.v[i2 - 1], g.w[i2 - 1]};
  const double y[3] = {g.u[i3 - 1], g.v[i3 - 1], g.w[i3 - 1]};
  double dsph = 0;
  rc = pair_raw_host(C, x, y, raw, &dsph);
  if (rc == -2) { set_global_error("unsupported nodes per cell"); return DGV_ERR_UNSUPPORTED; }
  const int D = g.D;
  if (rc == -1) {  // fails the prefilter: no record
    for (int l = 0; l < D; ++l) { raw[l] = 0.0; if (A) A[l] = 0.0; }
    return 1;
  }
  if (A)
    for (int l = 0; l < D; ++l) {
      double a = 0.0;
      if (std::fabs(raw[l]) > p.min_sens) {
        a = raw[l] * g.gwts[i3 - 1];
        a = a * g.gwts[i2 - 1];
        a = a * (dsph / 8.0);
      }
      A[l] = a;
    }
  return DGV_OK;
}
