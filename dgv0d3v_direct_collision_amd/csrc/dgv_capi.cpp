// dgv_capi.cpp -- the extern "C" entry points declared in include/dgv.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>

#include "dgv_internal.h"

namespace dgv {
const char* global_error();
int launch_lincomb(dgv_ctx* ctx, int64_t n, double* out, int nterms, const double* const* xs, const double* m1,
                   const double* m2, const double* den, const int* neg);
int get_basis_records(const dgv_ctx* ctx, int l, double* A, int32_t* xi, int32_t* xi1, int64_t cap,
                      int64_t* nrec, int index_base);
int export_bases_ordered(const dgv_ctx* ctx, const int32_t* ls, int nl, double* out_A, int32_t* out_xi,
                         int32_t* out_xi1, int index_base);
int get_all_records(const dgv_ctx* ctx, std::vector<int32_t>& capphi, std::vector<double>& A,
                    std::vector<int32_t>& xi, std::vector<int32_t>& xi1, int index_base);
int launch_eval_main_only(dgv_ctx* ctx, const double* d_f);
int workload(dgv_ctx* ctx, int64_t out[4]);
}  // namespace dgv

using namespace dgv;

#define CHECK_CTX(ctx) \
  do { if (!(ctx)) { set_global_error("null context"); return DGV_ERR_ARG; } } while (0)
#define CHECK_N(ctx, n)                                                                            \
  do {                                                                                             \
    if ((n) != (ctx)->g.N)                                                                         \
      return set_error((ctx), DGV_ERR_ARG, "array length " + std::to_string((long long)(n)) +      \
                                               " differs from the number of nodes " +               \
                                               std::to_string((long long)(ctx)->g.N));             \
  } while (0)

// Cholesky factor of the Gram matrix G_jk = sum_i w_i psi_j psi_k
// (psi = 1,u,v,w,|u|^2), accumulated in long double.
int dgv::gram_cholesky(const Grid& g, double L[25]) {
  long double G[5][5] = {};
  for (int64_t i = 0; i < g.N; ++i) {
    const long double u = g.u[i], v = g.v[i], w = g.w[i];
    const long double psi[5] = {1.0L, u, v, w, u * u + v * v + w * w};
    for (int a = 0; a < 5; ++a)
      for (int b = 0; b <= a; ++b) G[a][b] += (long double)g.gwts[i] * psi[a] * psi[b];
  }
  long double C[5][5] = {};
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j <= i; ++j) {
      long double s = G[i][j];
      for (int k = 0; k < j; ++k) s -= C[i][k] * C[j][k];
      if (i == j) {
        if (!(s > 0)) return -1;
        C[i][i] = std::sqrt(s);
      } else {
        C[i][j] = s / C[j][j];
      }
    }
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) L[i * 5 + j] = (double)C[i][j];
  return 0;
}

extern "C" {

const char* dgv_last_error(const dgv_ctx* ctx) { return ctx ? ctx->err.c_str() : global_error(); }

int dgv_ctx_create(const dgv_grid_desc* grid, int device, dgv_ctx** out) {
  if (!grid || !out) { set_global_error("null argument"); return DGV_ERR_ARG; }
  *out = nullptr;
  dgv_ctx* ctx = new (std::nothrow) dgv_ctx();
  if (!ctx) { set_global_error("out of host memory"); return DGV_ERR_NOMEM; }
  std::string err;
  int rc = build_grid(*grid, ctx->g, err);
  if (rc) { set_global_error(err); delete ctx; return rc; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_global_error("no HIP device available: the DGV evaluator runs on the GPU only");
    delete ctx;
    return DGV_ERR_HIP;
  }
  if (device < 0 || device >= ndev) { set_global_error("device index out of range"); delete ctx; return DGV_ERR_ARG; }
  ctx->device = device;
  const Grid& g = ctx->g;
  auto fail = [&](int code, const char* m) { set_global_error(m); dgv_ctx_destroy(ctx); return code; };
  if (hipSetDevice(device) != hipSuccess) return fail(DGV_ERR_HIP, "hipSetDevice failed");
  // the allocation-free evaluator's device array: its address on this device, once
  if (static_scratch_setup(device) != DGV_OK) return fail(DGV_ERR_HIP, "hipGetSymbolAddress(scratch) failed");
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(DGV_ERR_HIP, "hipStreamCreate failed");
  for (int i = 0; i < 4; ++i)
    if (hipEventCreate(&ctx->ev[i]) != hipSuccess) return fail(DGV_ERR_HIP, "hipEventCreate failed");
  const size_t nb = sizeof(double) * g.N;
  if (hipMalloc(&ctx->d_gwts, nb) || hipMalloc(&ctx->d_u, nb) || hipMalloc(&ctx->d_v, nb) ||
      hipMalloc(&ctx->d_w, nb) || hipMalloc(&ctx->d_f, nb) || hipMalloc(&ctx->d_q, nb))
    return fail(DGV_ERR_NOMEM, "hipMalloc failed");
  for (int i = 0; i < 6; ++i)
    if (hipMalloc(&ctx->d_tmp[i], nb)) return fail(DGV_ERR_NOMEM, "hipMalloc failed");
  if (hipMalloc(&ctx->d_red, sizeof(double) * (256 * 5 + 64))) return fail(DGV_ERR_NOMEM, "hipMalloc failed");
  if (hipMemcpy(ctx->d_gwts, g.gwts.data(), nb, hipMemcpyHostToDevice) ||
      hipMemcpy(ctx->d_u, g.u.data(), nb, hipMemcpyHostToDevice) ||
      hipMemcpy(ctx->d_v, g.v.data(), nb, hipMemcpyHostToDevice) ||
      hipMemcpy(ctx->d_w, g.w.data(), nb, hipMemcpyHostToDevice))
    return fail(DGV_ERR_HIP, "hipMemcpy failed");
  double L[25];
  if (gram_cholesky(g, L) == 0) {
    if (hipMemcpy(ctx->d_red + 256 * 5 + 16, L, sizeof(L), hipMemcpyHostToDevice))
      return fail(DGV_ERR_HIP, "hipMemcpy failed");
  }
  *out = ctx;
  return DGV_OK;
}

void dgv_ctx_destroy(dgv_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  free_plan(ctx);
  free_pairblock(ctx);
  free_linear_vectors(ctx);
  hipFree(ctx->d_gwts); hipFree(ctx->d_u); hipFree(ctx->d_v); hipFree(ctx->d_w);
  hipFree(ctx->d_f); hipFree(ctx->d_q); hipFree(ctx->d_red); hipFree(ctx->d_hi);
  for (int i = 0; i < 6; ++i) hipFree(ctx->d_tmp[i]);
  for (int i = 0; i < 4; ++i)
    if (ctx->ev[i]) hipEventDestroy(ctx->ev[i]);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

int dgv_ctx_num_nodes(const dgv_ctx* ctx, int64_t* nnodes) {
  CHECK_CTX(ctx);
  if (!nnodes) return DGV_ERR_ARG;
  *nnodes = ctx->g.N;
  return DGV_OK;
}

int dgv_grid_geometry(const dgv_grid_desc* grid, double* u, double* v, double* w, double* gwts, int32_t* ccell,
                      int64_t* nnodes) {
  if (!grid) { set_global_error("null grid"); return DGV_ERR_ARG; }
  Grid g;
  std::string err;
  int rc = build_grid(*grid, g, err);
  if (rc) { set_global_error(err); return rc; }
  if (u) std::copy(g.u.begin(), g.u.end(), u);
  if (v) std::copy(g.v.begin(), g.v.end(), v);
  if (w) std::copy(g.w.begin(), g.w.end(), w);
  if (gwts) std::copy(g.gwts.begin(), g.gwts.end(), gwts);
  if (ccell) *ccell = g.ccell + 1;
  if (nnodes) *nnodes = g.N;
  return DGV_OK;
}

int dgv_ctx_canonical_cell(const dgv_ctx* ctx, int32_t* ccell) {
  CHECK_CTX(ctx);
  if (!ccell) return DGV_ERR_ARG;
  *ccell = ctx->g.ccell + 1;
  return DGV_OK;
}

int dgv_ctx_get_nodes(const dgv_ctx* ctx, double* u, double* v, double* w, double* gwts) {
  CHECK_CTX(ctx);
  const Grid& g = ctx->g;
  if (u) std::copy(g.u.begin(), g.u.end(), u);
  if (v) std::copy(g.v.begin(), g.v.end(), v);
  if (w) std::copy(g.w.begin(), g.w.end(), w);
  if (gwts) std::copy(g.gwts.begin(), g.gwts.end(), gwts);
  return DGV_OK;
}

int dgv_ctx_get_shifts(const dgv_ctx* ctx, int32_t* phican, int32_t* dui, int32_t* dvi, int32_t* dwi) {
  CHECK_CTX(ctx);
  const Grid& g = ctx->g;
  for (int64_t i = 0; i < g.N; ++i) {
    const int64_t cell = i / g.D;
    const int l = (int)(i % g.D);
    const int cw = (int)(cell % g.Mw), cv = (int)((cell / g.Mw) % g.Mv), cu = (int)(cell / ((int64_t)g.Mw * g.Mv));
    if (phican) phican[i] = (int32_t)((int64_t)g.ccell * g.D + l + 1);
    if (dui) dui[i] = cu - g.ccu;
    if (dvi) dvi[i] = cv - g.ccv;
    if (dwi) dwi[i] = cw - g.ccw;
  }
  return DGV_OK;
}

int dgv_ctx_set_aarray(dgv_ctx* ctx, const int32_t* capphi, int64_t ncapphi, const double* A, const int32_t* A_xi,
                       const int32_t* A_xi1, int64_t nrec, int index_base) {
  CHECK_CTX(ctx);
  return build_pairblock_from_records(ctx, capphi, ncapphi, A, A_xi, A_xi1, nrec, index_base);
}

int dgv_ctx_generate_aarray(dgv_ctx* ctx, const dgv_agen_params* params) {
  CHECK_CTX(ctx);
  dgv_agen_params p;
  p.Trad = 100.0;
  p.min_sens = 0.0;
  p.nquad = 0;
  p.moment_mode = 0;
  if (params) p = *params;
  return generate_pairblock(ctx, p);
}

int dgv_ctx_aarray_size(const dgv_ctx* ctx, int64_t* nrec, int64_t* per_basis) {
  CHECK_CTX(ctx);
  if (!ctx->has_a) return set_error(const_cast<dgv_ctx*>(ctx), DGV_ERR_STATE, "no A-array set");
  if (nrec) *nrec = ctx->pb.nrec_ref;
  if (per_basis) std::copy(ctx->pb.nrec_basis.begin(), ctx->pb.nrec_basis.end(), per_basis);
  return DGV_OK;
}

int dgv_ctx_workload(const dgv_ctx* ctx, int64_t out[4]) {
  CHECK_CTX(ctx);
  dgv_ctx* c = const_cast<dgv_ctx*>(ctx);
  if (!ctx->has_a) return set_error(c, DGV_ERR_STATE, "no A-array set");
  if (!out) return set_error(c, DGV_ERR_ARG, "null output");
  DGV_HIP(c, hipSetDevice(ctx->device));
  return workload(c, out);
}

int dgv_ctx_get_aarray_basis(const dgv_ctx* ctx, int32_t l, double* A, int32_t* A_xi, int32_t* A_xi1, int64_t cap,
                             int64_t* nrec, int index_base) {
  CHECK_CTX(ctx);
  dgv_ctx* c = const_cast<dgv_ctx*>(ctx);
  if (!ctx->has_a) return set_error(c, DGV_ERR_STATE, "no A-array set");
  if (l < 0 || l >= ctx->g.D || !nrec || (index_base != 0 && index_base != 1))
    return set_error(c, DGV_ERR_ARG, "bad basis index, nrec pointer or index_base");
  int rc = get_basis_records(ctx, l, A, A_xi, A_xi1, cap, nrec, index_base);
  if (rc == DGV_ERR_ARG) return set_error(c, rc, "capacity smaller than the record count");
  if (rc) return set_error(c, rc, "A-array export failed");
  return DGV_OK;
}

int dgv_ctx_get_aarray_bases(const dgv_ctx* ctx, const int32_t* ls, int32_t nl, double* A, int32_t* A_xi,
                             int32_t* A_xi1, int64_t cap, int64_t* counts, int index_base) {
  CHECK_CTX(ctx);
  dgv_ctx* c = const_cast<dgv_ctx*>(ctx);
  if (!ctx->has_a) return set_error(c, DGV_ERR_STATE, "no A-array set");
  if (!ls || nl < 1 || !counts || (index_base != 0 && index_base != 1))
    return set_error(c, DGV_ERR_ARG, "bad basis list, counts pointer or index_base");
  int64_t total = 0;
  for (int k = 0; k < nl; ++k) {
    if (ls[k] < 0 || ls[k] >= ctx->g.D) return set_error(c, DGV_ERR_INDEX, "basis index outside 0..D-1");
    counts[k] = ctx->pb.nrec_basis[ls[k]];
    total += counts[k];
  }
  if (!A && !A_xi && !A_xi1) return DGV_OK;
  if (!A || !A_xi || !A_xi1) return set_error(c, DGV_ERR_ARG, "A, A_xi and A_xi1 are all needed");
  if (cap < total) return set_error(c, DGV_ERR_ARG, "capacity smaller than the record count");
  int rc = export_bases_ordered(ctx, ls, nl, A, A_xi, A_xi1, index_base);
  if (rc) return set_error(c, rc, "A-array export failed");
  return DGV_OK;
}

// WriteAarraysDGV / ReadAarraysDGV (readwrite.f90:1459-1484, 1066-1118): the
// file layout and the framing of records above 2^31 - 9 bytes are dgv_aio.cpp's.
int dgv_aarray_write_file(const dgv_ctx* ctx, const char* path) {
  CHECK_CTX(ctx);
  dgv_ctx* c = const_cast<dgv_ctx*>(ctx);
  if (!ctx->has_a) return set_error(c, DGV_ERR_STATE, "no A-array set");
  if (!path) return set_error(c, DGV_ERR_ARG, "null path");
  std::vector<int32_t> capphi, xi, xi1;
  std::vector<double> A;
  if (get_all_records(ctx, capphi, A, xi, xi1, 1)) return set_error(c, DGV_ERR_HIP, "A-array export failed");
  const int64_t mm = (int64_t)capphi.size(), nn = (int64_t)A.size();
  std::vector<int32_t> phi(nn);   // A_phi: the owning node of each record, 1-based
  for (int64_t k = 0, r = 0; k < mm; ++k)
    for (int32_t j = 0; j < capphi[k]; ++j, ++r) phi[r] = (int32_t)(k + 1);
  if (int rc = dgv_aarray_file_write(path, capphi.data(), mm, A.data(), xi.data(), xi1.data(), phi.data(), nn,
                                     DGV_FRAMING_SUBRECORDS))
    return set_error(c, rc, dgv_last_error(nullptr));
  return DGV_OK;
}

int dgv_ctx_read_aarray_file(dgv_ctx* ctx, const char* path) {
  CHECK_CTX(ctx);
  if (!path) return set_error(ctx, DGV_ERR_ARG, "null path");
  int64_t mm = 0, nn = 0;
  if (int rc = dgv_aarray_file_header(path, &mm, &nn)) return set_error(ctx, rc, dgv_last_error(nullptr));
  std::vector<int32_t> capphi(mm), xi(nn), xi1(nn), phi(nn);
  std::vector<double> A(nn);
  if (int rc = dgv_aarray_file_read(path, capphi.data(), mm, A.data(), xi.data(), xi1.data(), phi.data(), nn))
    return set_error(ctx, rc, dgv_last_error(nullptr));
  return build_pairblock_from_records(ctx, capphi.data(), mm, A.data(), xi.data(), xi1.data(), nn, 1);
}

int dgv_aarray_chunk_name(const dgv_grid_desc* grid, const char* dir, const char* base, int32_t chunk, int32_t k_c,
                          int32_t N_x, char* out, size_t cap) {
  if (!grid || !out || chunk < 0 || chunk > 999) { set_global_error("bad argument"); return DGV_ERR_ARG; }
  // MakeBaseNameAoperChnksDGV (readwrite.f90:2224-2281), uniform meshes
  char buf[1024];
  int n = std::snprintf(buf, sizeof(buf), "%s%sch%03d_%dkc%dsu%dsv%dsw%dNXU%dMuUU%dMvVU%dMwWU_Aarrs.dat",
                    dir ? dir : "", base ? base : "", chunk, k_c, grid->su, grid->sv, grid->sw, N_x, grid->Mu,
                    grid->Mv, grid->Mw);
  if (n < 0 || (size_t)n + 1 > cap) { set_global_error("output buffer too small"); return DGV_ERR_ARG; }
  std::memcpy(out, buf, (size_t)n + 1);
  return DGV_OK;
}

int dgv_ctx_set_kernel(dgv_ctx* ctx, int kernel) {
  CHECK_CTX(ctx);
  if (kernel < DGV_KERNEL_AUTO || kernel > DGV_KERNEL_MFMA) return set_error(ctx, DGV_ERR_ARG, "unknown kernel");
  ctx->kernel = kernel;
  if (ctx->rk_graph) { hipGraphExecDestroy(ctx->rk_graph); ctx->rk_graph = nullptr; }
  return DGV_OK;
}

int dgv_ctx_set_projection(dgv_ctx* ctx, int32_t on) {
  CHECK_CTX(ctx);
  ctx->project = on ? 1 : 0;
  if (ctx->rk_graph) { hipGraphExecDestroy(ctx->rk_graph); ctx->rk_graph = nullptr; }
  return DGV_OK;
}

int dgv_ctx_set_tile(dgv_ctx* ctx, int32_t pairs_per_item) {
  CHECK_CTX(ctx);
  if (pairs_per_item < 0) return set_error(ctx, DGV_ERR_ARG, "negative tile");
  if (pairs_per_item == 0) return DGV_OK;
  ctx->tile = pairs_per_item;
  if (ctx->has_a) return build_plan(ctx, pairs_per_item);
  return DGV_OK;
}

int dgv_eval_q_dev(dgv_ctx* ctx, const double* d_f, double* d_fc, int64_t n) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  return eval_q_dev(ctx, d_f, d_fc, 1.0);
}

int dgv_eval_q(dgv_ctx* ctx, const double* f, double* fc, int64_t n) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  if (!f || !fc) return set_error(ctx, DGV_ERR_ARG, "null array");
  const size_t nb = sizeof(double) * n;
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_f, f, nb, hipMemcpyHostToDevice, ctx->stream));
  int rc = eval_q_dev(ctx, ctx->d_f, ctx->d_q, 1.0);
  if (rc) return rc;
  DGV_HIP(ctx, hipMemcpyAsync(fc, ctx->d_q, nb, hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

int dgv_eval_q_a(dgv_ctx* ctx, const double* f, double* fc, int64_t n) { return dgv_eval_q(ctx, f, fc, n); }

int dgv_eval_mixed_dev(dgv_ctx* ctx, const double* d_f, const double* d_fm, double* d_fc, int64_t n) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  return eval_mixed_dev(ctx, d_f, d_fm, d_fc);
}

int dgv_eval_mixed(dgv_ctx* ctx, const double* f, const double* fm, double* fc, int64_t n) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  if (!f || !fm || !fc) return set_error(ctx, DGV_ERR_ARG, "null array");
  const size_t nb = sizeof(double) * n;
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_f, f, nb, hipMemcpyHostToDevice, ctx->stream));
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_tmp[5], fm, nb, hipMemcpyHostToDevice, ctx->stream));
  int rc = eval_mixed_dev(ctx, ctx->d_f, ctx->d_tmp[5], ctx->d_q);
  if (rc) return rc;
  DGV_HIP(ctx, hipMemcpyAsync(fc, ctx->d_q, nb, hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

int dgv_ctx_autotune(dgv_ctx* ctx, int32_t* best_pairs_per_item, double* best_ms) {
  CHECK_CTX(ctx);
  if (!ctx->has_a) return set_error(ctx, DGV_ERR_STATE, "no A-array set");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  // f = 1 everywhere is a valid input for timing
  std::vector<double> ones(ctx->g.N, 1.0);
  DGV_HIP(ctx, hipMemcpy(ctx->d_f, ones.data(), sizeof(double) * ones.size(), hipMemcpyHostToDevice));
  const int32_t cand[] = {512, 1024, 2048, 4096, 8192, 16384};
  int32_t best = 0;
  double bms = 1e300;
  for (int32_t t : cand) {
    int rc = build_plan(ctx, t);
    if (rc) return rc;
    rc = eval_q_dev(ctx, ctx->d_f, ctx->d_q, 1.0);  // warm-up
    if (rc) return rc;
    DGV_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    for (int k = 0; k < 3; ++k) {
      rc = eval_q_dev(ctx, ctx->d_f, ctx->d_q, 1.0);
      if (rc) return rc;
    }
    DGV_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    DGV_HIP(ctx, hipEventSynchronize(ctx->ev[1]));
    float ms = 0;
    DGV_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    if (ms / 3 < bms) { bms = ms / 3; best = t; }
  }
  ctx->tile = best;
  int rc = build_plan(ctx, best);
  if (rc) return rc;
  if (best_pairs_per_item) *best_pairs_per_item = best;
  if (best_ms) *best_ms = bms;
  return DGV_OK;
}

int dgv_spatial_oper_dev(dgv_ctx* ctx, const double* d_f, double* d_fcol, int64_t n, double dt, double coef) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  return spatial_oper_device(ctx, d_f, d_fcol, coef * dt);
}

int dgv_spatial_oper(dgv_ctx* ctx, const double* f, double* fcol, int64_t n, double dt, double coef) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  if (!f || !fcol) return set_error(ctx, DGV_ERR_ARG, "null array");
  const size_t nb = sizeof(double) * n;
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_f, f, nb, hipMemcpyHostToDevice, ctx->stream));
  int rc = spatial_oper_device(ctx, ctx->d_f, ctx->d_q, coef * dt);
  if (rc) return rc;
  DGV_HIP(ctx, hipMemcpyAsync(fcol, ctx->d_q, nb, hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

}  // extern "C"

namespace dgv {
// One TimeIntegratorORK_SH_DGV step on the device (time_integr_mod.f90:26-152).
static int rk_step_device(dgv_ctx* ctx, double* f, double dt, double coef, int rkmts, double* frhs1_out) {
  const int64_t n = ctx->g.N;
  double* k1 = ctx->d_tmp[0];
  double* ft2 = ctx->d_tmp[1];
  double* ft3 = ctx->d_tmp[2];
  double* ft4 = ctx->d_tmp[3];
  double* tmp = ctx->d_tmp[4];
  const double ct = coef * dt;
  int rc;
#define SOP(in, out) do { rc = spatial_oper_device(ctx, in, out, ct); if (rc) return rc; } while (0)
#define LC(out, N, ...)                                                                  \
  do {                                                                                   \
    const double* xs_[] = {__VA_ARGS__};                                                 \
    rc = launch_lincomb(ctx, n, out, N, xs_, m1_, m2_, den_, neg_);                      \
    if (rc) return rc;                                                                   \
  } while (0)
  switch (rkmts) {
    case 1: {
      SOP(f, k1);
      double m1_[] = {1, 1}, m2_[] = {1, 1}, den_[] = {1, 1};
      int neg_[] = {0, 0};
      LC(f, 2, f, k1);
      break;
    }
    case 2: {
      SOP(f, k1);
      {
        double m1_[] = {1, 2}, m2_[] = {1, 1}, den_[] = {1, 3.0};
        int neg_[] = {0, 0};
        LC(tmp, 2, f, k1);
      }
      SOP(tmp, ft2);
      {
        double m1_[] = {1, 1, 3.0}, m2_[] = {1, 1, 1}, den_[] = {1, 4.0, 4};
        int neg_[] = {0, 0, 0};
        LC(f, 3, f, k1, ft2);
      }
      break;
    }
    case 3: {
      SOP(f, k1);
      {
        double m1_[] = {1, 1}, m2_[] = {1, 1}, den_[] = {1, 2.0};
        int neg_[] = {0, 0};
        LC(tmp, 2, f, k1);
      }
      SOP(tmp, ft2);
      {
        double m1_[] = {1, 3.0}, m2_[] = {1, 1}, den_[] = {1, 4};
        int neg_[] = {0, 0};
        LC(tmp, 2, f, ft2);
      }
      SOP(tmp, ft3);
      {
        double m1_[] = {1, 2, 3.0, 4}, m2_[] = {1, 1, 1, 1}, den_[] = {1, 9.0, 9.0, 9.0};
        int neg_[] = {0, 0, 0, 0};
        LC(f, 4, f, k1, ft2, ft3);
      }
      break;
    }
    case 4: {
      SOP(f, k1);
      {
        double m1_[] = {1, 1}, m2_[] = {1, 1}, den_[] = {1, 2.0};
        int neg_[] = {0, 0};
        LC(tmp, 2, f, k1);
      }
      SOP(tmp, ft2);
      {
        double m1_[] = {1, 1, 1}, m2_[] = {1, 1, 1}, den_[] = {1, 4.0, 4.0};
        int neg_[] = {0, 0, 0};
        LC(tmp, 3, f, k1, ft2);
      }
      SOP(tmp, ft3);
      {
        double m1_[] = {1, 1, 2}, m2_[] = {1, 1, 1}, den_[] = {1, 1, 1};
        int neg_[] = {0, 1, 0};
        LC(tmp, 3, f, ft2, ft3);
      }
      SOP(tmp, ft2);
      {
        double m1_[] = {1, 1, 2, 1}, m2_[] = {1, 1, 1, 1}, den_[] = {1, 6.0, 3.0, 6.0};
        int neg_[] = {0, 0, 0, 0};
        LC(f, 4, f, k1, ft3, ft2);
      }
      break;
    }
    case 5: {
      SOP(f, k1);
      {
        double m1_[] = {1, 1}, m2_[] = {1, 1}, den_[] = {1, 2.0};
        int neg_[] = {0, 0};
        LC(tmp, 2, f, k1);
      }
      SOP(tmp, ft2);
      {
        double m1_[] = {1, 1, 1}, m2_[] = {1, 1, 1}, den_[] = {1, 4.0, 4.0};
        int neg_[] = {0, 0, 0};
        LC(tmp, 3, f, k1, ft2);
      }
      SOP(tmp, ft3);
      {
        double m1_[] = {1, 1, 2.0}, m2_[] = {1, 1, 1}, den_[] = {1, 1, 1};
        int neg_[] = {0, 1, 0};
        LC(tmp, 3, f, ft2, ft3);
      }
      SOP(tmp, ft4);
      {
        double m1_[] = {1, 546}, m2_[] = {1, 1}, den_[] = {5.0, 625.0};
        int neg_[] = {1, 0};
        LC(tmp, 2, ft2, ft3);  // -ft2/5 + ft3*546/625 (into tmp, then ft3)
      }
      std::swap(ctx->d_tmp[2], ctx->d_tmp[4]);
      ft3 = ctx->d_tmp[2];
      tmp = ctx->d_tmp[4];
      {
        double m1_[] = {1, 7.0, 10.0, 1}, m2_[] = {1, 1, 1, 1}, den_[] = {1, 27.0, 27.0, 27.0};
        int neg_[] = {0, 0, 0, 0};
        LC(tmp, 4, f, k1, ft2, ft4);
      }
      SOP(tmp, ft2);
      {
        double m1_[] = {1, 28.0, 1, 54, 378}, m2_[] = {1, 1, 1, 1, 1}, den_[] = {1, 625.0, 1, 625.0, 625.0};
        int neg_[] = {0, 0, 0, 0, 1};
        LC(tmp, 5, f, k1, ft3, ft4, ft2);
      }
      SOP(tmp, ft3);
      {
        double m1_[] = {1, 1, 5.0, 27.0, 125}, m2_[] = {1, 1, 1, 1, 1}, den_[] = {1, 24.0, 48.0, 56.0, 336.0};
        int neg_[] = {0, 0, 0, 0, 0};
        LC(f, 5, f, k1, ft4, ft2, ft3);
      }
      break;
    }
    default:
      return set_error(ctx, DGV_ERR_ARG, "Unsupported degree of Runge-Kutta method passed");
  }
  if (frhs1_out) {  // frhs1 = frhs1/dt (line 151)
    double m1_[] = {1}, m2_[] = {1}, den_[] = {dt};
    int neg_[] = {0};
    LC(frhs1_out, 1, k1);
  }
#undef SOP
#undef LC
  return DGV_OK;
}
}  // namespace dgv

extern "C" {

int dgv_rk_step_dev(dgv_ctx* ctx, double* d_f, int64_t n, double dt, double coef, int32_t rkmts, double* d_frhs1) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  if (rkmts < 1 || rkmts > 5) return set_error(ctx, DGV_ERR_ARG, "Unsupported degree of Runge-Kutta method passed");
  return rk_step_device(ctx, d_f, dt, coef, rkmts, d_frhs1);
}

int dgv_rk_step(dgv_ctx* ctx, double* f, int64_t n, double dt, double coef, int32_t rkmts, double* frhs1) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  if (!f) return set_error(ctx, DGV_ERR_ARG, "null array");
  if (rkmts < 1 || rkmts > 5) return set_error(ctx, DGV_ERR_ARG, "Unsupported degree of Runge-Kutta method passed");
  const size_t nb = sizeof(double) * n;
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_f, f, nb, hipMemcpyHostToDevice, ctx->stream));
  int rc = rk_step_device(ctx, ctx->d_f, dt, coef, rkmts, frhs1 ? ctx->d_q : nullptr);
  if (rc) return rc;
  DGV_HIP(ctx, hipMemcpyAsync(f, ctx->d_f, nb, hipMemcpyDeviceToHost, ctx->stream));
  if (frhs1) DGV_HIP(ctx, hipMemcpyAsync(frhs1, ctx->d_q, nb, hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

int dgv_rk_run_dev(dgv_ctx* ctx, double* d_f, int64_t n, double dt, double coef, int32_t rkmts, int64_t nsteps,
                   int use_graph) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  if (rkmts < 1 || rkmts > 5) return set_error(ctx, DGV_ERR_ARG, "Unsupported degree of Runge-Kutta method passed");
  if (!ctx->has_a) return set_error(ctx, DGV_ERR_STATE, "no A-array set");
  if (ctx->kernel == DGV_KERNEL_RECORD) {
    int rc = ensure_record_layout(ctx);
    if (rc) return rc;
  }
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  // RK5 swaps scratch buffers between steps; the regime check decides on the
  // host at every stage
  if (!use_graph || rkmts == 5 || ctx->lin.decomp_lev > 0.0) {
    for (int64_t s = 0; s < nsteps; ++s) {
      int rc = rk_step_device(ctx, d_f, dt, coef, rkmts, nullptr);
      if (rc) return rc;
    }
    return DGV_OK;
  }
  const bool same = ctx->rk_graph && ctx->rk_graph_key[0] == rkmts && ctx->rk_graph_key[1] == ctx->plan.pairs_per_item &&
                    ctx->rk_graph_dt == dt && ctx->rk_graph_coef == coef && ctx->rk_graph_f == d_f;
  if (!same) {
    if (ctx->rk_graph) { hipGraphExecDestroy(ctx->rk_graph); ctx->rk_graph = nullptr; }
    hipGraph_t graph;
    DGV_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = rk_step_device(ctx, d_f, dt, coef, rkmts, nullptr);
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc) return rc;
    if (e != hipSuccess) return set_error(ctx, DGV_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&ctx->rk_graph, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return set_error(ctx, DGV_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    ctx->rk_graph_key[0] = rkmts;
    ctx->rk_graph_key[1] = ctx->plan.pairs_per_item;
    ctx->rk_graph_dt = dt;
    ctx->rk_graph_coef = coef;
    ctx->rk_graph_f = d_f;
  }
  for (int64_t s = 0; s < nsteps; ++s) DGV_HIP(ctx, hipGraphLaunch(ctx->rk_graph, ctx->stream));
  return DGV_OK;
}

int dgv_ab_step(dgv_ctx* ctx, double* f, int64_t n, double dt, double coef, int32_t rkmts, double* hist) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  if (!f || !hist) return set_error(ctx, DGV_ERR_ARG, "null array");
  if (rkmts < 1 || rkmts > 5)
    return set_error(ctx, DGV_ERR_ARG, "Unsupported degree of Multiple Time Stepping method passed");
  // TimeIntegratorMTS_SH_DGV (time_integr_mod.f90:314-364): push the history
  // on the host, one spatial operator on the device, the AB combination on the
  // device, frhs1 = frhs1/dt.
  double* h[9];
  for (int k = 0; k < 9; ++k) h[k] = hist + (int64_t)k * n;
  const size_t nb = sizeof(double) * n;
  if (rkmts >= 5) { std::memcpy(h[3], h[2], nb); std::memcpy(h[8], h[7], nb); }
  if (rkmts >= 4) { std::memcpy(h[2], h[1], nb); std::memcpy(h[7], h[6], nb); }
  if (rkmts >= 3) { std::memcpy(h[1], h[0], nb); std::memcpy(h[6], h[5], nb); }
  if (rkmts >= 2) { std::memcpy(h[0], f, nb); std::memcpy(h[5], h[4], nb); }
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  double* d_r[5] = {ctx->d_tmp[0], ctx->d_tmp[1], ctx->d_tmp[2], ctx->d_tmp[3], ctx->d_tmp[4]};
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_f, f, nb, hipMemcpyHostToDevice, ctx->stream));
  for (int k = 1; k < rkmts; ++k)
    DGV_HIP(ctx, hipMemcpyAsync(d_r[k], h[4 + k], nb, hipMemcpyHostToDevice, ctx->stream));
  int rc = spatial_oper_device(ctx, ctx->d_f, d_r[0], coef * dt);
  if (rc) return rc;
  const double* xs[6] = {ctx->d_f, d_r[0], d_r[1], d_r[2], d_r[3], d_r[4]};
  double m1[6] = {1, 1, dt, dt, dt, dt}, m2[6] = {1, 1, 1, 1, 1, 1}, den[6] = {1, 1, 1, 1, 1, 1};
  int neg[6] = {0, 0, 0, 0, 0, 0};
  switch (rkmts) {
    case 1: break;
    case 2: m1[1] = 3.0; den[1] = 2.0; den[2] = 2.0; neg[2] = 1; break;
    case 3:
      m1[1] = 23.0; den[1] = 12.0; m2[2] = 4; den[2] = 3.0; neg[2] = 1; m2[3] = 5.0; den[3] = 12.0;
      break;
    case 4:
      m1[1] = 55.0; den[1] = 24.0; m2[2] = 59.0; den[2] = 24.0; neg[2] = 1; m2[3] = 37.0; den[3] = 24.0;
      m2[4] = 3.0; den[4] = 8; neg[4] = 1;
      break;
    case 5:
      m1[1] = 1901.0 / 720.0; m2[2] = 2774.0 / 720.0; neg[2] = 1; m2[3] = 2616.0 / 720.0;
      m2[4] = 1274.0 / 720.0; neg[4] = 1; m2[5] = 251.0 / 720.0;
      break;
  }
  rc = launch_lincomb(ctx, n, ctx->d_q, rkmts + 1, xs, m1, m2, den, neg);
  if (rc) return rc;
  DGV_HIP(ctx, hipMemcpyAsync(f, ctx->d_q, nb, hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipMemcpyAsync(h[4], d_r[0], nb, hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int64_t i = 0; i < n; ++i) h[4][i] = h[4][i] / dt;  // line 362
  return DGV_OK;
}

int dgv_moments(dgv_ctx* ctx, const double* f, int64_t n, double out[5]) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  if (!f || !out) return set_error(ctx, DGV_ERR_ARG, "null array");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_f, f, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  return device_moments(ctx, ctx->d_f, out);
}

int dgv_moments_high(dgv_ctx* ctx, const double* f, int64_t n, double out[20]) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  if (!f || !out) return set_error(ctx, DGV_ERR_ARG, "null array");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_f, f, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  return device_moments_high(ctx, ctx->d_f, out);
}

int dgv_moments_high_dev(dgv_ctx* ctx, const double* d_f, int64_t n, double out[20]) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  if (!d_f || !out) return set_error(ctx, DGV_ERR_ARG, "null array");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  return device_moments_high(ctx, d_f, out);
}

int dgv_collision_moments(dgv_ctx* ctx, const double* q, int64_t n, double out[5]) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  if (!q || !out) return set_error(ctx, DGV_ERR_ARG, "null array");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_q, q, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  return device_collision_moments(ctx, ctx->d_q, out);
}

int dgv_project_invariants_dev(dgv_ctx* ctx, double* d_q, int64_t n) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  return device_project(ctx, d_q);
}

int dgv_project_invariants(dgv_ctx* ctx, double* q, int64_t n) {
  CHECK_CTX(ctx);
  CHECK_N(ctx, n);
  if (!q) return set_error(ctx, DGV_ERR_ARG, "null array");
  const size_t nb = sizeof(double) * n;
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMemcpyAsync(ctx->d_q, q, nb, hipMemcpyHostToDevice, ctx->stream));
  int rc = device_project(ctx, ctx->d_q);
  if (rc) return rc;
  DGV_HIP(ctx, hipMemcpyAsync(q, ctx->d_q, nb, hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

int dgv_ctx_synchronize(dgv_ctx* ctx) {
  CHECK_CTX(ctx);
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

int dgv_time_eval_q_dev(dgv_ctx* ctx, const double* d_f, double* d_fc, int32_t niter, double* ms_per_call,
                        double* ms_main_kernel) {
  CHECK_CTX(ctx);
  if (niter < 1) return set_error(ctx, DGV_ERR_ARG, "niter < 1");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  int rc;
  DGV_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  for (int i = 0; i < niter; ++i)
    if ((rc = eval_q_dev(ctx, d_f, d_fc, 1.0))) return rc;
  DGV_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  for (int i = 0; i < niter; ++i)
    if ((rc = launch_eval_main_only(ctx, d_f))) return rc;
  DGV_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
  DGV_HIP(ctx, hipEventSynchronize(ctx->ev[2]));
  float a = 0, b = 0;
  DGV_HIP(ctx, hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]));
  DGV_HIP(ctx, hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]));
  if (ms_per_call) *ms_per_call = a / niter;
  if (ms_main_kernel) *ms_main_kernel = b / niter;
  return DGV_OK;
}

int dgv_device_alloc(dgv_ctx* ctx, size_t bytes, void** dptr) {
  CHECK_CTX(ctx);
  if (!dptr) return set_error(ctx, DGV_ERR_ARG, "null pointer");
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  if (hipMalloc(dptr, bytes) != hipSuccess) return set_error(ctx, DGV_ERR_NOMEM, "hipMalloc failed");
  return DGV_OK;
}
int dgv_device_free(dgv_ctx* ctx, void* dptr) {
  CHECK_CTX(ctx);
  DGV_HIP(ctx, hipFree(dptr));
  return DGV_OK;
}
int dgv_memcpy_htod(dgv_ctx* ctx, void* dst, const void* src, size_t bytes) {
  CHECK_CTX(ctx);
  DGV_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}
int dgv_memcpy_dtoh(dgv_ctx* ctx, void* dst, const void* src, size_t bytes) {
  CHECK_CTX(ctx);
  DGV_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

int dgv_device_copy(dgv_ctx* ctx, void* dst, const void* src, size_t bytes) {
  CHECK_CTX(ctx);
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  int rc = launch_copy16(ctx, dst, src, bytes);
  if (rc) return rc;
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

}  // extern "C"
