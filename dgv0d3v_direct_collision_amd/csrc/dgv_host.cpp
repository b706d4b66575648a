// dgv_host.cpp -- host side of the MI355X direct DGV collision library:
// geometry, A-array layout conversion, work planning, file I/O.
#include <atomic>
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>

#include "dgv_internal.h"

namespace dgv {

static thread_local std::string g_last_error;

int set_error(dgv_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  g_last_error = msg;
  // a failed hipMalloc also leaves HIP's sticky last error set: consume it here,
  // where it is reported, so that the next launch check does not blame it
  if (code == DGV_ERR_NOMEM) (void)hipGetLastError();
  return code;
}
void set_global_error(const std::string& msg) { g_last_error = msg; }
const char* global_error() { return g_last_error.c_str(); }

// Gauss-Legendre nodes/weights on [-1,1] by Newton's method on P_n, the roots
// refined independently from cos(pi (i - 1/4) / (n + 1/2)) until the update is
// below 3e-14 -- the same iteration as GauLeg (gaussian_mod.f90:4-53), so the
// weights agree with the reference's to the last bits.
int gauss_legendre(int n, double* x, double* w) {
  if (n == 1) {  // SetDGblzmGnodes special case (miscset.f90:76-77)
    x[0] = 0.0;
    w[0] = 2.0;
    return 0;
  }
  const double pi = 3.141592653589793238462643383279502884197;
  const int m = (n + 1) / 2;
  for (int i = 1; i <= m; ++i) {
    double z = std::cos(pi * ((double)i - 0.25) / ((double)n + 0.5));
    double pp = 0.0;
    bool conv = false;
    for (int it = 0; it < 40 && !conv; ++it) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      const double zold = z;
      z = zold - p1 / pp;
      conv = !(std::fabs(z - zold) > 3.0e-14);
    }
    if (!conv) return -1;
    x[i - 1] = -z;
    x[n - i] = z;
    w[i - 1] = 2.0 / ((1.0 - z * z) * (pp * pp));
    w[n - i] = w[i - 1];
  }
  return 0;
}

// First index (1-based) of a uniform 1-D grid whose point lies above x by more
// than the 1e-10 tolerance of finduiviwi (dgvtools_mod.f90:302-337); 0 when x
// is off the grid.
static int locate(double lo, double hi, int M, double x) {
  const double eps = 10e-11;
  std::vector<double> g(M + 1);
  for (int i = 0; i <= M; ++i) g[i] = lo + (hi - lo) / (double)M * i;
  if (x < g[0] - eps || x > g[M] + eps) return 0;
  int k = 1;
  while (x >= g[k - 1] - eps && k < M + 1) ++k;
  return k;
}

int build_grid(const dgv_grid_desc& d, Grid& g, std::string& err) {
  if (d.Mu < 1 || d.Mv < 1 || d.Mw < 1) { err = "number of cells must be >= 1"; return DGV_ERR_ARG; }
  if (d.su < 1 || d.sv < 1 || d.sw < 1 || d.su > kMaxS || d.sv > kMaxS || d.sw > kMaxS) {
    err = "nodes per cell must be in 1..8";
    return DGV_ERR_GRID;
  }
  if (!(d.u_R > d.u_L) || !(d.v_R > d.v_L) || !(d.w_R > d.w_L)) {
    err = "velocity box must have right endpoint > left endpoint";
    return DGV_ERR_ARG;
  }
  const int64_t nc = (int64_t)d.Mu * d.Mv * d.Mw;
  const int64_t D = (int64_t)d.su * d.sv * d.sw;
  if (nc * D > (int64_t)INT32_MAX) { err = "grid too large for 32-bit node ids"; return DGV_ERR_GRID; }
  g.Mu = d.Mu; g.Mv = d.Mv; g.Mw = d.Mw;
  g.su = d.su; g.sv = d.sv; g.sw = d.sw;
  g.D = (int)D;
  g.box[0] = d.u_L; g.box[1] = d.u_R; g.box[2] = d.v_L; g.box[3] = d.v_R;
  g.box[4] = d.w_L; g.box[5] = d.w_R;
  g.ncells = nc;
  g.N = nc * D;
  const int s[3] = {d.su, d.sv, d.sw};
  for (int k = 0; k < 3; ++k)
    if (gauss_legendre(s[k], g.gn[k], g.gw[k])) { err = "Gauss-Legendre iteration failed"; return DGV_ERR_GRID; }
  // cells: u outer, w inner (FillCellsArrsDGV, miscset.f90:915-927)
  auto edge = [](double lo, double hi, int M, int i) { return lo + (hi - lo) / (double)M * i; };
  g.lu.resize(nc); g.ru.resize(nc); g.lv.resize(nc); g.rv.resize(nc); g.lw.resize(nc); g.rw.resize(nc);
  int64_t c = 0;
  for (int iu = 0; iu < d.Mu; ++iu)
    for (int iv = 0; iv < d.Mv; ++iv)
      for (int iw = 0; iw < d.Mw; ++iw, ++c) {
        g.lu[c] = edge(d.u_L, d.u_R, d.Mu, iu); g.ru[c] = edge(d.u_L, d.u_R, d.Mu, iu + 1);
        g.lv[c] = edge(d.v_L, d.v_R, d.Mv, iv); g.rv[c] = edge(d.v_L, d.v_R, d.Mv, iv + 1);
        g.lw[c] = edge(d.w_L, d.w_R, d.Mw, iw); g.rw[c] = edge(d.w_L, d.w_R, d.Mw, iw + 1);
      }
  // nodes: cell outer, local u->v->w inner (SetNodesDGV, miscset.f90:1150-1176)
  g.u.resize(g.N); g.v.resize(g.N); g.w.resize(g.N); g.gwts.resize(g.N);
  int64_t n = 0;
  for (c = 0; c < nc; ++c) {
    const double du = g.ru[c] - g.lu[c], dv = g.rv[c] - g.lv[c], dw = g.rw[c] - g.lw[c];
    for (int a = 0; a < d.su; ++a)
      for (int b = 0; b < d.sv; ++b)
        for (int e = 0; e < d.sw; ++e, ++n) {
          g.u[n] = g.gn[0][a] * du / 2.0 + (g.ru[c] + g.lu[c]) / 2.0;
          g.v[n] = g.gn[1][b] * dv / 2.0 + (g.rv[c] + g.lv[c]) / 2.0;
          g.w[n] = g.gn[2][e] * dw / 2.0 + (g.rw[c] + g.lw[c]) / 2.0;
          g.gwts[n] = g.gw[0][a] * g.gw[1][b] * g.gw[2][e] * du * dv * dw / 8.0;
        }
  }
  // canonical cell: the cell containing the origin (miscmpiset.f90:553)
  const int ku = locate(d.u_L, d.u_R, d.Mu, 0.0), kv = locate(d.v_L, d.v_R, d.Mv, 0.0),
            kw = locate(d.w_L, d.w_R, d.Mw, 0.0);
  if (ku == 0 || kv == 0 || kw == 0) {
    err = "the velocity box does not contain the origin: no canonical cell";
    return DGV_ERR_GRID;
  }
  g.ccu = ku - 2; g.ccv = kv - 2; g.ccw = kw - 2;
  g.ccell = (g.ccu * d.Mv + g.ccv) * d.Mw + g.ccw;
  return DGV_OK;
}

void bump_aarray_generation(dgv_ctx* ctx) {
  static std::atomic<uint64_t> counter{0};
  ctx->a_gen = ++counter;
}

void free_pairblock(dgv_ctx* ctx) {
  bump_aarray_generation(ctx);
  free_linear(ctx);
  PairBlock& pb = ctx->pb;
  hipFree(pb.d_pairs); hipFree(pb.d_blocks); hipFree(pb.d_group_blocks); hipFree(pb.d_groups);
  pb.d_pairs = nullptr; pb.d_blocks = nullptr; pb.d_group_blocks = nullptr; pb.d_groups = nullptr;
  pb.blocks.clear(); pb.groups.clear(); pb.group_blocks.clear();
  pb.npairs = 0; pb.nrec_ref = 0; pb.nrec_basis.clear();
  hipFree(ctx->d_rec_off); hipFree(ctx->d_rec_A); hipFree(ctx->d_rec_xi); hipFree(ctx->d_rec_xi1);
  ctx->d_rec_off = nullptr; ctx->d_rec_A = nullptr; ctx->d_rec_xi = nullptr; ctx->d_rec_xi1 = nullptr;
  ctx->has_rec = false;
  ctx->has_a = false;
  ctx->exp_valid = false;
  ctx->exp_idx.clear();
  ctx->exp_val.clear();
}

void free_plan(dgv_ctx* ctx) {
  Plan& p = ctx->plan;
  hipFree(p.d_items); hipFree(p.d_partial); hipFree(p.d_cell_ptr); hipFree(p.d_cell_rows);
  p.d_items = nullptr; p.d_partial = nullptr; p.d_cell_ptr = nullptr; p.d_cell_rows = nullptr;
  p.items.clear(); p.nrows = 0; p.ncell_rows = 0; p.pairs_per_item = 0;
  if (ctx->rk_graph) { hipGraphExecDestroy(ctx->rk_graph); ctx->rk_graph = nullptr; }
}

// Block id of a (a, b) cell-offset pair; offsets lie in [-(M-1), M-1].
struct BlockKey {
  int8_t a[3], b[3];
  bool operator<(const BlockKey& o) const {
    return std::lexicographical_compare(a, a + 3, o.a, o.a + 3) ||
           (std::equal(a, a + 3, o.a) && std::lexicographical_compare(b, b + 3, o.b, o.b + 3));
  }
};

// Sorts (block, p, q) pair keys and fills the pair-block structures from a
// list of (key, l, A) entries.  Entries with equal (block, p, q, l) are summed.
namespace {
struct Entry {
  uint64_t key;  // block index (upper bits) | p << 6 | q
  double A;
  uint8_t l;
};
}  // namespace

static uint64_t encode_block(const Grid& g, const int a[3], const int b[3]) {
  const int64_t R[3] = {2 * g.Mu - 1, 2 * g.Mv - 1, 2 * g.Mw - 1};
  uint64_t k = 0;
  for (int d = 0; d < 3; ++d) k = k * R[d] + (uint64_t)(a[d] + (R[d] - 1) / 2);
  for (int d = 0; d < 3; ++d) k = k * R[d] + (uint64_t)(b[d] + (R[d] - 1) / 2);
  return k;
}
static void decode_block(const Grid& g, uint64_t k, int a[3], int b[3]) {
  const int64_t R[3] = {2 * g.Mu - 1, 2 * g.Mv - 1, 2 * g.Mw - 1};
  for (int d = 2; d >= 0; --d) { b[d] = (int)(k % R[d]) - (int)((R[d] - 1) / 2); k /= R[d]; }
  for (int d = 2; d >= 0; --d) { a[d] = (int)(k % R[d]) - (int)((R[d] - 1) / 2); k /= R[d]; }
}

// Valid box of output cells of a block: both c+a and c+b on the grid.
static bool block_box(const Grid& g, const int a[3], const int b[3], int lo[3], int ext[3]) {
  const int M[3] = {g.Mu, g.Mv, g.Mw};
  for (int d = 0; d < 3; ++d) {
    const int l = std::max(0, std::max(-a[d], -b[d]));
    const int h = std::min(M[d], std::min(M[d] - a[d], M[d] - b[d]));
    lo[d] = l;
    ext[d] = h - l;
    if (ext[d] <= 0) return false;
  }
  return true;
}

// Builds blocks/groups from pairs sorted by (block key, p, q).  `pairs` holds
// npairs*(D+1) doubles, `pkeys` the sorted 64-bit keys of each pair.
static int finish_pairblock(dgv_ctx* ctx, std::vector<uint64_t>& pkeys, std::vector<double>& pairs) {
  const Grid& g = ctx->g;
  const int D = g.D;
  PairBlock& pb = ctx->pb;
  pb.D = D;
  const int64_t np = (int64_t)pkeys.size();
  // blocks = runs of equal block id; drop blocks with an empty box
  std::vector<BlockHdr> blocks;
  std::vector<uint64_t> bkeys;
  std::vector<double> kept;
  kept.reserve(pairs.size());
  int64_t out_pairs = 0;
  int64_t i = 0;
  while (i < np) {
    const uint64_t bk = pkeys[i] >> 12;
    int64_t j = i;
    while (j < np && (pkeys[j] >> 12) == bk) ++j;
    int a[3], b[3], lo[3], ext[3];
    decode_block(g, bk, a, b);
    if (block_box(g, a, b, lo, ext)) {
      BlockHdr h;
      std::memset(&h, 0, sizeof(h));
      h.pbeg = out_pairs;
      for (int d = 0; d < 3; ++d) { h.a[d] = (int8_t)a[d]; h.b[d] = (int8_t)b[d]; }
      h.npairs = (int16_t)(j - i);
      // seg[p] = first pair with local index of xi >= p
      int64_t k = i;
      for (int p = 0; p <= D; ++p) {
        while (k < j && (int)((pkeys[k] >> 6) & 63) < p) ++k;
        h.seg[p] = (int32_t)(k - i);
      }
      h.seg[D] = (int32_t)(j - i);
      blocks.push_back(h);
      bkeys.push_back(bk);
      kept.insert(kept.end(), pairs.begin() + i * (D + 1), pairs.begin() + j * (D + 1));
      out_pairs += j - i;
    }
    i = j;
  }
  pairs.swap(kept);
  pb.npairs = out_pairs;
  // groups by identical box
  std::map<std::array<int, 6>, int> gid;
  std::vector<std::vector<int32_t>> members;
  for (size_t b = 0; b < blocks.size(); ++b) {
    int a[3], bb[3], lo[3], ext[3];
    for (int d = 0; d < 3; ++d) { a[d] = blocks[b].a[d]; bb[d] = blocks[b].b[d]; }
    block_box(g, a, bb, lo, ext);
    std::array<int, 6> key = {lo[0], lo[1], lo[2], ext[0], ext[1], ext[2]};
    auto it = gid.find(key);
    if (it == gid.end()) {
      it = gid.emplace(key, (int)members.size()).first;
      members.emplace_back();
    }
    members[it->second].push_back((int32_t)b);
  }
  pb.groups.clear();
  pb.group_blocks.clear();
  for (auto& kv : gid) {
    const auto& key = kv.first;
    Group G;
    for (int d = 0; d < 3; ++d) { G.lo[d] = key[d]; G.ext[d] = key[3 + d]; }
    G.ncells = G.ext[0] * G.ext[1] * G.ext[2];
    G.bbeg = (int32_t)pb.group_blocks.size();
    G.npairs = 0;
    for (int32_t b : members[kv.second]) {
      pb.group_blocks.push_back(b);
      G.npairs += blocks[b].npairs;
    }
    G.bend = (int32_t)pb.group_blocks.size();
    pb.groups.push_back(G);
  }
  pb.blocks.swap(blocks);
  // record counts of the reference layout
  pb.nrec_basis.assign(D, 0);
  for (int64_t p = 0; p < pb.npairs; ++p)
    for (int l = 0; l < D; ++l)
      if (pairs[p * (D + 1) + l] != 0.0) pb.nrec_basis[l]++;
  pb.nrec_ref = std::accumulate(pb.nrec_basis.begin(), pb.nrec_basis.end(), (int64_t)0);
  return upload_pairblock(ctx, pairs);
}

int upload_pairblock(dgv_ctx* ctx, std::vector<double>& pairs) {
  PairBlock& pb = ctx->pb;
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  const size_t pad = (size_t)kPairPad * (pb.D + 1);
  if (hipMalloc(&pb.d_pairs, sizeof(double) * (pairs.size() + pad)) != hipSuccess)
    return set_error(ctx, DGV_ERR_NOMEM, "hipMalloc(pairs) failed");
  DGV_HIP(ctx, hipMemcpy(pb.d_pairs, pairs.data(), sizeof(double) * pairs.size(), hipMemcpyHostToDevice));
  const std::vector<double> zeros(pad, 0.0);
  DGV_HIP(ctx, hipMemcpy(pb.d_pairs + pairs.size(), zeros.data(), sizeof(double) * pad, hipMemcpyHostToDevice));
  DGV_HIP(ctx, hipMalloc(&pb.d_blocks, sizeof(BlockHdr) * std::max<size_t>(pb.blocks.size(), 1)));
  DGV_HIP(ctx, hipMemcpy(pb.d_blocks, pb.blocks.data(), sizeof(BlockHdr) * pb.blocks.size(), hipMemcpyHostToDevice));
  DGV_HIP(ctx, hipMalloc(&pb.d_group_blocks, sizeof(int32_t) * std::max<size_t>(pb.group_blocks.size(), 1)));
  DGV_HIP(ctx, hipMemcpy(pb.d_group_blocks, pb.group_blocks.data(), sizeof(int32_t) * pb.group_blocks.size(),
                         hipMemcpyHostToDevice));
  DGV_HIP(ctx, hipMalloc(&pb.d_groups, sizeof(Group) * std::max<size_t>(pb.groups.size(), 1)));
  DGV_HIP(ctx, hipMemcpy(pb.d_groups, pb.groups.data(), sizeof(Group) * pb.groups.size(), hipMemcpyHostToDevice));
  ctx->has_a = true;
  return build_plan(ctx, ctx->tile);
}

int build_pairblock_from_records(dgv_ctx* ctx, const int32_t* capphi, int64_t ncap, const double* A,
                                 const int32_t* xi, const int32_t* xi1, int64_t nrec, int index_base) {
  const Grid& g = ctx->g;
  const int D = g.D;
  if (ncap != g.N) return set_error(ctx, DGV_ERR_ARG, "A_capphi length differs from the number of nodes");
  if (index_base != 0 && index_base != 1) return set_error(ctx, DGV_ERR_ARG, "index_base must be 0 or 1");
  if (D > 64) return set_error(ctx, DGV_ERR_UNSUPPORTED, "more than 64 nodes per cell");
  if (nrec < 0 || (nrec > 0 && (!A || !xi || !xi1)) || !capphi)
    return set_error(ctx, DGV_ERR_ARG, "null A-array pointer");
  std::vector<int64_t> pref(g.N + 1, 0);
  for (int64_t k = 0; k < g.N; ++k) {
    if (capphi[k] < 0) return set_error(ctx, DGV_ERR_ARG, "negative A_capphi entry");
    pref[k + 1] = pref[k] + capphi[k];
  }
  if (pref[g.N] != nrec) return set_error(ctx, DGV_ERR_ARG, "sum(A_capphi) differs from the record count");
  // node -> (cell coords, local)
  auto cell_of = [&](int64_t node, int c3[3]) {
    const int64_t cell = node / D;
    c3[2] = (int)(cell % g.Mw);
    c3[1] = (int)((cell / g.Mw) % g.Mv);
    c3[0] = (int)(cell / ((int64_t)g.Mw * g.Mv));
  };
  std::vector<Entry> ent;
  int64_t used = 0;
  for (int l = 0; l < D; ++l) used += capphi[(int64_t)g.ccell * D + l];
  ent.reserve(used);
  for (int l = 0; l < D; ++l) {
    const int64_t k = (int64_t)g.ccell * D + l;
    for (int64_t r = pref[k]; r < pref[k + 1]; ++r) {
      const int64_t x = (int64_t)xi[r] - index_base, y = (int64_t)xi1[r] - index_base;
      if (x < 0 || x >= g.N || y < 0 || y >= g.N)
        return set_error(ctx, DGV_ERR_INDEX, "A_xi/A_xi1 entry out of range at record " + std::to_string(r + index_base));
      int cx[3], cy[3], a[3], b[3];
      cell_of(x, cx);
      cell_of(y, cy);
      const int cc[3] = {g.ccu, g.ccv, g.ccw};
      for (int d = 0; d < 3; ++d) { a[d] = cx[d] - cc[d]; b[d] = cy[d] - cc[d]; }
      Entry e;
      e.key = (encode_block(g, a, b) << 12) | ((uint64_t)(x % D) << 6) | (uint64_t)(y % D);
      e.A = A[r];
      e.l = (uint8_t)l;
      ent.push_back(e);
    }
  }
  std::stable_sort(ent.begin(), ent.end(), [](const Entry& p, const Entry& q) { return p.key < q.key; });
  std::vector<uint64_t> pkeys;
  std::vector<double> pairs;
  for (size_t i = 0; i < ent.size();) {
    size_t j = i;
    pkeys.push_back(ent[i].key);
    const size_t base = pairs.size();
    pairs.resize(base + D + 1, 0.0);
    while (j < ent.size() && ent[j].key == ent[i].key) {
      pairs[base + ent[j].l] += ent[j].A;
      ++j;
    }
    int64_t qbits = pair_code_bits((int)((ent[i].key >> 6) & 63), (int)(ent[i].key & 63));
    std::memcpy(&pairs[base + D], &qbits, sizeof(double));
    i = j;
  }
  std::vector<Entry>().swap(ent);
  free_plan(ctx);
  free_pairblock(ctx);
  return finish_pairblock(ctx, pkeys, pairs);
}

// Work items: every group's box is cut into 64-cell chunks and its block list
// into runs of about `ppi` pairs; one wave per (chunk, run).  Partial sums of
// item i live in rows [64 i, 64 i + 64); the per-cell row lists are in item
// order, which fixes the order of the final FP64 sum.
int build_plan(dgv_ctx* ctx, int32_t ppi) {
  free_plan(ctx);
  const PairBlock& pb = ctx->pb;
  const Grid& g = ctx->g;
  Plan& P = ctx->plan;
  if (ppi <= 0) ppi = 2048;
  P.pairs_per_item = ppi;
  std::vector<WorkItem> items;
  std::vector<int64_t> item_pairs;
  for (size_t gi = 0; gi < pb.groups.size(); ++gi) {
    const Group& G = pb.groups[gi];
    const int nchunk = (G.ncells + 63) / 64;
    int32_t b0 = G.bbeg;
    while (b0 < G.bend) {
      int32_t b1 = b0;
      int64_t acc = 0;
      while (b1 < G.bend && (b1 == b0 || acc + pb.blocks[pb.group_blocks[b1]].npairs <= ppi)) {
        acc += pb.blocks[pb.group_blocks[b1]].npairs;
        ++b1;
      }
      for (int ch = 0; ch < nchunk; ++ch) {
        WorkItem it;
        it.group = (int32_t)gi;
        it.chunk = ch;
        it.bbeg = b0;
        it.bend = b1;
        it.prow = 0;
        items.push_back(it);
        item_pairs.push_back(acc);
      }
      b0 = b1;
    }
  }
  // XCD-aware order.  Workgroup i is dispatched to XCD i % 8, and the chunks
  // of one (group, block run) read the same pair records, so they are placed
  // on one XCD (ids x, x+8, x+16, ...) to share its L2.  Runs are dealt to the
  // least-loaded XCD, largest first; columns are padded with empty slots.
  constexpr int kXcd = 8;
  struct Run { size_t first; int n; int64_t work; };
  std::vector<Run> runs;
  for (size_t i = 0; i < items.size();) {
    size_t j = i;
    while (j < items.size() && items[j].group == items[i].group && items[j].bbeg == items[i].bbeg) ++j;
    runs.push_back({i, (int)(j - i), item_pairs[i] * (int64_t)(j - i)});
    i = j;
  }
  std::stable_sort(runs.begin(), runs.end(), [](const Run& x, const Run& y) { return x.work > y.work; });
  std::vector<std::vector<size_t>> col(kXcd);
  std::vector<int64_t> load(kXcd, 0);
  for (const Run& r : runs) {
    int c = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    for (int k = 0; k < r.n; ++k) col[c].push_back(r.first + k);
    load[c] += r.work;
  }
  size_t nslot = 0;
  for (auto& cvec : col) nslot = std::max(nslot, cvec.size());
  P.items.assign(nslot * kXcd, WorkItem{0, -1, 0, 0, 0});
  for (size_t rr = 0; rr < nslot; ++rr)
    for (int c = 0; c < kXcd; ++c) {
      const size_t pos = rr * kXcd + c;
      if (rr < col[c].size()) P.items[pos] = items[col[c][rr]];
      P.items[pos].prow = (int64_t)pos * 64;
    }
  P.nrows = (int64_t)P.items.size() * 64;
  // per-cell row lists
  std::vector<int64_t> cnt(g.ncells + 1, 0);
  for (const WorkItem& it : P.items) {
    if (it.chunk < 0) continue;
    const Group& G = pb.groups[it.group];
    const int n = std::min(64, G.ncells - it.chunk * 64);
    for (int l = 0; l < n; ++l) {
      int r = it.chunk * 64 + l;
      const int cw = G.lo[2] + r % G.ext[2]; r /= G.ext[2];
      const int cv = G.lo[1] + r % G.ext[1]; r /= G.ext[1];
      const int cu = G.lo[0] + r;
      cnt[((int64_t)cu * g.Mv + cv) * g.Mw + cw + 1]++;
    }
  }
  for (int64_t c = 0; c < g.ncells; ++c) cnt[c + 1] += cnt[c];
  std::vector<int32_t> rows(cnt[g.ncells]);
  std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
  for (const WorkItem& it : P.items) {
    if (it.chunk < 0) continue;
    const Group& G = pb.groups[it.group];
    const int n = std::min(64, G.ncells - it.chunk * 64);
    for (int l = 0; l < n; ++l) {
      int r = it.chunk * 64 + l;
      const int cw = G.lo[2] + r % G.ext[2]; r /= G.ext[2];
      const int cv = G.lo[1] + r % G.ext[1]; r /= G.ext[1];
      const int cu = G.lo[0] + r;
      const int64_t cell = ((int64_t)cu * g.Mv + cv) * g.Mw + cw;
      if (it.prow + l > INT32_MAX) return set_error(ctx, DGV_ERR_UNSUPPORTED, "too many partial rows");
      rows[pos[cell]++] = (int32_t)(it.prow + l);
    }
  }
  P.ncell_rows = (int64_t)rows.size();
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMalloc(&P.d_items, sizeof(WorkItem) * std::max<size_t>(P.items.size(), 1)));
  DGV_HIP(ctx, hipMemcpy(P.d_items, P.items.data(), sizeof(WorkItem) * P.items.size(), hipMemcpyHostToDevice));
  if (hipMalloc(&P.d_partial, sizeof(double) * std::max<int64_t>(P.nrows * g.D, 1)) != hipSuccess)
    return set_error(ctx, DGV_ERR_NOMEM, "hipMalloc(partial sums) failed");
  DGV_HIP(ctx, hipMalloc(&P.d_cell_ptr, sizeof(int64_t) * (g.ncells + 1)));
  DGV_HIP(ctx, hipMemcpy(P.d_cell_ptr, cnt.data(), sizeof(int64_t) * (g.ncells + 1), hipMemcpyHostToDevice));
  DGV_HIP(ctx, hipMalloc(&P.d_cell_rows, sizeof(int32_t) * std::max<size_t>(rows.size(), 1)));
  DGV_HIP(ctx, hipMemcpy(P.d_cell_rows, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice));
  return DGV_OK;
}

// Reference layout (per canonical basis function, records ordered by xi then
// xi1 -- the serial order of SetA_DGV's i2/i3 loops, dgvtools_mod.f90:484-488)
// rebuilt from the pair-block layout.
static int export_records_uncached(const dgv_ctx* ctx, std::vector<std::vector<std::array<int64_t, 2>>>* idx,
                                   std::vector<std::vector<double>>* val) {
  const Grid& g = ctx->g;
  const PairBlock& pb = ctx->pb;
  const int D = g.D;
  std::vector<double> pairs((size_t)pb.npairs * (D + 1));
  if (pb.npairs > 0 &&
      hipMemcpy(pairs.data(), pb.d_pairs, sizeof(double) * pairs.size(), hipMemcpyDeviceToHost) != hipSuccess)
    return DGV_ERR_HIP;
  idx->assign(D, {});
  val->assign(D, {});
  for (int l = 0; l < D; ++l) { (*idx)[l].reserve(pb.nrec_basis[l]); (*val)[l].reserve(pb.nrec_basis[l]); }
  for (const BlockHdr& B : pb.blocks) {
    const int64_t xc = ((int64_t)(g.ccu + B.a[0]) * g.Mv + (g.ccv + B.a[1])) * g.Mw + (g.ccw + B.a[2]);
    const int64_t yc = ((int64_t)(g.ccu + B.b[0]) * g.Mv + (g.ccv + B.b[1])) * g.Mw + (g.ccw + B.b[2]);
    for (int p = 0; p < D; ++p)
      for (int j = B.seg[p]; j < B.seg[p + 1]; ++j) {
        const double* R = &pairs[(size_t)(B.pbeg + j) * (D + 1)];
        int64_t q;
        std::memcpy(&q, &R[D], sizeof(q));
        q &= 255;
        for (int l = 0; l < D; ++l)
          if (R[l] != 0.0) {
            (*idx)[l].push_back({xc * D + p, yc * D + q});
            (*val)[l].push_back(R[l]);
          }
      }
  }
  for (int l = 0; l < D; ++l) {
    std::vector<size_t> o((*idx)[l].size());
    std::iota(o.begin(), o.end(), 0);
    const auto& I = (*idx)[l];
    std::sort(o.begin(), o.end(), [&](size_t x, size_t y) { return I[x] < I[y]; });
    std::vector<std::array<int64_t, 2>> I2(o.size());
    std::vector<double> V2(o.size());
    for (size_t k = 0; k < o.size(); ++k) { I2[k] = I[o[k]]; V2[k] = (*val)[l][o[k]]; }
    (*idx)[l].swap(I2);
    (*val)[l].swap(V2);
  }
  return DGV_OK;
}

using IdxT = std::vector<std::vector<std::array<int64_t, 2>>>;
using ValT = std::vector<std::vector<double>>;
static int export_records(const dgv_ctx* cctx, const IdxT** idx, const ValT** val) {
  dgv_ctx* ctx = const_cast<dgv_ctx*>(cctx);
  if (!ctx->exp_valid) {
    int rc = export_records_uncached(ctx, &ctx->exp_idx, &ctx->exp_val);
    if (rc) return rc;
    ctx->exp_valid = true;
  }
  *idx = &ctx->exp_idx;
  *val = &ctx->exp_val;
  return DGV_OK;
}

int ensure_record_layout(dgv_ctx* ctx) {
  if (ctx->has_rec) return DGV_OK;
  if (!ctx->has_a) return set_error(ctx, DGV_ERR_STATE, "no A-array set");
  const IdxT* pidx = nullptr;
  const ValT* pval = nullptr;
  int rc = export_records(ctx, &pidx, &pval);
  if (rc) return set_error(ctx, rc, "export of the A-array failed");
  const IdxT& idx = *pidx;
  const ValT& val = *pval;
  const int D = ctx->g.D;
  std::vector<int64_t> off(D + 1, 0);
  for (int l = 0; l < D; ++l) off[l + 1] = off[l] + (int64_t)val[l].size();
  const int64_t n = off[D];
  std::vector<double> A(n);
  std::vector<int32_t> xi(n), xi1(n);
  for (int l = 0; l < D; ++l)
    for (size_t k = 0; k < val[l].size(); ++k) {
      A[off[l] + k] = val[l][k];
      xi[off[l] + k] = (int32_t)idx[l][k][0];
      xi1[off[l] + k] = (int32_t)idx[l][k][1];
    }
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  DGV_HIP(ctx, hipMalloc(&ctx->d_rec_off, sizeof(int64_t) * (D + 1)));
  DGV_HIP(ctx, hipMemcpy(ctx->d_rec_off, off.data(), sizeof(int64_t) * (D + 1), hipMemcpyHostToDevice));
  if (hipMalloc(&ctx->d_rec_A, sizeof(double) * std::max<int64_t>(n, 1)) != hipSuccess ||
      hipMalloc(&ctx->d_rec_xi, sizeof(int32_t) * std::max<int64_t>(n, 1)) != hipSuccess ||
      hipMalloc(&ctx->d_rec_xi1, sizeof(int32_t) * std::max<int64_t>(n, 1)) != hipSuccess)
    return set_error(ctx, DGV_ERR_NOMEM, "hipMalloc(record layout) failed");
  DGV_HIP(ctx, hipMemcpy(ctx->d_rec_A, A.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  DGV_HIP(ctx, hipMemcpy(ctx->d_rec_xi, xi.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  DGV_HIP(ctx, hipMemcpy(ctx->d_rec_xi1, xi1.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  ctx->has_rec = true;
  return DGV_OK;
}

// Records of several basis functions straight from the device layout (no host
// cache), for A-arrays too large to export whole.  The pair array comes to the
// host once.  The reference order (xi ascending, then xi1) is produced by
// construction, without a sort: xi = xc * D + p and xi1 = yc * D + q with
// xc <= yc, so walking the blocks in (xc, yc) order, and for each xc the local
// index p over every block of that xc before the next p, emits xi in order; the
// pairs of one (xc, p) come from blocks of increasing yc, each sorted by q.
// out_A / out_xi / out_xi1 receive the lists of ls[0], ls[1], ... back to back.
int export_bases_ordered(const dgv_ctx* ctx, const int32_t* ls, int nl, double* out_A, int32_t* out_xi,
                         int32_t* out_xi1, int index_base) {
  const Grid& g = ctx->g;
  const PairBlock& pb = ctx->pb;
  const int D = g.D;
  std::vector<double> pairs;
  try {
    pairs.resize((size_t)pb.npairs * (D + 1));
  } catch (const std::bad_alloc&) {
    return DGV_ERR_NOMEM;
  }
  if (pb.npairs > 0 &&
      hipMemcpy(pairs.data(), pb.d_pairs, sizeof(double) * pairs.size(), hipMemcpyDeviceToHost) != hipSuccess)
    return DGV_ERR_HIP;
  auto cell_of = [&](const int8_t* o) {
    return ((int64_t)(g.ccu + o[0]) * g.Mv + (g.ccv + o[1])) * g.Mw + (g.ccw + o[2]);
  };
  std::vector<int32_t> order(pb.blocks.size());
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
    const int64_t ax = cell_of(pb.blocks[x].a), ay = cell_of(pb.blocks[y].a);
    if (ax != ay) return ax < ay;
    return cell_of(pb.blocks[x].b) < cell_of(pb.blocks[y].b);
  });
  std::vector<int64_t> pos(nl, 0);
  for (int k = 1; k < nl; ++k) pos[k] = pos[k - 1] + pb.nrec_basis[ls[k - 1]];
  std::vector<int64_t> end(nl);
  for (int k = 0; k < nl; ++k) end[k] = pos[k] + pb.nrec_basis[ls[k]];
  for (size_t b0 = 0; b0 < order.size();) {
    const int64_t xc = cell_of(pb.blocks[order[b0]].a);
    size_t b1 = b0;
    while (b1 < order.size() && cell_of(pb.blocks[order[b1]].a) == xc) ++b1;
    for (int p = 0; p < D; ++p)
      for (size_t bi = b0; bi < b1; ++bi) {
        const BlockHdr& B = pb.blocks[order[bi]];
        const int64_t yc = cell_of(B.b);
        for (int j = B.seg[p]; j < B.seg[p + 1]; ++j) {
          const double* R = &pairs[(size_t)(B.pbeg + j) * (D + 1)];
          int64_t q;
          std::memcpy(&q, &R[D], sizeof(q));
          q &= 255;
          for (int k = 0; k < nl; ++k) {
            const double a = R[ls[k]];
            if (a == 0.0) continue;
            if (pos[k] >= end[k]) return DGV_ERR_STATE;   // counts and layout disagree
            out_A[pos[k]] = a;
            out_xi[pos[k]] = (int32_t)(xc * D + p + index_base);
            out_xi1[pos[k]] = (int32_t)(yc * D + q + index_base);
            ++pos[k];
          }
        }
      }
    b0 = b1;
  }
  for (int k = 0; k < nl; ++k)
    if (pos[k] != end[k]) return DGV_ERR_STATE;
  return DGV_OK;
}

int get_basis_records(const dgv_ctx* ctx, int l, double* A, int32_t* xi, int32_t* xi1, int64_t cap,
                      int64_t* nrec, int index_base) {
  const int64_t total = ctx->pb.npairs * (ctx->g.D + 1);
  if (!ctx->exp_valid && total > (int64_t)1 << 28) {  // > 2 GiB of pairs: per-basis path
    *nrec = ctx->pb.nrec_basis[l];
    if (!A && !xi && !xi1) return DGV_OK;
    if (cap < *nrec) return DGV_ERR_ARG;
    std::vector<double> tA(A ? 0 : *nrec);
    std::vector<int32_t> tx(xi ? 0 : *nrec), tx1(xi1 ? 0 : *nrec);
    const int32_t one[1] = {l};
    return export_bases_ordered(ctx, one, 1, A ? A : tA.data(), xi ? xi : tx.data(), xi1 ? xi1 : tx1.data(),
                                index_base);
  }
  const IdxT* pidx = nullptr;
  const ValT* pval = nullptr;
  int rc = export_records(ctx, &pidx, &pval);
  if (rc) return rc;
  const IdxT& idx = *pidx;
  const ValT& val = *pval;
  const int64_t n = (int64_t)val[l].size();
  *nrec = n;
  if (!A && !xi && !xi1) return DGV_OK;
  if (cap < n) return DGV_ERR_ARG;
  for (int64_t k = 0; k < n; ++k) {
    if (A) A[k] = val[l][k];
    if (xi) xi[k] = (int32_t)(idx[l][k][0] + index_base);
    if (xi1) xi1[k] = (int32_t)(idx[l][k][1] + index_base);
  }
  return DGV_OK;
}

// Whole reference-layout A-array (all nodes; only canonical nodes have records).
int get_all_records(const dgv_ctx* ctx, std::vector<int32_t>& capphi, std::vector<double>& A,
                    std::vector<int32_t>& xi, std::vector<int32_t>& xi1, int index_base) {
  const IdxT* pidx = nullptr;
  const ValT* pval = nullptr;
  int rc = export_records(ctx, &pidx, &pval);
  if (rc) return rc;
  const IdxT& idx = *pidx;
  const ValT& val = *pval;
  const Grid& g = ctx->g;
  capphi.assign(g.N, 0);
  A.clear(); xi.clear(); xi1.clear();
  for (int l = 0; l < g.D; ++l) {
    capphi[(int64_t)g.ccell * g.D + l] = (int32_t)val[l].size();
    for (size_t k = 0; k < val[l].size(); ++k) {
      A.push_back(val[l][k]);
      xi.push_back((int32_t)(idx[l][k][0] + index_base));
      xi1.push_back((int32_t)(idx[l][k][1] + index_base));
    }
  }
  return DGV_OK;
}

}  // namespace dgv
