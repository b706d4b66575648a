// dgv_kernels.hip -- gfx950 kernels of the direct DGV collision evaluator (FP64)
// and the device arithmetic of the spatial operator and the time steppers.
#include <mutex>
#include <string>
#include <cmath>
#include <algorithm>
#include <cstring>
#include <vector>

#include "dgv_internal.h"

namespace dgv {

// ---------------------------------------------------------------------------
// Pair-block evaluator.  One wave per work item (group chunk x block run).
// Lane = one output cell of the group's box.  Per block the lane holds
// f[c+a][0..D-1] in registers and f[c+b][0..D-1] in its own LDS column; per
// pair it forms g = f[c+a][p] * f[c+b][q] once and accumulates the pair's D
// coefficients (wave-uniform, read with scalar loads) into D accumulators:
//   acc[l] += T[l] * g.
// Bounds tests of collision_mod.f90:196-211 are unnecessary: every cell of a
// group's box is in range for every block of the group by construction.
// MIX: the cross term of EvalCollisionPeriodicMixedTermsA_DGV
// (collision_mod.f90:244-332) with g = fm: the pair's product is
// g[c+a][p] f[c+b][q] + f[c+a][p] g[c+b][q].
template <int D, bool MIX>
__global__ __launch_bounds__(64) void k_pb_eval(const WorkItem* __restrict__ items,
                                                const Group* __restrict__ groups,
                                                const int32_t* __restrict__ gblocks,
                                                const BlockHdr* __restrict__ blocks,
                                                const double* __restrict__ pairs,
                                                const double* __restrict__ f,
                                                const double* __restrict__ gm,
                                                double* __restrict__ partial, int Mv, int Mw) {
  __shared__ double fb[D * 64];
  __shared__ double gb[MIX ? D * 64 : 1];
  const WorkItem it = items[blockIdx.x];
  if (it.chunk < 0) return;  // padding slot of the XCD-aware order
  const Group g = groups[it.group];
  const int lane = threadIdx.x;
  const int cl = it.chunk * 64 + lane;
  const bool active = cl < g.ncells;
  int r = active ? cl : 0;  // idle lanes shadow the box's first cell (in range)
  const int cw = g.lo[2] + r % g.ext[2];
  r /= g.ext[2];
  const int cv = g.lo[1] + r % g.ext[1];
  r /= g.ext[1];
  const int cu = g.lo[0] + r;
  constexpr int64_t kRec = (D + 1) * 8;     // bytes per pair record
  constexpr int64_t kWin = 64 * 128;        // one prefetch instruction: 64 lanes x 128-B lines
  // L2 prefetch of the pair records: an ordinary (lane-divergent, so vector)
  // load of one dword per 128-B line, kept alive by folding it into `sink`
  // one segment later, when it has long arrived (no wait on the critical path).
  int sink = 0, pending = 0;

  double acc[D];
#pragma unroll
  for (int l = 0; l < D; ++l) acc[l] = 0.0;

  for (int bi = it.bbeg; bi < it.bend; ++bi) {
    const BlockHdr* B = blocks + gblocks[bi];
    const int xc = ((cu + B->a[0]) * Mv + (cv + B->a[1])) * Mw + (cw + B->a[2]);
    const int yc = ((cu + B->b[0]) * Mv + (cv + B->b[1])) * Mw + (cw + B->b[2]);
    const double* P = pairs + B->pbeg * (D + 1);
    const char* Pb = reinterpret_cast<const char*>(P);
    const int64_t bend = (int64_t)B->seg[D] * kRec;
    int64_t pf = 0;
    const double* fx = f + (int64_t)xc * D;
    const double* fy = f + (int64_t)yc * D;
    double fa[D], ga[MIX ? D : 1];
#pragma unroll
    for (int p = 0; p < D; ++p) fa[p] = fx[p];
#pragma unroll
    for (int q = 0; q < D; ++q) fb[q * 64 + lane] = fy[q];
    if constexpr (MIX) {
#pragma unroll
      for (int p = 0; p < D; ++p) ga[p] = gm[(int64_t)xc * D + p];
#pragma unroll
      for (int q = 0; q < D; ++q) gb[q * 64 + lane] = gm[(int64_t)yc * D + q];
    }
#pragma unroll
    for (int p = 0; p < D; ++p) {
      const int j0 = B->seg[p], j1 = B->seg[p + 1];
      // keep the next 16 KB of records on their way to L2
      sink ^= pending;
      if (pf < bend && pf < (int64_t)j1 * kRec + 2 * kWin) {
        const int64_t o = pf + (int64_t)lane * 128;
        pending = o < bend ? *reinterpret_cast<const int*>(Pb + o) : 0;
        pf += kWin;
      }
      const double fap = fa[p];
      const double gap = MIX ? ga[p] : 0.0;
      // q of the next pair is read together with the current coefficients,
      // so each pair costs one scalar-load round trip
      int qn = j0 < j1 ? (int)(__double_as_longlong(P[(int64_t)j0 * (D + 1) + D]) & 255) : 0;
      for (int j = j0; j < j1; ++j) {
        const double* R = P + (int64_t)j * (D + 1);
        const double fbq = fb[qn * 64 + lane];
        const double gbq = MIX ? gb[qn * 64 + lane] : 0.0;
        if (j + 1 < j1) qn = (int)(__double_as_longlong(R[(D + 1) + D]) & 255);
        const double gq = MIX ? __dadd_rn(__dmul_rn(gap, fbq), __dmul_rn(fap, gbq)) : fap * fbq;
#pragma unroll
        for (int l = 0; l < D; ++l) acc[l] = __fma_rn(R[l], gq, acc[l]);
      }
    }
  }
  sink ^= pending;
  if (active) {
    double* out = partial + (it.prow + lane) * D;
#pragma unroll
    for (int l = 0; l < D; ++l) out[l] = acc[l];
  }
  if (sink == 0x5eed5eed && it.prow < 0) partial[0] = (double)sink;  // never true; keeps the prefetch
}

// Pair-block evaluator on the matrix cores.  For one block and the 64 cells
// of a chunk the sum is a small matrix product,
//   Q[l, cell] += sum_pairs T[l, pair] * g[pair, cell],
//   g[pair, cell] = f[cell + a][p] * f[cell + b][q],
// done with v_mfma_f64_16x16x4_f64 (A = 16 basis functions x 4 pairs,
// B = 4 pairs x 16 cells, C = 16 x 16): RT row tiles cover the D basis
// functions (rows >= D are zero), column tiles the cells.  Two waves share a
// work item and its LDS-staged f values; wave w owns cells 32 w .. 32 w + 31
// (two column tiles).  The coefficients reach the A operand through ordinary
// vector loads (lane i holds T[pair s + i/16][16 rt + i%16]) issued kAhead
// steps ahead; g is formed on the VALU from the staged f.  Operand layouts of
// the instruction: A[i][k] and B[k][j] in lane i|j + 16 k; D[i][j] in lane
// j + 16 (i%4), element i/4 (the FP64 output layout).
typedef double dgv_d4 __attribute__((ext_vector_type(4)));
#define AS1 __attribute__((address_space(1)))   // global memory: plain global loads, not flat

constexpr int kPbWaves = 2;

typedef double dgv_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double lane_el(const dgv_d2& v, int i) { return v[i]; }
__device__ __forceinline__ double lane_el(double v, int) { return v; }

template <int D, bool MIX>
__global__ __launch_bounds__(64 * kPbWaves) void k_pb_mfma(const WorkItem* __restrict__ items,
                                                           const Group* __restrict__ groups,
                                                           const int32_t* __restrict__ gblocks,
                                                           const BlockHdr* __restrict__ blocks,
                                                           const double* __restrict__ pairs,
                                                           const double* __restrict__ f,
                                                           const double* __restrict__ gm,
                                                           double* __restrict__ partial, int Mv, int Mw) {
  constexpr int RT = (D + 15) / 16;
  constexpr int CT = 4 / kPbWaves;
  constexpr int kRing = 8;
  static_assert(4 * kRing + 3 <= kPairPad, "the ring reads at most 4 kRing + 3 pairs past a block");
  __shared__ double lfa[D * 64], lfb[D * 64];
  __shared__ double lga[MIX ? D * 64 : 1], lgb[MIX ? D * 64 : 1];
  const WorkItem it = items[blockIdx.x];
  if (it.chunk < 0) return;
  const Group g = groups[it.group];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int cl = it.chunk * 64 + lane;
  int r = cl < g.ncells ? cl : 0;   // idle lanes stage the box's first cell (in range)
  const int cw = g.lo[2] + r % g.ext[2];
  r /= g.ext[2];
  const int cv = g.lo[1] + r % g.ext[1];
  r /= g.ext[1];
  const int cu = g.lo[0] + r;
  const int kk = lane >> 4, jj = lane & 15;
  const int col0 = wave * 16 * CT;
  // Column jj of tile ct is cell col0 + CT jj + ct: a lane's CT staged values
  // of one row are one 16-byte LDS read.
  const int lcol = col0 + CT * jj;
  const bool wave_cells = it.chunk * 64 + col0 < g.ncells;   // the wave owns at least one cell
  // the lane's byte offsets inside a step of 4 pair records (rows 16 rt + jj
  // of the coefficient operand, clamped to the code slot D, and the high word
  // of the code slot: the pair's two LDS row offsets)
  uint32_t loffb[RT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) loffb[rt] = (uint32_t)((kk * (D + 1) + min(16 * rt + jj, D)) * 8);
  const uint32_t coffb = (uint32_t)((kk * (D + 1) + D) * 8 + 4);
  typedef double d2 __attribute__((ext_vector_type(2)));
  using OpT = typename std::conditional<CT == 2, d2, double>::type;   // a lane's CT values of a row
  static_assert(CT == 1 || CT == 2, "one LDS read per operand and step");

  dgv_d4 acc[RT][CT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = dgv_d4{0.0, 0.0, 0.0, 0.0};

  const char* la_lane = reinterpret_cast<const char*>(lfa + lcol);
  const char* lb_lane = reinterpret_cast<const char*>(lfb + lcol);
  const char* ga_lane = reinterpret_cast<const char*>(lga + (MIX ? lcol : 0));
  const char* gb_lane = reinterpret_cast<const char*>(lgb + (MIX ? lcol : 0));
  int prev_a = -1;   // offset a whose f is staged in lfa (a group's blocks come ordered by a)
  for (int bi = it.bbeg; bi < it.bend; ++bi) {
    const BlockHdr* B = blocks + gblocks[bi];
    const int xc = ((cu + B->a[0]) * Mv + (cv + B->a[1])) * Mw + (cw + B->a[2]);
    const int yc = ((cu + B->b[0]) * Mv + (cv + B->b[1])) * Mw + (cw + B->b[2]);
    const int akey = ((B->a[0] + 64) << 16) | ((B->a[1] + 64) << 8) | (B->a[2] + 64);
    const bool new_a = akey != prev_a;
    prev_a = akey;
    const int n = B->npairs;
    // a wave whose cells all lie past the group's box (a chunk's partial end)
    // only stages: its products would all be dropped
    const bool compute = n > 0 && wave_cells;
    // Ring of kRing steps of coefficient loads in flight.  The pair array
    // carries zero records past its end, so the ring reads straight ahead from
    // a uniform step base; only the block's last, partial step masks the pairs
    // past the block's end.  The ring is filled before the f staging, so the
    // two memory latencies of a block overlap.
    const __amdgpu_buffer_rsrc_t P = ring_rsrc(pairs + B->pbeg * (D + 1));
    constexpr uint32_t kStepBytes = 4 * (D + 1) * 8;
    double ab[kRing][RT];
    int cbits[kRing];
#define PB_LOAD(slot, step)                                                               \
  do {                                                                                    \
    const uint32_t s_ = kStepBytes * (uint32_t)(step);                                    \
    _Pragma("unroll") for (int rt = 0; rt < RT; ++rt) ab[slot][rt] = ring_f64(P, loffb[rt], s_); \
    cbits[slot] = ring_i32(P, coffb, s_);                                                 \
  } while (0)
    if (compute) {
#pragma unroll
      for (int u = 0; u < kRing; ++u) PB_LOAD(u, u);
    }
    // thread (wave, lane) stages the nodes p = wave, wave + kPbWaves, ... of its
    // lane's cell: all loads are issued into registers first (one memory
    // latency per block), then written to LDS after the barrier
    constexpr int NS = (D + kPbWaves - 1) / kPbWaves;
    double sa[NS], sb[NS];
#define PB_STAGE_LOAD(src)                                                                \
  do {                                                                                    \
    _Pragma("unroll") for (int i = 0; i < NS; ++i) {                                      \
      const int p_ = min(wave + i * kPbWaves, D - 1);                                     \
      if (new_a) sa[i] = (src)[(int64_t)xc * D + p_];                                     \
      sb[i] = (src)[(int64_t)yc * D + p_];                                                \
    }                                                                                     \
  } while (0)
#define PB_STAGE_STORE(da, db)                                                            \
  do {                                                                                    \
    _Pragma("unroll") for (int i = 0; i < NS; ++i) {                                      \
      const int p_ = wave + i * kPbWaves;                                                 \
      if (p_ < D) {                                                                       \
        if (new_a) (da)[p_ * 64 + lane] = sa[i];                                          \
        (db)[p_ * 64 + lane] = sb[i];                                                     \
      }                                                                                   \
    }                                                                                     \
  } while (0)
    PB_STAGE_LOAD(f);
    __syncthreads();  // previous block's LDS reads are done
    PB_STAGE_STORE(lfa, lfb);
    if constexpr (MIX) {   // the second VDF through the same registers
      PB_STAGE_LOAD(gm);
      PB_STAGE_STORE(lga, lgb);
    }
#undef PB_STAGE_STORE
#undef PB_STAGE_LOAD
    __syncthreads();
    if (!compute) continue;
    // The LDS reads of the next step's B operand are issued before the current
    // step's products.
    OpT xa, xb, ya, yb;
    double bcur[CT];
#define PB_ISSUE(slot)                                                                    \
  do {                                                                                    \
    const uint32_t c_ = (uint32_t)cbits[slot];                                            \
    const uint32_t pb_ = c_ & 0xffffu, qb_ = c_ >> 16;   /* pair_code_bits */             \
    xa = *reinterpret_cast<const OpT*>(la_lane + pb_);                                    \
    xb = *reinterpret_cast<const OpT*>(lb_lane + qb_);                                    \
    if constexpr (MIX) {                                                                  \
      ya = *reinterpret_cast<const OpT*>(ga_lane + pb_);                                  \
      yb = *reinterpret_cast<const OpT*>(gb_lane + qb_);                                  \
    }                                                                                     \
  } while (0)
#define PB_FINISH()                                                                       \
  do {                                                                                    \
    _Pragma("unroll") for (int ct = 0; ct < CT; ++ct) {                                   \
      if constexpr (MIX)                                                                  \
        bcur[ct] = __dadd_rn(__dmul_rn(lane_el(ya, ct), lane_el(xb, ct)), __dmul_rn(lane_el(xa, ct), lane_el(yb, ct))); \
      else bcur[ct] = lane_el(xa, ct) * lane_el(xb, ct);                                  \
    }                                                                                     \
  } while (0)
#define PB_MFMA(u)                                                                        \
  do {                                                                                    \
    _Pragma("unroll") for (int ct = 0; ct < CT; ++ct)                                     \
    _Pragma("unroll") for (int rt = 0; rt < RT; ++rt)                                     \
      acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(ab[u][rt], bcur[ct], acc[rt][ct], 0, 0, 0); \
  } while (0)
    const int nfull = n >> 2, nst = (n + 3) >> 2;
    PB_ISSUE(0);
    PB_FINISH();
    int st = 0;
    for (; st + kRing <= nfull; st += kRing) {
#pragma unroll
      for (int u = 0; u < kRing; ++u) {
        PB_ISSUE((u + 1) % kRing);
        // keep the next step's LDS reads ahead of this step's products, and
        // the wait for them behind (the scheduler would otherwise sink the reads
        // below the products and wait at once)
        __builtin_amdgcn_sched_barrier(0);
        PB_MFMA(u);
        __builtin_amdgcn_sched_barrier(0);
        PB_LOAD(u, st + u + kRing);
        PB_FINISH();
      }
    }
#pragma unroll
    for (int u = 0; u < kRing; ++u) {   // the last nst - st <= kRing steps, already in the ring
      if (st + u < nst) {
        if (st + u == nfull) {   // partial step: records past the block's end contribute 0
          const bool ok = 4 * nfull + kk < n;
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) ab[u][rt] = ok ? ab[u][rt] : 0.0;
        }
        PB_ISSUE((u + 1) % kRing);
        PB_MFMA(u);
        PB_FINISH();
      }
    }
#undef PB_MFMA
#undef PB_FINISH
#undef PB_ISSUE
#undef PB_LOAD
  }
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      const int cell = col0 + CT * jj + ct;
      if (it.chunk * 64 + cell >= g.ncells) continue;
      double* out = partial + (it.prow + cell) * D;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int l = 16 * rt + 4 * e + kk;
        if (l < D) out[l] = acc[rt][ct][e];
      }
    }
}

// Final FP64 reduction, one wave per output cell: lane j sums the cell's
// partial rows j, j+64, ... in order (all D basis functions at once), then a
// fixed binary tree over the 64 lanes in LDS; Q = (2 * sum) / w, times
// `scale` (the spatial operator's (d/L)^2 N dt, spat_oper_mod.f90:65-66).
template <int D>
__global__ __launch_bounds__(64) void k_pb_reduce(int64_t ncells, const int64_t* __restrict__ cell_ptr,
                                                  const int32_t* __restrict__ cell_rows,
                                                  const double* __restrict__ partial,
                                                  const double* __restrict__ gwts, double scale,
                                                  double* __restrict__ q) {
  __shared__ double red[64 * (D + 1)];
  const int64_t c = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t b = cell_ptr[c], e = cell_ptr[c + 1];
  double s[D];
#pragma unroll
  for (int l = 0; l < D; ++l) s[l] = 0.0;
  for (int64_t k = b + lane; k < e; k += 64) {
    const double* row = partial + (int64_t)cell_rows[k] * D;
#pragma unroll
    for (int l = 0; l < D; ++l) s[l] = s[l] + row[l];
  }
#pragma unroll
  for (int l = 0; l < D; ++l) red[lane * (D + 1) + l] = s[l];
  __syncthreads();
  for (int h = 32; h > 0; h >>= 1) {
    if (lane < h) {
#pragma unroll
      for (int l = 0; l < D; ++l) red[lane * (D + 1) + l] = red[lane * (D + 1) + l] + red[(lane + h) * (D + 1) + l];
    }
    __syncthreads();
  }
  for (int l = lane; l < D; l += 64) {
    double v = (2.0 * red[l]) / gwts[c * D + l];
    if (scale != 1.0) v = v * scale;
    q[c * D + l] = v;
  }
}

// ---------------------------------------------------------------------------
// Record kernel: the reference's loop structure on the device
// (collision_mod.f90:186-227).  One workgroup per node iphi; the threads stride
// over the records of the node's canonical basis function, apply the shift and
// the six bounds tests, accumulate 2*f(xi)*f(xi1)*A, and reduce in a fixed tree.
__global__ __launch_bounds__(256) void k_rec_eval(int D, int Mu, int Mv, int Mw, int ccu, int ccv,
                                                  int ccw, const int64_t* __restrict__ off,
                                                  const double* __restrict__ A,
                                                  const int32_t* __restrict__ xi,
                                                  const int32_t* __restrict__ xi1,
                                                  const double* __restrict__ f,
                                                  const double* __restrict__ gwts, double scale,
                                                  double* __restrict__ q) {
  __shared__ double red[256];
  const int64_t iphi = blockIdx.x;
  const int64_t cell = iphi / D;
  const int l = (int)(iphi - cell * D);
  const int cw = (int)(cell % Mw), cv = (int)((cell / Mw) % Mv), cu = (int)(cell / ((int64_t)Mw * Mv));
  const int du = cu - ccu, dv = cv - ccv, dw = cw - ccw;
  double acc = 0.0;
  for (int64_t j = off[l] + threadIdx.x; j < off[l + 1]; j += 256) {
    const int32_t x = xi[j], y = xi1[j];
    const int32_t xc = x / D, yc = y / D;
    const int xu = xc / (Mv * Mw) + du, xv = (xc / Mw) % Mv + dv, xw = xc % Mw + dw;
    if (xu < 0 || xu >= Mu || xv < 0 || xv >= Mv || xw < 0 || xw >= Mw) continue;
    const int yu = yc / (Mv * Mw) + du, yv = (yc / Mw) % Mv + dv, yw = yc % Mw + dw;
    if (yu < 0 || yu >= Mu || yv < 0 || yv >= Mv || yw < 0 || yw >= Mw) continue;
    const int64_t kx = ((int64_t)(xu * Mv + xv) * Mw + xw) * D + (x - xc * D);
    const int64_t ky = ((int64_t)(yu * Mv + yv) * Mw + yw) * D + (y - yc * D);
    acc = acc + 2.0 * f[kx] * f[ky] * A[j];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double v = red[0] / gwts[iphi];
    if (scale != 1.0) v = v * scale;
    q[iphi] = v;
  }
}

// d_g != nullptr: the mixed term with g (EvalCollisionPeriodicMixedTermsA_DGV)
// partial: where the items' partial rows go (the plan's buffer unless given)
static int launch_main_pairblock(dgv_ctx* ctx, const double* d_f, bool mfma, const double* d_g = nullptr,
                                 double* partial = nullptr) {
  const Grid& g = ctx->g;
  const PairBlock& pb = ctx->pb;
  const Plan& P = ctx->plan;
  double* const part = partial ? partial : P.d_partial;
  const int nitems = (int)P.items.size();
  if (nitems == 0) return DGV_OK;
  if (mfma) {
    switch (g.D) {
#define DGV_MF_CASE(DD)                                                                               \
  case DD:                                                                                            \
    if (d_g)                                                                                          \
      hipLaunchKernelGGL((k_pb_mfma<DD, true>), dim3(nitems), dim3(64 * kPbWaves), 0, ctx->stream,      \
                         P.d_items,                                                                   \
                         pb.d_groups, pb.d_group_blocks, pb.d_blocks, pb.d_pairs, d_f, d_g,           \
                         part, g.Mv, g.Mw);                                                    \
    else                                                                                              \
      hipLaunchKernelGGL((k_pb_mfma<DD, false>), dim3(nitems), dim3(64 * kPbWaves), 0, ctx->stream,     \
                         P.d_items,                                                                   \
                         pb.d_groups, pb.d_group_blocks, pb.d_blocks, pb.d_pairs, d_f, d_f,           \
                         part, g.Mv, g.Mw);                                                    \
    break;
      DGV_MF_CASE(8)
      DGV_MF_CASE(27)
#undef DGV_MF_CASE
      default:
        return set_error(ctx, DGV_ERR_UNSUPPORTED, "matrix-core kernel supports 8 or 27 nodes per cell");
    }
    DGV_HIP(ctx, hipGetLastError());
    return DGV_OK;
  }
  switch (g.D) {
#define DGV_PB_CASE(DD)                                                                               \
  case DD:                                                                                            \
    if (d_g)                                                                                          \
      hipLaunchKernelGGL((k_pb_eval<DD, true>), dim3(nitems), dim3(64), 0, ctx->stream, P.d_items,    \
                         pb.d_groups, pb.d_group_blocks, pb.d_blocks, pb.d_pairs, d_f, d_g,           \
                         part, g.Mv, g.Mw);                                                    \
    else                                                                                              \
      hipLaunchKernelGGL((k_pb_eval<DD, false>), dim3(nitems), dim3(64), 0, ctx->stream, P.d_items,   \
                         pb.d_groups, pb.d_group_blocks, pb.d_blocks, pb.d_pairs, d_f, d_f,           \
                         part, g.Mv, g.Mw);                                                    \
    break;
    DGV_PB_CASE(1)
    DGV_PB_CASE(8)
    DGV_PB_CASE(27)
#undef DGV_PB_CASE
    default:
      return set_error(ctx, DGV_ERR_UNSUPPORTED, "pair-block kernel supports 1, 8 or 27 nodes per cell");
  }
  DGV_HIP(ctx, hipGetLastError());
  return DGV_OK;
}

static int launch_reduce(dgv_ctx* ctx, double* d_q, double scale, const double* partial = nullptr) {
  const Grid& g = ctx->g;
  const Plan& P = ctx->plan;
  const double* const part = partial ? partial : P.d_partial;
  switch (g.D) {
#define DGV_RED_CASE(DD)                                                                               \
  case DD:                                                                                             \
    hipLaunchKernelGGL(k_pb_reduce<DD>, dim3((unsigned)g.ncells), dim3(64), 0, ctx->stream, g.ncells,  \
                       P.d_cell_ptr, P.d_cell_rows, part, ctx->d_gwts, scale, d_q);                    \
    break;
    DGV_RED_CASE(1)
    DGV_RED_CASE(8)
    DGV_RED_CASE(27)
#undef DGV_RED_CASE
  }
  DGV_HIP(ctx, hipGetLastError());
  return DGV_OK;
}

int launch_eval_pairblock(dgv_ctx* ctx, const double* d_f, double* d_q, double scale, bool mfma) {
  int rc = launch_main_pairblock(ctx, d_f, mfma);
  if (rc) return rc;
  return launch_reduce(ctx, d_q, scale);
}

int launch_eval_record(dgv_ctx* ctx, const double* d_f, double* d_q, double scale) {
  int rc = ensure_record_layout(ctx);
  if (rc) return rc;
  const Grid& g = ctx->g;
  hipLaunchKernelGGL(k_rec_eval, dim3((unsigned)g.N), dim3(256), 0, ctx->stream, g.D, g.Mu, g.Mv, g.Mw, g.ccu,
                     g.ccv, g.ccw, ctx->d_rec_off, ctx->d_rec_A, ctx->d_rec_xi, ctx->d_rec_xi1, d_f, ctx->d_gwts,
                     scale, d_q);
  DGV_HIP(ctx, hipGetLastError());
  return DGV_OK;
}

static bool pairblock_supported(int D) { return D == 1 || D == 8 || D == 27; }
static bool mfma_supported(int D) { return D == 8 || D == 27; }

static int resolve_kernel(const dgv_ctx* ctx) {
  int k = ctx->kernel;
  const int D = ctx->g.D;
  if (k == DGV_KERNEL_AUTO)
    k = mfma_supported(D) ? DGV_KERNEL_MFMA : (pairblock_supported(D) ? DGV_KERNEL_PAIRBLOCK : DGV_KERNEL_RECORD);
  if (k == DGV_KERNEL_MFMA && !mfma_supported(D)) k = DGV_KERNEL_PAIRBLOCK;
  if (k == DGV_KERNEL_PAIRBLOCK && !pairblock_supported(D)) k = DGV_KERNEL_RECORD;
  return k;
}

// The dominant kernel alone (no final reduction), for timing.
int launch_eval_main_only(dgv_ctx* ctx, const double* d_f) {
  const int k = resolve_kernel(ctx);
  if (k == DGV_KERNEL_RECORD) return launch_eval_record(ctx, d_f, ctx->d_q, 1.0);
  return launch_main_pairblock(ctx, d_f, k == DGV_KERNEL_MFMA);
}

// EvalCollisionPeriodicMixedTermsA_DGV(f, fm, fc) on the device (pair-block kernels)
int eval_mixed_dev(dgv_ctx* ctx, const double* d_f, const double* d_fm, double* d_q) {
  if (!ctx->has_a) return set_error(ctx, DGV_ERR_STATE, "no A-array set (dgv_ctx_set_aarray / generate)");
  int k = resolve_kernel(ctx);
  if (k == DGV_KERNEL_RECORD) {
    if (!pairblock_supported(ctx->g.D))
      return set_error(ctx, DGV_ERR_UNSUPPORTED, "the mixed term needs 1, 8 or 27 nodes per cell");
    k = DGV_KERNEL_PAIRBLOCK;
  }
  int rc = launch_main_pairblock(ctx, d_f, k == DGV_KERNEL_MFMA, d_fm);
  if (rc) return rc;
  return launch_reduce(ctx, d_q, 1.0);
}

// ---- the allocation-free evaluator -------------------------------------------------
// A file-scope device array holds the partial sums (and the projection's
// moments) of dgv_eval_q_dev_noalloc, for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops).  Each GPU holds its
// own copy from the library's load to the end of the process; its address on a
// device comes from hipGetSymbolAddress, looked up once per device when the
// first context is created there.  Every call overwrites what it reads, so no
// clearing is needed; calls that overlap on one GPU (two streams, two threads)
// share the array and race.
constexpr size_t kStaticScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) double scratch[kStaticScratchBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

int static_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return DGV_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return DGV_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return DGV_ERR_HIP;
  g_scratch_ptr[device] = static_cast<double*>(p);
  return DGV_OK;
}

int eval_q_dev(dgv_ctx* ctx, const double* d_f, double* d_q, double scale) {
  if (!ctx->has_a) return set_error(ctx, DGV_ERR_STATE, "no A-array set (dgv_ctx_set_aarray / generate)");
  const int k = resolve_kernel(ctx);
  int rc = k == DGV_KERNEL_RECORD ? launch_eval_record(ctx, d_f, d_q, scale)
                                  : launch_eval_pairblock(ctx, d_f, d_q, scale, k == DGV_KERNEL_MFMA);
  if (rc == DGV_OK && ctx->project) rc = device_project(ctx, d_q);
  return rc;
}

// Nonzero coefficients per block (records of the reference layout).
__global__ void k_block_nnz(int D, int64_t nblocks, const BlockHdr* __restrict__ blocks,
                            const double* __restrict__ pairs, int64_t* __restrict__ out) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const BlockHdr& B = blocks[b];
  int64_t n = 0;
  for (int j = 0; j < B.npairs; ++j)
    for (int l = 0; l < D; ++l) n += pairs[(B.pbeg + j) * (D + 1) + l] != 0.0;
  out[b] = n;
}

int workload(dgv_ctx* ctx, int64_t out[4]) {
  const Grid& g = ctx->g;
  const PairBlock& pb = ctx->pb;
  const int64_t nb = (int64_t)pb.blocks.size();
  std::vector<int64_t> nnz(nb);
  int64_t* d = nullptr;
  DGV_HIP(ctx, hipMalloc(&d, sizeof(int64_t) * std::max<int64_t>(nb, 1)));
  hipLaunchKernelGGL(k_block_nnz, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, ctx->stream, g.D, nb,
                     pb.d_blocks, pb.d_pairs, d);
  hipError_t e = hipMemcpyAsync(nnz.data(), d, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return set_error(ctx, DGV_ERR_HIP, std::string("workload: ") + hipGetErrorString(e));
  const int M[3] = {g.Mu, g.Mv, g.Mw};
  int64_t R = 0, P = 0, V = 0, PV = 0;
  for (int64_t b = 0; b < nb; ++b) {
    const BlockHdr& B = pb.blocks[b];
    int64_t box = 1;
    for (int dd = 0; dd < 3; ++dd) {
      const int lo = std::max(0, std::max(-(int)B.a[dd], -(int)B.b[dd]));
      const int hi = std::min(M[dd], std::min(M[dd] - B.a[dd], M[dd] - B.b[dd]));
      box *= std::max(0, hi - lo);
    }
    R += nnz[b];
    P += B.npairs;
    V += nnz[b] * box;
    PV += (int64_t)B.npairs * box;
  }
  out[0] = R; out[1] = P; out[2] = V; out[3] = PV;
  return DGV_OK;
}

// ---------------------------------------------------------------------------
// Linear combinations of the steppers, evaluated term by term left to right
// as the reference's Fortran expressions are: out = t0 + t1 + ... with
// t_k = sign_k * ((x_k * m1_k) * m2_k) / den_k  (time_integr_mod.f90:47-145,
// 325-357).
struct LinTerm {
  const double* x;
  double m1, m2, den;
  int neg;
};
struct LinComb {
  LinTerm t[6];
  int n;
};

// Plain device copy of 16-byte words (the byte counters of a profiler are
// calibrated against it: n16 * 16 bytes read and written).
__global__ void k_copy16(int64_t n16, const uint4* __restrict__ src, uint4* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int launch_copy16(dgv_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes % 16 || (reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16)
    return set_error(ctx, DGV_ERR_ARG, "device copy needs 16-byte aligned sizes and pointers");
  const int64_t n16 = (int64_t)(bytes / 16);
  if (n16 > 0)
    hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, ctx->stream, n16, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst));
  DGV_HIP(ctx, hipGetLastError());
  return DGV_OK;
}

__global__ void k_lincomb(int64_t n, double* __restrict__ out, LinComb c) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
    double s = 0.0;
    for (int k = 0; k < c.n; ++k) {
      const LinTerm& t = c.t[k];
      double v = t.x[i];
      if (t.m1 != 1.0) v = v * t.m1;
      if (t.m2 != 1.0) v = v * t.m2;
      if (t.den != 1.0) v = v / t.den;
      if (k == 0) s = t.neg ? -v : v;
      else s = t.neg ? s - v : s + v;
    }
    out[i] = s;
  }
}

int launch_lincomb(dgv_ctx* ctx, int64_t n, double* out, int nterms, const double* const* xs, const double* m1,
                   const double* m2, const double* den, const int* neg) {
  LinComb c;
  std::memset(&c, 0, sizeof(c));
  c.n = nterms;
  for (int k = 0; k < nterms; ++k) {
    c.t[k].x = xs[k];
    c.t[k].m1 = m1 ? m1[k] : 1.0;
    c.t[k].m2 = m2 ? m2[k] : 1.0;
    c.t[k].den = den ? den[k] : 1.0;
    c.t[k].neg = neg ? neg[k] : 0;
  }
  hipLaunchKernelGGL(k_lincomb, dim3(stride_blocks(n, 256)), dim3(256), 0, ctx->stream, n, out, c);
  DGV_HIP(ctx, hipGetLastError());
  return DGV_OK;
}

// ---------------------------------------------------------------------------
// Deterministic FP64 reductions: 256 workgroups x 256 threads, each thread a
// fixed strided subset, fixed LDS tree, then one workgroup sums the 256
// partials in order.
constexpr int kRedBlocks = 256;

// mode 0: {sum f w, sum f w u, sum f w v, sum f w w, 0}
// mode 1: {0,0,0,0, sum f w |u - ubar|^2} with ubar read from mean[]
// mode 2: {sum q w psi_k}, psi = 1, u, v, w, |u|^2
__global__ __launch_bounds__(256) void k_moments_part(int mode, int64_t n, const double* __restrict__ f,
                                                      const double* __restrict__ w, const double* __restrict__ u,
                                                      const double* __restrict__ v, const double* __restrict__ ww,
                                                      const double* __restrict__ mean, double* __restrict__ part) {
  __shared__ double red[5][256];
  double s[5] = {0, 0, 0, 0, 0};
  double ub = 0, vb = 0, wb = 0;
  if (mode == 1) { ub = mean[1]; vb = mean[2]; wb = mean[3]; }
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)kRedBlocks * 256) {
    const double fw = f[i] * w[i];
    if (mode == 0) {
      s[0] += fw;
      s[1] += fw * u[i];
      s[2] += fw * v[i];
      s[3] += fw * ww[i];
    } else if (mode == 1) {
      const double du = u[i] - ub, dv = v[i] - vb, dw = ww[i] - wb;
      s[4] += fw * (du * du + dv * dv + dw * dw);
    } else {
      s[0] += fw;
      s[1] += fw * u[i];
      s[2] += fw * v[i];
      s[3] += fw * ww[i];
      s[4] += fw * (u[i] * u[i] + v[i] * v[i] + ww[i] * ww[i]);
    }
  }
  for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h)
      for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; ++k) part[blockIdx.x * 5 + k] = red[k][0];
}

// Sums the partials; mode 0 turns the sums into (n, ubar, vbar, wbar) as
// MassCheckRec does (dgvtools_mod.f90:1448-1452); mode 1 finishes T (line 1455).
__global__ void k_moments_final(int mode, const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s[5] = {0, 0, 0, 0, 0};
  for (int b = 0; b < kRedBlocks; ++b)
    for (int k = 0; k < 5; ++k) s[k] += part[b * 5 + k];
  if (mode == 0) {
    out[0] = s[0];
    out[1] = s[1] / s[0];
    out[2] = s[2] / s[0];
    out[3] = s[3] / s[0];
  } else if (mode == 1) {
    out[4] = s[4] / out[0] / 3.0 * 2.0;
  } else {
    for (int k = 0; k < 5; ++k) out[k] = s[k];
  }
}

int device_moments(dgv_ctx* ctx, const double* d_f, double out[5]) {
  const Grid& g = ctx->g;
  double* part = ctx->d_red;
  double* res = ctx->d_red + kRedBlocks * 5;
  hipLaunchKernelGGL(k_moments_part, dim3(kRedBlocks), dim3(256), 0, ctx->stream, 0, g.N, d_f, ctx->d_gwts,
                     ctx->d_u, ctx->d_v, ctx->d_w, res, part);
  hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(1), 0, ctx->stream, 0, part, res);
  hipLaunchKernelGGL(k_moments_part, dim3(kRedBlocks), dim3(256), 0, ctx->stream, 1, g.N, d_f, ctx->d_gwts,
                     ctx->d_u, ctx->d_v, ctx->d_w, res, part);
  hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(1), 0, ctx->stream, 1, part, res);
  DGV_HIP(ctx, hipGetLastError());
  DGV_HIP(ctx, hipMemcpyAsync(out, res, 5 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

int device_collision_moments(dgv_ctx* ctx, const double* d_q, double out[5]) {
  const Grid& g = ctx->g;
  double* part = ctx->d_red;
  double* res = ctx->d_red + kRedBlocks * 5;
  hipLaunchKernelGGL(k_moments_part, dim3(kRedBlocks), dim3(256), 0, ctx->stream, 2, g.N, d_q, ctx->d_gwts,
                     ctx->d_u, ctx->d_v, ctx->d_w, res, part);
  hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(1), 0, ctx->stream, 2, part, res);
  DGV_HIP(ctx, hipGetLastError());
  DGV_HIP(ctx, hipMemcpyAsync(out, res, 5 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

// MassCheckRecHighCMoments (dgvtools_mod.f90:1501-1549).  After the mass and
// momentum pass of k_moments_part, one pass of the central sums:
//   s[0] = sum f w |u - ubar|^2,  s[1 + 3 (k - 2) + d] = sum f w (u_d - ubar_d)^k, k = 2..6.
constexpr int kHiBlocks = 64;
constexpr int kHiSums = 16;

__global__ __launch_bounds__(256) void k_high_part(int64_t n, const double* __restrict__ f,
                                                   const double* __restrict__ w, const double* __restrict__ u,
                                                   const double* __restrict__ v, const double* __restrict__ ww,
                                                   const double* __restrict__ mean, double* __restrict__ part) {
  __shared__ double red[kHiSums][256];
  double s[kHiSums];
#pragma unroll
  for (int k = 0; k < kHiSums; ++k) s[k] = 0.0;
  const double ub = mean[1], vb = mean[2], wb = mean[3];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)kHiBlocks * 256) {
    const double fw = f[i] * w[i];
    const double du = u[i] - ub, dv = v[i] - vb, dw = ww[i] - wb;
    s[0] += fw * (du * du + dv * dv + dw * dw);
    double pu = du * du, pv = dv * dv, pw = dw * dw;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      s[1 + 3 * k] += fw * pu;
      s[2 + 3 * k] += fw * pv;
      s[3 + 3 * k] += fw * pw;
      pu *= du;
      pv *= dv;
      pw *= dw;
    }
  }
#pragma unroll
  for (int k = 0; k < kHiSums; ++k) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h)
      for (int k = 0; k < kHiSums; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < kHiSums; ++k) part[blockIdx.x * kHiSums + k] = red[k][0];
}

// out = {n, ubar, vbar, wbar, T, Tu, Tv, Tw, mom3_uvw, mom4_uvw, mom5_uvw, mom6_uvw}
__global__ void k_high_final(const double* __restrict__ part, const double* __restrict__ mean,
                             double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s[kHiSums];
  for (int k = 0; k < kHiSums; ++k) s[k] = 0.0;
  for (int b = 0; b < kHiBlocks; ++b)
    for (int k = 0; k < kHiSums; ++k) s[k] += part[b * kHiSums + k];
  const double n = mean[0];
  for (int k = 0; k < 4; ++k) out[k] = mean[k];
  out[4] = s[0] / n / 3.0 * 2.0;
  for (int d = 0; d < 3; ++d) out[5 + d] = s[1 + d] / n / 3.0 * 2.0;  // directional T (lines 1525-1527)
  for (int k = 1; k < 5; ++k)
    for (int d = 0; d < 3; ++d) out[8 + 3 * (k - 1) + d] = s[1 + 3 * k + d] / n;
}

int device_moments_high(dgv_ctx* ctx, const double* d_f, double out[20]) {
  const Grid& g = ctx->g;
  if (!ctx->d_hi && hipMalloc(&ctx->d_hi, sizeof(double) * (kHiBlocks * kHiSums + 32)) != hipSuccess) {
    ctx->d_hi = nullptr;
    return set_error(ctx, DGV_ERR_NOMEM, "hipMalloc(moment scratch) failed");
  }
  double* part = ctx->d_red;
  double* mean = ctx->d_red + kRedBlocks * 5;
  double* res = ctx->d_hi + kHiBlocks * kHiSums;
  hipLaunchKernelGGL(k_moments_part, dim3(kRedBlocks), dim3(256), 0, ctx->stream, 0, g.N, d_f, ctx->d_gwts,
                     ctx->d_u, ctx->d_v, ctx->d_w, mean, part);
  hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(1), 0, ctx->stream, 0, part, mean);
  hipLaunchKernelGGL(k_high_part, dim3(kHiBlocks), dim3(256), 0, ctx->stream, g.N, d_f, ctx->d_gwts, ctx->d_u,
                     ctx->d_v, ctx->d_w, mean, ctx->d_hi);
  hipLaunchKernelGGL(k_high_final, dim3(1), dim3(1), 0, ctx->stream, ctx->d_hi, mean, res);
  DGV_HIP(ctx, hipGetLastError());
  DGV_HIP(ctx, hipMemcpyAsync(out, res, 20 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  DGV_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DGV_OK;
}

// Projection onto the complement of span{1,u,v,w,|u|^2} in the w-weighted
// inner product: lambda = G^{-1} b with G_jk = sum w psi_j psi_k (Cholesky
// factor precomputed on the host in extended precision), q -= sum lambda_k psi_k.
__global__ void k_project_solve(const double* __restrict__ part, const double* __restrict__ L,
                                double* __restrict__ lam) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double b[5] = {0, 0, 0, 0, 0};
  for (int blk = 0; blk < kRedBlocks; ++blk)
    for (int k = 0; k < 5; ++k) b[k] += part[blk * 5 + k];
  double y[5];
  for (int i = 0; i < 5; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= L[i * 5 + k] * y[k];
    y[i] = s / L[i * 5 + i];
  }
  for (int i = 4; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < 5; ++k) s -= L[k * 5 + i] * lam[k];
    lam[i] = s / L[i * 5 + i];
  }
}

__global__ void k_project_apply(int64_t n, double* __restrict__ q, const double* __restrict__ u,
                                const double* __restrict__ v, const double* __restrict__ w,
                                const double* __restrict__ lam) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double uu = u[i], vv = v[i], ww = w[i];
  q[i] = q[i] - (lam[0] + lam[1] * uu + lam[2] * vv + lam[3] * ww + lam[4] * (uu * uu + vv * vv + ww * ww));
}

// scratch: kRedBlocks * 5 + 16 doubles for the partial moments and the
// multipliers (the context's reduction buffer unless given); the Cholesky
// factor is read from the context's buffer either way
int device_project(dgv_ctx* ctx, double* d_q, double* scratch) {
  const Grid& g = ctx->g;
  double* part = scratch ? scratch : ctx->d_red;
  double* lam = part + kRedBlocks * 5 + 8;
  double* L = ctx->d_red + kRedBlocks * 5 + 16;
  hipLaunchKernelGGL(k_moments_part, dim3(kRedBlocks), dim3(256), 0, ctx->stream, 2, g.N, d_q, ctx->d_gwts,
                     ctx->d_u, ctx->d_v, ctx->d_w, lam, part);
  hipLaunchKernelGGL(k_project_solve, dim3(1), dim3(1), 0, ctx->stream, part, L, lam);
  hipLaunchKernelGGL(k_project_apply, dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, ctx->stream, g.N, d_q,
                     ctx->d_u, ctx->d_v, ctx->d_w, lam);
  DGV_HIP(ctx, hipGetLastError());
  return DGV_OK;
}

}  // namespace dgv

using namespace dgv;

extern "C" int dgv_eval_q_dev_noalloc(dgv_ctx* ctx, const double* d_f, double* d_fc, int64_t n) {
  if (!ctx) {
    set_global_error("null context");
    return DGV_ERR_ARG;
  }
  if (n != ctx->g.N) return set_error(ctx, DGV_ERR_ARG, "array length differs from the number of nodes");
  if (!d_f || !d_fc) return set_error(ctx, DGV_ERR_ARG, "null device pointer");
  if (!ctx->has_a) return set_error(ctx, DGV_ERR_STATE, "no A-array set (dgv_ctx_set_aarray / generate)");
  double* const sp = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_ptr[ctx->device] : nullptr;
  if (!sp) return set_error(ctx, DGV_ERR_STATE, "the static scratch of this device was not set up");
  const Grid& g = ctx->g;
  const int k = resolve_kernel(ctx);
  DGV_HIP(ctx, hipSetDevice(ctx->device));
  int rc;
  size_t used = 0;
  if (k == DGV_KERNEL_RECORD) {   // the record kernel keeps its sums on chip
    rc = launch_eval_record(ctx, d_f, d_fc, 1.0);
  } else {
    used = (size_t)ctx->plan.nrows * (size_t)g.D;
    const size_t need = (used + (ctx->project ? kRedBlocks * 5 + 16 : 0)) * sizeof(double);
    if (need > kStaticScratchBytes)
      return set_error(ctx, DGV_ERR_UNSUPPORTED,
                       "partial sums of " + std::to_string(need) + " bytes exceed the static scratch of " +
                           std::to_string(kStaticScratchBytes) + " bytes: raise the tile (dgv_ctx_set_tile)");
    rc = launch_main_pairblock(ctx, d_f, k == DGV_KERNEL_MFMA, nullptr, sp);
    if (!rc) rc = launch_reduce(ctx, d_fc, 1.0, sp);
  }
  if (!rc && ctx->project) rc = device_project(ctx, d_fc, sp + used);
  return rc;
}
