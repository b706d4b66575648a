/*
 * dgv_oracle.c -- CPU restatement of the reference's direct DGV collision path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle, never the product): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  See dgv_oracle.h.
 *
 * Arithmetic follows the reference's Fortran expressions term by term (same
 * operand order, left-to-right evaluation, no FMA contraction: build with
 * -ffp-contract=off), so that on the same inputs the per-node sequential sums
 * round like the reference's serial code.  Citations are file:line of
 * alexmalekseenko/DGV0D3V_direct_collision.
 */
#include "dgv_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_D 3.141592653589793238462643383279502884197   /* nrtype.f90:17 */
#define PI25DT 3.141592653589793238462643                /* dgvtools_mod.f90:386 */

/* ------------------------------------------------------------------------- */
/* GauLeg (gaussian_mod.f90:4-53): Newton iteration on Legendre polynomials,
 * roots iterated independently (the WHERE mask of the Fortran), EPS=3e-14,
 * MAXIT=40.                                                                  */
int orc_gauleg(int n, double x1, double x2, double* x, double* w) {
  const double EPS = 3.0e-14;
  int m = (n + 1) / 2;
  double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  for (int i = 1; i <= m; i++) {
    double z = cos(PI_D * ((double)i - 0.25) / ((double)n + 0.5));
    double p1 = 0, p2 = 0, p3, pp = 0, z1;
    int its, done = 0;
    for (its = 1; its <= 40; its++) {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
      if (!(fabs(z - z1) > EPS)) { done = 1; break; }
    }
    if (!done) return -1;
    x[i - 1] = xm - xl * z;
    x[n - i] = xm + xl * z;
    w[i - 1] = 2.0 * xl / ((1.0 - z * z) * (pp * pp));
    w[n - i] = w[i - 1];
  }
  return 0;
}

/* g_nds_all(:,s) (miscset.f90:76-98): order 1 is special-cased. */
static int gauss_nodes(int s, double* x, double* w) {
  if (s < 1 || s > 16) return -1;
  if (s == 1) { x[0] = 0.0; w[0] = 2.0; return 0; }
  return orc_gauleg(s, -1.0, 1.0, x, w);
}

/* finduiviwi (dgvtools_mod.f90:302-337) on one 1-D grid of size ncap. */
static int find_index(const double* grid, int ncap, double u) {
  const double epsill = 10e-11;
  if (u < grid[0] - epsill || u > grid[ncap - 1] + epsill) return 0;
  int ui = 1;
  while (u >= grid[ui - 1] - epsill && ui < ncap) ui++;
  return ui;
}

void orc_geom_free(orc_geom* G) {
  if (!G) return;
  free(G->cells_lu); free(G->cells_ru); free(G->cells_lv); free(G->cells_rv);
  free(G->cells_lw); free(G->cells_rw);
  free(G->cells_ugi); free(G->cells_vgi); free(G->cells_wgi);
  free(G->nodes_u); free(G->nodes_v); free(G->nodes_w); free(G->nodes_gwts);
  free(G->nodes_pcell); free(G->nodes_ui); free(G->nodes_vi); free(G->nodes_wi);
  free(G->nodes_dui); free(G->nodes_dvi); free(G->nodes_dwi); free(G->nodes_phican);
  free(G);
}

int orc_geom_build(const orc_grid* g, orc_geom** out) {
  *out = NULL;
  if (g->Mu < 1 || g->Mv < 1 || g->Mw < 1 || g->su < 1 || g->sv < 1 || g->sw < 1) return -1;
  orc_geom* G = (orc_geom*)calloc(1, sizeof(orc_geom));
  G->g = *g;
  if (gauss_nodes(g->su, G->gn_u, G->gw_u) || gauss_nodes(g->sv, G->gn_v, G->gw_v) ||
      gauss_nodes(g->sw, G->gn_w, G->gw_w)) { orc_geom_free(G); return -2; }
  int Mu = g->Mu, Mv = g->Mv, Mw = g->Mw;
  /* OldCreateUmesh uniform branch (miscset.f90:458-459) */
  double* gu = (double*)malloc(sizeof(double) * (Mu + 1));
  double* gv = (double*)malloc(sizeof(double) * (Mv + 1));
  double* gw = (double*)malloc(sizeof(double) * (Mw + 1));
  for (int i = 0; i <= Mu; i++) gu[i] = g->u_L + (g->u_R - g->u_L) / (double)Mu * i;
  for (int i = 0; i <= Mv; i++) gv[i] = g->v_L + (g->v_R - g->v_L) / (double)Mv * i;
  for (int i = 0; i <= Mw; i++) gw[i] = g->w_L + (g->w_R - g->w_L) / (double)Mw * i;
  int64_t nc = (int64_t)Mu * Mv * Mw;
  G->ncells = nc;
  G->cells_lu = malloc(sizeof(double) * nc); G->cells_ru = malloc(sizeof(double) * nc);
  G->cells_lv = malloc(sizeof(double) * nc); G->cells_rv = malloc(sizeof(double) * nc);
  G->cells_lw = malloc(sizeof(double) * nc); G->cells_rw = malloc(sizeof(double) * nc);
  /* FillCellsArrsDGV (miscset.f90:915-927): u outer, w inner */
  int64_t mm = 0;
  for (int iu = 0; iu < Mu; iu++)
    for (int iv = 0; iv < Mv; iv++)
      for (int iw = 0; iw < Mw; iw++) {
        G->cells_lu[mm] = gu[iu]; G->cells_ru[mm] = gu[iu + 1];
        G->cells_lv[mm] = gv[iv]; G->cells_rv[mm] = gv[iv + 1];
        G->cells_lw[mm] = gw[iw]; G->cells_rw[mm] = gw[iw + 1];
        mm++;
      }
  /* SetNodesDGV (miscset.f90:1150-1176) */
  int dofc = g->su * g->sv * g->sw;
  int64_t nn = nc * dofc;
  G->nnodes = nn;
  G->nodes_u = malloc(sizeof(double) * nn); G->nodes_v = malloc(sizeof(double) * nn);
  G->nodes_w = malloc(sizeof(double) * nn); G->nodes_gwts = malloc(sizeof(double) * nn);
  G->nodes_pcell = malloc(sizeof(int32_t) * nn);
  G->nodes_ui = malloc(sizeof(int32_t) * nn); G->nodes_vi = malloc(sizeof(int32_t) * nn);
  G->nodes_wi = malloc(sizeof(int32_t) * nn);
  int64_t jm = 0;
  for (int64_t ic = 0; ic < nc; ic++) {
    double lu = G->cells_lu[ic], ru = G->cells_ru[ic];
    double lv = G->cells_lv[ic], rv = G->cells_rv[ic];
    double lw = G->cells_lw[ic], rw = G->cells_rw[ic];
    for (int iu = 1; iu <= g->su; iu++) {
      double rru = G->gn_u[iu - 1] * (ru - lu) / 2.0 + (ru + lu) / 2.0;
      double qu = G->gw_u[iu - 1];
      for (int iv = 1; iv <= g->sv; iv++) {
        double rrv = G->gn_v[iv - 1] * (rv - lv) / 2.0 + (rv + lv) / 2.0;
        double qv = G->gw_v[iv - 1];
        for (int iw = 1; iw <= g->sw; iw++) {
          double rrw = G->gn_w[iw - 1] * (rw - lw) / 2.0 + (rw + lw) / 2.0;
          double qw = G->gw_w[iw - 1];
          G->nodes_pcell[jm] = (int32_t)(ic + 1);
          G->nodes_u[jm] = rru; G->nodes_v[jm] = rrv; G->nodes_w[jm] = rrw;
          G->nodes_gwts[jm] = qu * qv * qw * (ru - lu) * (rv - lv) * (rw - lw) / 8.0;
          G->nodes_ui[jm] = iu; G->nodes_vi[jm] = iv; G->nodes_wi[jm] = iw;
          jm++;
        }
      }
    }
  }
  /* FindCellContainsPoint_DGV(0,0,0) (dgvtools_mod.f90:1175-1247), one grid */
  int ui = find_index(gu, Mu + 1, 0.0), vi = find_index(gv, Mv + 1, 0.0),
      wi = find_index(gw, Mw + 1, 0.0);
  free(gu); free(gv); free(gw);
  if (ui * vi * wi == 0) { orc_geom_free(G); return -3; }
  G->ccell = 1 + (ui - 2) * Mv * Mw + (vi - 2) * Mw + (wi - 2);
  /* SetCellsUGINdsSftArrs (miscset.f90:1202-1311): the cell-address loop walks
   * the enumeration w fastest, exactly as lines 1248-1277. */
  G->cells_ugi = malloc(sizeof(int32_t) * nc); G->cells_vgi = malloc(sizeof(int32_t) * nc);
  G->cells_wgi = malloc(sizeof(int32_t) * nc);
  {
    int cu = 1, cv = 1, cw = 1;
    for (int64_t phicell = 1; phicell <= nc; phicell++) {
      if (phicell > 1) {
        if (cw == Mw) {
          cw = 1;
          if (cv == Mv) { cv = 1; cu++; } else cv++;
        } else cw++;
      }
      G->cells_ugi[phicell - 1] = cu; G->cells_vgi[phicell - 1] = cv;
      G->cells_wgi[phicell - 1] = cw;
    }
  }
  G->nodes_dui = malloc(sizeof(int32_t) * nn); G->nodes_dvi = malloc(sizeof(int32_t) * nn);
  G->nodes_dwi = malloc(sizeof(int32_t) * nn); G->nodes_phican = malloc(sizeof(int32_t) * nn);
  int32_t cc = G->ccell;
  int64_t my_cshift = (int64_t)(cc - 1) * dofc;
  for (int64_t i = 0; i < nn; i++) {
    int32_t pc = G->nodes_pcell[i];
    G->nodes_dui[i] = G->cells_ugi[pc - 1] - G->cells_ugi[cc - 1];
    G->nodes_dvi[i] = G->cells_vgi[pc - 1] - G->cells_vgi[cc - 1];
    G->nodes_dwi[i] = G->cells_wgi[pc - 1] - G->cells_wgi[cc - 1];
    G->nodes_phican[i] = (int32_t)(my_cshift + (G->nodes_ui[i] - 1) * g->sv * g->sw +
                                   (G->nodes_vi[i] - 1) * g->sw + G->nodes_wi[i]);
  }
  *out = G;
  return 0;
}

int orc_ashift(const orc_geom* G, const int32_t* capphi, int64_t* ashift) {
  /* nodes_Ashift(iphi) = sum(A_capphi(1:phishift-1)) (miscset.f90:1307) */
  int64_t nn = G->nnodes;
  int64_t* pref = (int64_t*)malloc(sizeof(int64_t) * (nn + 1));
  pref[0] = 0;
  for (int64_t k = 0; k < nn; k++) pref[k + 1] = pref[k] + capphi[k];
  for (int64_t i = 0; i < nn; i++) ashift[i] = pref[G->nodes_phican[i] - 1];
  free(pref);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Shared bounds-tested translation of one record for node iphi
 * (collision_mod.f90:196-211).  Returns 1 and the two 1-based f indices when
 * both translated nodes stay on the grid.                                    */
static inline int shift_record(const orc_geom* G, int32_t ixi, int32_t ixi1, int dui, int dvi,
                               int dwi, int64_t* kxi, int64_t* kxi1) {
  const int pgcu = G->g.Mu, pgcv = G->g.Mv, pgcw = G->g.Mw;
  const int gov = G->g.sv, gow = G->g.sw, dofc = G->g.su * G->g.sv * G->g.sw;
  int32_t xicell = G->nodes_pcell[ixi - 1];
  int iux = G->cells_ugi[xicell - 1] + dui;
  if (!(iux >= 1 && iux <= pgcu)) return 0;
  int ivx = G->cells_vgi[xicell - 1] + dvi;
  if (!(ivx >= 1 && ivx <= pgcv)) return 0;
  int iwx = G->cells_wgi[xicell - 1] + dwi;
  if (!(iwx >= 1 && iwx <= pgcw)) return 0;
  int32_t xi1cell = G->nodes_pcell[ixi1 - 1];
  int iux1 = G->cells_ugi[xi1cell - 1] + dui;
  if (!(iux1 >= 1 && iux1 <= pgcu)) return 0;
  int ivx1 = G->cells_vgi[xi1cell - 1] + dvi;
  if (!(ivx1 >= 1 && ivx1 <= pgcv)) return 0;
  int iwx1 = G->cells_wgi[xi1cell - 1] + dwi;
  if (!(iwx1 >= 1 && iwx1 <= pgcw)) return 0;
  *kxi = ((int64_t)(iux - 1) * pgcw * pgcv + (int64_t)(ivx - 1) * pgcw + iwx - 1) * dofc +
         (G->nodes_ui[ixi - 1] - 1) * gow * gov + (G->nodes_vi[ixi - 1] - 1) * gow +
         G->nodes_wi[ixi - 1];
  *kxi1 = ((int64_t)(iux1 - 1) * pgcw * pgcv + (int64_t)(ivx1 - 1) * pgcw + iwx1 - 1) * dofc +
          (G->nodes_ui[ixi1 - 1] - 1) * gow * gov + (G->nodes_vi[ixi1 - 1] - 1) * gow +
          G->nodes_wi[ixi1 - 1];
  return 1;
}

static int set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
  return nthreads;
#else
  (void)nthreads;
  return 1;
#endif
}

int orc_eval_aplus(const orc_geom* G, const int32_t* capphi, const double* A,
                   const int32_t* A_xi, const int32_t* A_xi1, const double* f, double* fc,
                   int64_t node_begin, int64_t node_end, int nthreads) {
  int64_t* ashift = (int64_t*)malloc(sizeof(int64_t) * G->nnodes);
  orc_ashift(G, capphi, ashift);
  int nt = set_threads(nthreads);
  /* fc=0 (line 184), then one node at a time (lines 186-227) */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t iphi = node_begin; iphi < node_end; iphi++) {
    int64_t Ashift = ashift[iphi];
    int32_t phicap = capphi[G->nodes_phican[iphi] - 1];
    int dui = G->nodes_dui[iphi], dvi = G->nodes_dvi[iphi], dwi = G->nodes_dwi[iphi];
    double acc = 0.0;
    for (int32_t j = 1; j <= phicap; j++) {
      int64_t k1, k2;
      if (shift_record(G, A_xi[Ashift + j - 1], A_xi1[Ashift + j - 1], dui, dvi, dwi, &k1, &k2))
        acc = acc + 2 * f[k1 - 1] * f[k2 - 1] * A[Ashift + j - 1];
    }
    fc[iphi] = acc / G->nodes_gwts[iphi];
  }
  free(ashift);
  return 0;
}

int orc_eval_mixed(const orc_geom* G, const int32_t* capphi, const double* A,
                   const int32_t* A_xi, const int32_t* A_xi1, const double* f, const double* fm,
                   double* fc, int nthreads) {
  int64_t* ashift = (int64_t*)malloc(sizeof(int64_t) * G->nnodes);
  orc_ashift(G, capphi, ashift);
  int nt = set_threads(nthreads);
  /* EvalCollisionPeriodicMixedTermsA_DGV: fc=0, per node the shifted records,
   * fc += (fm(xi) f(xi1) + f(xi) fm(xi1)) A, then fc = 2 fc / w (lines 283-328) */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t iphi = 0; iphi < G->nnodes; iphi++) {
    int64_t Ashift = ashift[iphi];
    int32_t phicap = capphi[G->nodes_phican[iphi] - 1];
    int dui = G->nodes_dui[iphi], dvi = G->nodes_dvi[iphi], dwi = G->nodes_dwi[iphi];
    double acc = 0.0;
    for (int32_t j = 1; j <= phicap; j++) {
      int64_t k1, k2;
      if (shift_record(G, A_xi[Ashift + j - 1], A_xi1[Ashift + j - 1], dui, dvi, dwi, &k1, &k2))
        acc = acc + (fm[k1 - 1] * f[k2 - 1] + f[k1 - 1] * fm[k2 - 1]) * A[Ashift + j - 1];
    }
    fc[iphi] = 2 * acc / G->nodes_gwts[iphi];
  }
  free(ashift);
  return 0;
}

int orc_prepare_fma(const orc_geom* G, const int32_t* capphi, const double* A,
                    const int32_t* A_xi, const int32_t* A_xi1, const double* fm, double* fmA) {
  /* PrepareFMA_DGV (collision_mod.f90:377-464): column iphi of the dense N x N
   * fmA (column-major, fmA[row + N col]) collects fm(xi1) A at row xi and
   * fm(xi) A at row xi1 over the node's shifted records, then is divided by w */
  const int64_t N = G->nnodes;
  int64_t* ashift = (int64_t*)malloc(sizeof(int64_t) * N);
  orc_ashift(G, capphi, ashift);
  memset(fmA, 0, sizeof(double) * N * N);
  for (int64_t iphi = 0; iphi < N; iphi++) {
    double* col = fmA + iphi * N;
    int64_t Ashift = ashift[iphi];
    int32_t phicap = capphi[G->nodes_phican[iphi] - 1];
    int dui = G->nodes_dui[iphi], dvi = G->nodes_dvi[iphi], dwi = G->nodes_dwi[iphi];
    for (int32_t j = 1; j <= phicap; j++) {
      int64_t k1, k2;
      if (shift_record(G, A_xi[Ashift + j - 1], A_xi1[Ashift + j - 1], dui, dvi, dwi, &k1, &k2)) {
        col[k1 - 1] = col[k1 - 1] + fm[k2 - 1] * A[Ashift + j - 1];
        col[k2 - 1] = col[k2 - 1] + fm[k1 - 1] * A[Ashift + j - 1];
      }
    }
    for (int64_t r = 0; r < N; r++) col[r] = col[r] / G->nodes_gwts[iphi];
  }
  free(ashift);
  return 0;
}

void orc_eval_linear(int64_t N, const double* fmA, const double* f, double* fc) {
  /* EvalCollisionLinear (collision_mod.f90:346-359): fc = 2 MATMUL(f, fmA) */
  for (int64_t i = 0; i < N; i++) {
    const double* col = fmA + i * N;
    double s = 0.0;
    for (int64_t r = 0; r < N; r++) s = s + f[r] * col[r];
    fc[i] = 2 * s;
  }
}

int orc_eval_aplus_list(const orc_geom* G, const int32_t* capphi, const double* A,
                        const int32_t* A_xi, const int32_t* A_xi1, const double* f, double* fc,
                        const int64_t* nodes, int64_t nsel, int nthreads) {
  int64_t* ashift = (int64_t*)malloc(sizeof(int64_t) * G->nnodes);
  orc_ashift(G, capphi, ashift);
  int nt = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t t = 0; t < nsel; t++) {
    const int64_t iphi = nodes[t];
    int64_t Ashift = ashift[iphi];
    int32_t phicap = capphi[G->nodes_phican[iphi] - 1];
    int dui = G->nodes_dui[iphi], dvi = G->nodes_dvi[iphi], dwi = G->nodes_dwi[iphi];
    double acc = 0.0;
    for (int32_t j = 1; j <= phicap; j++) {
      int64_t k1, k2;
      if (shift_record(G, A_xi[Ashift + j - 1], A_xi1[Ashift + j - 1], dui, dvi, dwi, &k1, &k2))
        acc = acc + 2 * f[k1 - 1] * f[k2 - 1] * A[Ashift + j - 1];
    }
    fc[t] = acc / G->nodes_gwts[iphi];
  }
  free(ashift);
  return 0;
}

int orc_eval_a(const orc_geom* G, const int32_t* capphi, const double* A,
               const int32_t* A_xi, const int32_t* A_xi1, const double* f, double* fc,
               int64_t node_begin, int64_t node_end, int nthreads) {
  int64_t* ashift = (int64_t*)malloc(sizeof(int64_t) * G->nnodes);
  orc_ashift(G, capphi, ashift);
  int nt = set_threads(nthreads);
  int err = 0;
#pragma omp parallel num_threads(nt)
  {
    int64_t cap = 0;
    double *fxi = NULL, *fxi1 = NULL, *Aphi = NULL;
#pragma omp for schedule(dynamic, 1)
    for (int64_t iphi = node_begin; iphi < node_end; iphi++) {
      int64_t Ashift = ashift[iphi];
      int32_t phicap = capphi[G->nodes_phican[iphi] - 1];
      if (phicap > cap) { /* scratch arrays of line 76 */
        free(fxi); free(fxi1); free(Aphi);
        cap = phicap;
        fxi = malloc(sizeof(double) * cap); fxi1 = malloc(sizeof(double) * cap);
        Aphi = malloc(sizeof(double) * cap);
        if (!fxi || !fxi1 || !Aphi) { err = 1; cap = 0; continue; }
      }
      int dui = G->nodes_dui[iphi], dvi = G->nodes_dvi[iphi], dwi = G->nodes_dwi[iphi];
      int64_t k = 0;
      for (int32_t j = 1; j <= phicap; j++) {
        int64_t k1, k2;
        if (shift_record(G, A_xi[Ashift + j - 1], A_xi1[Ashift + j - 1], dui, dvi, dwi, &k1, &k2)) {
          fxi[k] = f[k1 - 1]; fxi1[k] = f[k2 - 1]; Aphi[k] = A[Ashift + j - 1];
          k++;
        }
      }
      /* fc(iphi)=2*sum(Aphi(1:k)*fxi(1:k)*fxi1(1:k))/nodes_gwts(iphi) (line 125) */
      double s = 0.0;
      for (int64_t t = 0; t < k; t++) s = s + Aphi[t] * fxi[t] * fxi1[t];
      fc[iphi] = 2 * s / G->nodes_gwts[iphi];
    }
    free(fxi); free(fxi1); free(Aphi);
  }
  free(ashift);
  return err ? -1 : 0;
}

int orc_eval_aplus_comp(const orc_geom* G, const int32_t* capphi, const double* A,
                        const int32_t* A_xi, const int32_t* A_xi1, const double* f, double* fc,
                        double* abs_sum, int64_t node_begin, int64_t node_end, int nthreads) {
  int64_t* ashift = (int64_t*)malloc(sizeof(int64_t) * G->nnodes);
  orc_ashift(G, capphi, ashift);
  int nt = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t iphi = node_begin; iphi < node_end; iphi++) {
    int64_t Ashift = ashift[iphi];
    int32_t phicap = capphi[G->nodes_phican[iphi] - 1];
    int dui = G->nodes_dui[iphi], dvi = G->nodes_dvi[iphi], dwi = G->nodes_dwi[iphi];
    double s = 0.0, c = 0.0, sa = 0.0;
    for (int32_t j = 1; j <= phicap; j++) {
      int64_t k1, k2;
      if (shift_record(G, A_xi[Ashift + j - 1], A_xi1[Ashift + j - 1], dui, dvi, dwi, &k1, &k2)) {
        double t = 2 * f[k1 - 1] * f[k2 - 1] * A[Ashift + j - 1];
        /* Neumaier's improved Kahan summation */
        double u = s + t;
        if (fabs(s) >= fabs(t)) c += (s - u) + t; else c += (t - u) + s;
        s = u;
        sa += fabs(t);
      }
    }
    fc[iphi] = (s + c) / G->nodes_gwts[iphi];
    if (abs_sum) abs_sum[iphi] = sa / G->nodes_gwts[iphi];
  }
  free(ashift);
  return 0;
}

/* a1 at a list of nodes with every term and the sum carried in double-double:
 * each product f(xi') f(xi1') A is formed exactly (two-product with fma), and
 * the terms are added with two-sum into a (hi, lo) pair, so the result is Q
 * rounded once, whatever the conditioning.  The reference value for kernels whose
 * distance from the reference's sequential FP64 sum is dominated by that sum's
 * own rounding.  abs_sum[t] = sum |terms| / w (the terms' scale). */
static inline void two_prod(double a, double b, double* p, double* e) {
  *p = a * b;
  *e = fma(a, b, -*p);
}
static inline void two_sum_dd(double a, double b, double* s, double* e) {
  *s = a + b;
  double bp = *s - a;
  *e = (a - (*s - bp)) + (b - bp);
}
int orc_eval_aplus_dd_list(const orc_geom* G, const int32_t* capphi, const double* A,
                           const int32_t* A_xi, const int32_t* A_xi1, const double* f, double* fc,
                           double* abs_sum, const int64_t* nodes, int64_t nsel, int nthreads) {
  int64_t* ashift = (int64_t*)malloc(sizeof(int64_t) * G->nnodes);
  orc_ashift(G, capphi, ashift);
  int nt = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t t = 0; t < nsel; t++) {
    const int64_t iphi = nodes[t];
    int64_t Ashift = ashift[iphi];
    int32_t phicap = capphi[G->nodes_phican[iphi] - 1];
    int dui = G->nodes_dui[iphi], dvi = G->nodes_dvi[iphi], dwi = G->nodes_dwi[iphi];
    double sh = 0.0, sl = 0.0, sa = 0.0;
    for (int32_t j = 1; j <= phicap; j++) {
      int64_t k1, k2;
      if (shift_record(G, A_xi[Ashift + j - 1], A_xi1[Ashift + j - 1], dui, dvi, dwi, &k1, &k2)) {
        double ph, pl, th, tl, s, e;
        two_prod(f[k1 - 1], f[k2 - 1], &ph, &pl);          /* f f = ph + pl exactly */
        two_prod(ph, A[Ashift + j - 1], &th, &tl);           /* ph A = th + tl exactly */
        tl += pl * A[Ashift + j - 1];                        /* + pl A (below 2^-106 relative) */
        two_sum_dd(sh, th, &s, &e);
        sh = s;
        sl += e + tl;
        sa += fabs(th);
      }
    }
    /* factor 2 is exact; (sh + sl) / w rounded once */
    double q = 2.0 * sh, ql = 2.0 * sl;
    double qh = q + ql;
    fc[t] = (qh + ((q - qh) + ql)) / G->nodes_gwts[iphi];
    if (abs_sum) abs_sum[t] = 2.0 * sa / G->nodes_gwts[iphi];
  }
  free(ashift);
  return 0;
}

int orc_iphi(int64_t nnodes, const int32_t* capphi, const double* A, const int32_t* A_xi,
             const int32_t* A_xi1, const double* gwts, const double* f, double coef,
             double* out) {
  int64_t A_ct = 0;
  for (int64_t i = 0; i < nnodes; i++) {
    double s = 0.0;
    for (int64_t j = A_ct; j < A_ct + capphi[i]; j++)
      s = s + A[j] * f[A_xi[j] - 1] * f[A_xi1[j] - 1];
    out[i] = 2 * s / gwts[i] * coef;
    A_ct += capphi[i];
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
void orc_masscheck(const orc_geom* G, const double* f, double out[5]) {
  int64_t nn = G->nnodes;
  const double *w = G->nodes_gwts, *u = G->nodes_u, *v = G->nodes_v, *ww = G->nodes_w;
  double n = 0, su = 0, sv = 0, sw = 0, st = 0;
  for (int64_t i = 0; i < nn; i++) n = n + f[i] * w[i];
  for (int64_t i = 0; i < nn; i++) su = su + f[i] * w[i] * u[i];
  for (int64_t i = 0; i < nn; i++) sv = sv + f[i] * w[i] * v[i];
  for (int64_t i = 0; i < nn; i++) sw = sw + f[i] * w[i] * ww[i];
  double ub = su / n, vb = sv / n, wb = sw / n;
  for (int64_t i = 0; i < nn; i++) {
    double du = u[i] - ub, dv = v[i] - vb, dw = ww[i] - wb;
    st = st + f[i] * w[i] * (du * du + dv * dv + dw * dw);
  }
  out[0] = n; out[1] = ub; out[2] = vb; out[3] = wb;
  out[4] = st / n / 3.0 * 2.0;
}

void orc_masscheck_high(const orc_geom* G, const double* f, double out[20]) {
  /* MassCheckRecHighCMoments (dgvtools_mod.f90:1501-1549): every sum its own
   * loop in node order, (x - xbar)**k as repeated products */
  int64_t nn = G->nnodes;
  const double *w = G->nodes_gwts, *u = G->nodes_u, *v = G->nodes_v, *ww = G->nodes_w;
  double base[5];
  orc_masscheck(G, f, base);
  const double n = base[0], mean[3] = {base[1], base[2], base[3]};
  const double* x[3] = {u, v, ww};
  for (int k = 0; k < 5; k++) out[k] = base[k];
  for (int d = 0; d < 3; d++) {
    double s = 0;
    for (int64_t i = 0; i < nn; i++) {
      double t = x[d][i] - mean[d];
      s = s + f[i] * w[i] * (t * t);
    }
    out[5 + d] = s / n / 3.0 * 2.0;
  }
  for (int k = 3; k <= 6; k++)
    for (int d = 0; d < 3; d++) {
      double s = 0;
      for (int64_t i = 0; i < nn; i++) {
        double t = x[d][i] - mean[d], p = t;
        for (int j = 1; j < k; j++) p = p * t;
        s = s + f[i] * w[i] * p;
      }
      out[8 + 3 * (k - 3) + d] = s / n;
    }
}

double orc_maxwell(double T, double u0, double v0, double w0, double n, double u, double v,
                   double w) {
  double beta = sqrt(PI25DT * T) * (PI25DT * T);
  double tt = T > 0.0000001 ? T : 0.0000001;
  return n * exp(-((u - u0) * (u - u0) + (v - v0) * (v - v0) + (w - w0) * (w - w0)) / tt) / beta;
}

int orc_spatial_oper(orc_q_fn q, void* user, int64_t n, const double* f, double* fcol,
                     double dt, double coef) {
  int rc = q(user, f, fcol);
  if (rc) return rc;
  double coef_temp = coef * dt; /* (mol_diam/L_inf)**2*N_inf*dt, line 65 */
  for (int64_t i = 0; i < n; i++) fcol[i] = fcol[i] * coef_temp;
  return 0;
}

int orc_rk_step(orc_q_fn q, void* user, int64_t n, double* f, double dt, double coef,
                int rkmts, double* frhs1) {
  double *ft2 = NULL, *ft3 = NULL, *ft4 = NULL, *tmp = NULL;
  int rc = 0;
  if (rkmts < 1 || rkmts > 5) return -1;
  ft2 = malloc(sizeof(double) * n); ft3 = malloc(sizeof(double) * n);
  ft4 = malloc(sizeof(double) * n); tmp = malloc(sizeof(double) * n);
#define SOP(in, out) do { rc = orc_spatial_oper(q, user, n, in, out, dt, coef); if (rc) goto done; } while (0)
  switch (rkmts) {
    case 1: /* lines 37-40 */
      SOP(f, frhs1);
      for (int64_t i = 0; i < n; i++) f[i] = f[i] + frhs1[i];
      break;
    case 2: /* lines 47-61 */
      SOP(f, frhs1);
      for (int64_t i = 0; i < n; i++) tmp[i] = f[i] + frhs1[i] * 2 / 3.0;
      SOP(tmp, ft2);
      for (int64_t i = 0; i < n; i++) f[i] = f[i] + frhs1[i] / 4.0 + ft2[i] * 3.0 / 4;
      break;
    case 3: /* lines 64-83 */
      SOP(f, frhs1);
      for (int64_t i = 0; i < n; i++) tmp[i] = f[i] + frhs1[i] / 2.0;
      SOP(tmp, ft2);
      for (int64_t i = 0; i < n; i++) tmp[i] = f[i] + ft2[i] * 3.0 / 4;
      SOP(tmp, ft3);
      for (int64_t i = 0; i < n; i++)
        f[i] = f[i] + frhs1[i] * 2 / 9.0 + ft2[i] * 3.0 / 9.0 + ft3[i] * 4 / 9.0;
      break;
    case 4: /* lines 85-105 */
      SOP(f, frhs1);
      for (int64_t i = 0; i < n; i++) tmp[i] = f[i] + frhs1[i] / 2.0;
      SOP(tmp, ft2);
      for (int64_t i = 0; i < n; i++) tmp[i] = f[i] + frhs1[i] / 4.0 + ft2[i] / 4.0;
      SOP(tmp, ft3);
      for (int64_t i = 0; i < n; i++) tmp[i] = f[i] - ft2[i] + ft3[i] * 2;
      SOP(tmp, ft2);
      for (int64_t i = 0; i < n; i++)
        f[i] = f[i] + frhs1[i] / 6.0 + ft3[i] * 2 / 3.0 + ft2[i] / 6.0;
      break;
    case 5: /* lines 107-145 */
      SOP(f, frhs1);
      for (int64_t i = 0; i < n; i++) tmp[i] = f[i] + frhs1[i] / 2.0;
      SOP(tmp, ft2);
      for (int64_t i = 0; i < n; i++) tmp[i] = f[i] + frhs1[i] / 4.0 + ft2[i] / 4.0;
      SOP(tmp, ft3);
      for (int64_t i = 0; i < n; i++) tmp[i] = f[i] - ft2[i] + ft3[i] * 2.0;
      SOP(tmp, ft4);
      for (int64_t i = 0; i < n; i++) ft3[i] = -ft2[i] / 5.0 + ft3[i] * 546 / 625.0;
      for (int64_t i = 0; i < n; i++)
        tmp[i] = f[i] + frhs1[i] * 7.0 / 27.0 + ft2[i] * 10.0 / 27.0 + ft4[i] / 27.0;
      SOP(tmp, ft2);
      for (int64_t i = 0; i < n; i++)
        tmp[i] = f[i] + frhs1[i] * 28.0 / 625.0 + ft3[i] + ft4[i] * 54 / 625.0 - ft2[i] * 378 / 625.0;
      SOP(tmp, ft3);
      for (int64_t i = 0; i < n; i++)
        f[i] = f[i] + frhs1[i] / 24.0 + ft4[i] * 5.0 / 48.0 + ft2[i] * 27.0 / 56.0 + ft3[i] * 125 / 336.0;
      break;
  }
  for (int64_t i = 0; i < n; i++) frhs1[i] = frhs1[i] / dt; /* line 151 */
done:
#undef SOP
  free(ft2); free(ft3); free(ft4); free(tmp);
  return rc;
}

int orc_ab_step(orc_q_fn q, void* user, int64_t n, double* f, double dt, double coef,
                int rkmts, double* hist) {
  double *f1 = hist, *f2 = hist + n, *f3 = hist + 2 * n, *f4 = hist + 3 * n;
  double *r1 = hist + 4 * n, *r2 = hist + 5 * n, *r3 = hist + 6 * n, *r4 = hist + 7 * n,
         *r5 = hist + 8 * n;
  int rc;
  if (rkmts < 1 || rkmts > 5) return -1;
  /* push the history (time_integr_mod.f90:333,339,345,352-353) */
  for (int64_t i = 0; i < n; i++) {
    if (rkmts >= 5) { f4[i] = f3[i]; r5[i] = r4[i]; }
    if (rkmts >= 4) { f3[i] = f2[i]; r4[i] = r3[i]; }
    if (rkmts >= 3) { f2[i] = f1[i]; r3[i] = r2[i]; }
    if (rkmts >= 2) { f1[i] = f[i]; r2[i] = r1[i]; }
  }
  rc = orc_spatial_oper(q, user, n, f, r1, dt, coef);
  if (rc) return rc;
  for (int64_t i = 0; i < n; i++) {
    switch (rkmts) {
      case 1: f[i] = f[i] + r1[i]; break;
      case 2: f[i] = f[i] + r1[i] * 3.0 / 2.0 - dt * r2[i] / 2.0; break;
      case 3:
        f[i] = f[i] + r1[i] * 23.0 / 12.0 - dt * r2[i] * 4 / 3.0 + dt * r3[i] * 5.0 / 12.0;
        break;
      case 4:
        f[i] = f[i] + r1[i] * 55.0 / 24.0 - dt * r2[i] * 59.0 / 24.0 + dt * r3[i] * 37.0 / 24.0 -
               dt * r4[i] * 3.0 / 8;
        break;
      case 5:
        f[i] = f[i] + r1[i] * (1901.0 / 720.0) - dt * r2[i] * (2774.0 / 720.0) +
               dt * r3[i] * (2616.0 / 720.0) - dt * r4[i] * (1274.0 / 720.0) +
               dt * r5[i] * (251.0 / 720.0);
        break;
    }
  }
  for (int64_t i = 0; i < n; i++) r1[i] = r1[i] / dt; /* line 362 */
  return 0;
}

/* ------------------------------------------------------------------------- */
/* A-array generation restatement (dgvtools_mod.f90).                         */

static double lagrbasfun(int i, double x, const double* kappa, int n) {
  /* dgvtools_mod.f90:31-56 */
  double y = 1.0;
  for (int j = 1; j <= i - 1; j++) y = y * (kappa[j - 1] - x) / (kappa[j - 1] - kappa[i - 1]);
  for (int j = i + 1; j <= n; j++) y = y * (kappa[j - 1] - x) / (kappa[j - 1] - kappa[i - 1]);
  return y;
}

double orc_basis(const orc_geom* G, int mode, double u, double v, double w, int32_t i) {
  /* EvalLagrBasisFunByNdsDGblzmEZ (dgvtools_mod.f90:232-289) */
  switch (mode) {
    case 1: return 1.0;
    case 2: return u;
    case 3: return v;
    case 4: return w;
    case 5: return u * u + v * v + w * w;
    default: break;
  }
  int32_t cellp = G->nodes_pcell[i - 1];
  double lu = G->cells_lu[cellp - 1], ru = G->cells_ru[cellp - 1];
  double lv = G->cells_lv[cellp - 1], rv = G->cells_rv[cellp - 1];
  double lw = G->cells_lw[cellp - 1], rw = G->cells_rw[cellp - 1];
  if (u < lu || ru < u || v < lv || rv < v || w < lw || rw < w) return 0.0;
  double y = 1.0;
  double unor = (u - (ru + lu) / 2.0) / (ru - lu) * 2.0;
  y = y * lagrbasfun(G->nodes_ui[i - 1], unor, G->gn_u, G->g.su);
  double vnor = (v - (rv + lv) / 2.0) / (rv - lv) * 2.0;
  y = y * lagrbasfun(G->nodes_vi[i - 1], vnor, G->gn_v, G->g.sv);
  double wnor = (w - (rw + lw) / 2.0) / (rw - lw) * 2.0;
  y = y * lagrbasfun(G->nodes_wi[i - 1], wnor, G->gn_w, G->g.sw);
  return y;
}

typedef struct {
  const orc_geom* G;
  int mode;
  int32_t i1;
  double xiu, xiv, xiw, xi1u, xi1v, xi1w;
  double ugu, ugv, ugw, ugux2, ugvx2, ugwx2, ugux3, ugvx3, ugwx3, g1, g2, dsph;
  double varphi_xi, varphi_xi1;
  double ires_e, ErrEps;
  /* scratch of A_IntEpsilon (NdsEps=10000) */
  double *N1, *N2, *F1, *F2;
  int *R1, *R2;
} seta_ctx;

enum { NDS = 10000 };

static double integrand(const seta_ctx* c, double epsil, double sinchi, double coschi) {
  /* EvalPhiPostCollIntegrand (dgvtools_mod.f90:938-975) */
  double sinchicoseps = cos(epsil) * sinchi / c->g1;
  double sinchisineps = sin(epsil) * sinchi / c->g2;
  double ku = (c->ugu) * coschi - c->ugux2 * sinchicoseps - c->ugux3 * sinchisineps;
  double kv = (c->ugv) * coschi - c->ugvx2 * sinchicoseps - c->ugvx3 * sinchisineps;
  double kw = (c->ugw) * coschi - c->ugwx2 * sinchicoseps - c->ugwx3 * sinchisineps;
  double xiupr = ((c->xiu + c->xi1u) + c->dsph * ku) / 2.0;
  double xi1upr = ((c->xiu + c->xi1u) - c->dsph * ku) / 2.0;
  double xivpr = ((c->xiv + c->xi1v) + c->dsph * kv) / 2.0;
  double xi1vpr = ((c->xiv + c->xi1v) - c->dsph * kv) / 2.0;
  double xiwpr = ((c->xiw + c->xi1w) + c->dsph * kw) / 2.0;
  double xi1wpr = ((c->xiw + c->xi1w) - c->dsph * kw) / 2.0;
  return orc_basis(c->G, c->mode, xiupr, xivpr, xiwpr, c->i1) +
         orc_basis(c->G, c->mode, xi1upr, xi1vpr, xi1wpr, c->i1) - c->varphi_xi - c->varphi_xi1;
}

/* A_IntEpsilon (dgvtools_mod.f90:988-1122).  *err set on node overflow.   */
static double a_int_epsilon(seta_ctx* c, double chi, int* err) {
  double sinchi = sin(chi), coschi = cos(chi);
  double *N1 = c->N1, *N2 = c->N2, *F1 = c->F1, *F2 = c->F2;
  int *R1 = c->R1, *R2 = c->R2;
  int Neps1 = (int)floor(PI25DT * c->dsph * sinchi / c->ires_e) + 1;
  if (Neps1 > NDS) Neps1 = NDS;
  for (int f = 1; f <= Neps1; f++) {
    N1[2 * (f - 1)] = (f - 1) * 2 * PI25DT / (double)Neps1;
    N1[2 * f - 1] = f * 2 * PI25DT / (double)Neps1;
  }
  for (int f = 1; f <= Neps1; f++) {
    if (f > 1 && N1[2 * f - 2] == N1[2 * f - 3]) F1[3 * f - 3] = F1[3 * f - 4];
    else F1[3 * f - 3] = integrand(c, N1[2 * f - 2], sinchi, coschi);
    F1[3 * f - 1] = integrand(c, N1[2 * f - 1], sinchi, coschi);
    F1[3 * f - 2] = integrand(c, (N1[2 * f - 1] + N1[2 * f - 2]) / 2, sinchi, coschi);
  }
  for (int f = 1; f <= Neps1; f++) R1[f - 1] = 1;
  double my_int1 = 10 * c->ErrEps;
  int flag = 1;
  double at = 0;
  while (flag) {
    flag = 0;
    int g = 0;
    if (my_int1 > c->ErrEps) {
      my_int1 = 0;
      for (int f = 1; f <= Neps1; f++) {
        if ((fabs(F1[3 * f - 3]) + fabs(F1[3 * f - 2]) + fabs(F1[3 * f - 1]) > 1.0e-15) &&
            R1[f - 1] == 1) {
          if (g > NDS - 2) { *err = 1; return 0.0; }
          N2[2 * g] = N1[2 * f - 2];
          F2[3 * g] = F1[3 * f - 3];
          N2[2 * g + 3] = N1[2 * f - 1];
          F2[3 * g + 5] = F1[3 * f - 1];
          N2[2 * g + 1] = (N1[2 * f - 2] + N1[2 * f - 1]) / 2;
          N2[2 * g + 2] = N2[2 * g + 1];
          F2[3 * g + 2] = F1[3 * f - 2];
          F2[3 * g + 3] = F2[3 * g + 2];
          F2[3 * g + 1] = integrand(c, (N2[2 * g + 1] + N2[2 * g]) / 2.0, sinchi, coschi);
          F2[3 * g + 4] = integrand(c, (N2[2 * g + 3] + N2[2 * g + 2]) / 2.0, sinchi, coschi);
          double quad = (F1[3 * f - 3] + 4 * F1[3 * f - 2] + F1[3 * f - 1]) *
                        (N1[2 * f - 1] - N1[2 * f - 2]) / 6;
          double quad1 = (F2[3 * g] + 4 * F2[3 * g + 1] + F2[3 * g + 2]) * (N2[2 * g + 1] - N2[2 * g]) / 6;
          double quad2 = (F2[3 * g + 3] + 4 * F2[3 * g + 4] + F2[3 * g + 5]) *
                         (N2[2 * g + 3] - N2[2 * g + 2]) / 6;
          double es = fabs(quad1 + quad2 - quad) / 15 / (N1[2 * f - 1] - N1[2 * f - 2]);
          my_int1 = my_int1 + fabs(quad1) + fabs(quad2);
          if (es > c->ErrEps) {
            flag = 1;
            R2[g] = 1; R2[g + 1] = 1;
          } else {
            R2[g] = 0; R2[g + 1] = 0;
            at = at + quad1 + quad2;
          }
          g = g + 2;
        }
      }
    }
    memcpy(F1, F2, sizeof(double) * 3 * (size_t)g);
    memcpy(N1, N2, sizeof(double) * 2 * (size_t)g);
    memcpy(R1, R2, sizeof(int) * (size_t)g);
    Neps1 = g;
  }
  return at * sinchi;
}

static int pair_setup(const orc_geom* G, seta_ctx* c, int mode, int32_t i1, int32_t i2,
                      int32_t i3, double* dist2_out, double* dphi_out) {
  int32_t pci = G->nodes_pcell[i1 - 1];
  double uu1 = G->cells_lu[pci - 1], uu2 = G->cells_ru[pci - 1];
  double vv1 = G->cells_lv[pci - 1], vv2 = G->cells_rv[pci - 1];
  double ww1 = G->cells_lw[pci - 1], ww2 = G->cells_rw[pci - 1];
  double dphi = sqrt((uu2 - uu1) * (uu2 - uu1) + (vv2 - vv1) * (vv2 - vv1) + (ww2 - ww1) * (ww2 - ww1));
  double m1 = uu2 - uu1, m2 = vv2 - vv1, m3 = ww2 - ww1;
  double mx = m1 > m2 ? m1 : m2;
  mx = mx > m3 ? mx : m3;
  c->ires_e = mx / 32; /* ires_t = ires_e (lines 476-477) */
  c->G = G; c->mode = mode; c->i1 = i1;
  c->xiu = G->nodes_u[i2 - 1]; c->xiv = G->nodes_v[i2 - 1]; c->xiw = G->nodes_w[i2 - 1];
  c->xi1u = G->nodes_u[i3 - 1]; c->xi1v = G->nodes_v[i3 - 1]; c->xi1w = G->nodes_w[i3 - 1];
  c->varphi_xi = orc_basis(G, mode, c->xiu, c->xiv, c->xiw, i1);
  c->varphi_xi1 = orc_basis(G, mode, c->xi1u, c->xi1v, c->xi1w, i1);
  double du = c->xiu - c->xi1u, dv = c->xiv - c->xi1v, dw = c->xiw - c->xi1w;
  c->dsph = sqrt(du * du + dv * dv + dw * dw);
  double s1 = c->xiu + c->xi1u - uu1 - uu2, s2 = c->xiv + c->xi1v - vv1 - vv2,
         s3 = c->xiw + c->xi1w - ww1 - ww2;
  double dist2 = sqrt(s1 * s1 + s2 * s2 + s3 * s3);
  c->ugu = (c->xiu - c->xi1u) / c->dsph;
  c->ugv = (c->xiv - c->xi1v) / c->dsph;
  c->ugw = (c->xiw - c->xi1w) / c->dsph;
  double ugu = c->ugu, ugv = c->ugv, ugw = c->ugw;
  if (fabs(ugw - ugv) + fabs(ugu - ugw) + fabs(ugv - ugu) < 1.0e-6) { /* lines 500-511 */
    c->ugux2 = ugv; c->ugvx2 = -ugu; c->ugwx2 = 0;
    c->ugux3 = ugu * ugw; c->ugvx3 = ugv * ugw; c->ugwx3 = -(ugu * ugu) - (ugv * ugv);
    c->g1 = sqrt(c->ugux2 * c->ugux2 + c->ugvx2 * c->ugvx2);
    c->g2 = sqrt(c->ugux3 * c->ugux3 + c->ugvx3 * c->ugvx3 + c->ugwx3 * c->ugwx3);
  } else { /* lines 512-523 */
    c->ugux2 = (ugw - ugv); c->ugvx2 = (ugu - ugw); c->ugwx2 = (ugv - ugu);
    c->ugux3 = ((ugv * ugv) - (ugv) * (ugu) - (ugw) * (ugu) + (ugw * ugw));
    c->ugvx3 = ((ugw * ugw) - (ugw) * (ugv) - (ugu) * (ugv) + (ugu * ugu));
    c->ugwx3 = ((ugu * ugu) - (ugu) * (ugw) - (ugv) * (ugw) + (ugv * ugv));
    c->g1 = sqrt(c->ugux2 * c->ugux2 + c->ugvx2 * c->ugvx2 + c->ugwx2 * c->ugwx2);
    c->g2 = sqrt(c->ugux3 * c->ugux3 + c->ugvx3 * c->ugvx3 + c->ugwx3 * c->ugwx3);
  }
  *dist2_out = dist2;
  *dphi_out = dphi;
  return 0;
}

int orc_prefilter(const orc_geom* G, int32_t i1, int32_t i2, int32_t i3, double Trad) {
  seta_ctx c;
  double dist2, dphi;
  pair_setup(G, &c, 0, i1, i2, i3, &dist2, &dphi);
  /* line 527 */
  if ((dist2 > c.dsph + dphi) || (dist2 + dphi < c.dsph) || (c.dsph > Trad)) return 0;
  return 1;
}

static int seta_raw_ctx(seta_ctx* c, double ErrChi, double* atemp) {
  /* integration in chi (dgvtools_mod.f90:536-637) */
  double *N1 = (double*)malloc(sizeof(double) * 2 * NDS), *N2 = (double*)malloc(sizeof(double) * 2 * NDS);
  double *F1 = (double*)malloc(sizeof(double) * 3 * NDS), *F2 = (double*)malloc(sizeof(double) * 3 * NDS);
  int *R1 = (int*)malloc(sizeof(int) * NDS), *R2 = (int*)malloc(sizeof(int) * NDS);
  int err = 0;
  double ires_t = c->ires_e;
  int Nchi1 = (int)floor(PI25DT * c->dsph / ires_t / 2.0) + 1;
  if (Nchi1 > NDS) Nchi1 = NDS;
  for (int d = 1; d <= Nchi1; d++) {
    N1[2 * (d - 1)] = (d - 1) * PI25DT / (double)Nchi1;
    N1[2 * d - 1] = d * PI25DT / (double)Nchi1;
  }
  for (int d = 1; d <= Nchi1 && !err; d++) {
    if (d > 1 && N1[2 * d - 2] == N1[2 * d - 3]) F1[3 * d - 3] = F1[3 * d - 4];
    else F1[3 * d - 3] = a_int_epsilon(c, N1[2 * d - 2], &err);
    F1[3 * d - 1] = a_int_epsilon(c, N1[2 * d - 1], &err);
    F1[3 * d - 2] = a_int_epsilon(c, (N1[2 * d - 2] + N1[2 * d - 1]) / 2.0, &err);
  }
  for (int d = 1; d <= Nchi1; d++) R1[d - 1] = 1;
  double at = 0;
  int flag = 1;
  double my_int1 = 100 * ErrChi;
  while (flag && !err) {
    flag = 0;
    int e = 0;
    if (my_int1 > ErrChi) {
      my_int1 = 0;
      for (int d = 1; d <= Nchi1 && !err; d++) {
        if ((fabs(F1[3 * d - 3]) + fabs(F1[3 * d - 2]) + fabs(F1[3 * d - 1]) > 1.0e-15) &&
            R1[d - 1] == 1) {
          if (e > NDS - 2) { err = 1; break; }
          N2[2 * e] = N1[2 * d - 2];
          F2[3 * e] = F1[3 * d - 3];
          N2[2 * e + 3] = N1[2 * d - 1];
          F2[3 * e + 5] = F1[3 * d - 1];
          N2[2 * e + 1] = (N1[2 * d - 2] + N1[2 * d - 1]) / 2;
          N2[2 * e + 2] = N2[2 * e + 1];
          F2[3 * e + 2] = F1[3 * d - 2];
          F2[3 * e + 3] = F2[3 * e + 2];
          F2[3 * e + 1] = a_int_epsilon(c, (N2[2 * e + 1] + N2[2 * e]) / 2.0, &err);
          F2[3 * e + 4] = a_int_epsilon(c, (N2[2 * e + 3] + N2[2 * e + 2]) / 2.0, &err);
          double quad = (F1[3 * d - 3] + 4 * F1[3 * d - 2] + F1[3 * d - 1]) *
                        (N1[2 * d - 1] - N1[2 * d - 2]) / 6;
          double quad1 = (F2[3 * e] + 4 * F2[3 * e + 1] + F2[3 * e + 2]) * (N2[2 * e + 1] - N2[2 * e]) / 6;
          double quad2 = (F2[3 * e + 3] + 4 * F2[3 * e + 4] + F2[3 * e + 5]) *
                         (N2[2 * e + 3] - N2[2 * e + 2]) / 6;
          double es = fabs(quad1 + quad2 - quad) / 15 / (N1[2 * d - 1] - N1[2 * d - 2]);
          my_int1 = my_int1 + fabs(quad1) + fabs(quad2);
          if (es > ErrChi) {
            flag = 1;
            R2[e] = 1; R2[e + 1] = 1;
          } else {
            R2[e] = 0; R2[e + 1] = 0;
            at = at + quad1 + quad2;
          }
          e = e + 2;
        }
      }
    }
    memcpy(F1, F2, sizeof(double) * 3 * (size_t)e);
    memcpy(N1, N2, sizeof(double) * 2 * (size_t)e);
    memcpy(R1, R2, sizeof(int) * (size_t)e);
    Nchi1 = e;
  }
  free(N1); free(N2); free(F1); free(F2); free(R1); free(R2);
  *atemp = at;
  return err ? -1 : 0;
}

static void ctx_alloc(seta_ctx* c) {
  c->N1 = (double*)malloc(sizeof(double) * 2 * NDS); c->N2 = (double*)malloc(sizeof(double) * 2 * NDS);
  c->F1 = (double*)malloc(sizeof(double) * 3 * NDS); c->F2 = (double*)malloc(sizeof(double) * 3 * NDS);
  c->R1 = (int*)malloc(sizeof(int) * NDS); c->R2 = (int*)malloc(sizeof(int) * NDS);
}
static void ctx_free(seta_ctx* c) {
  free(c->N1); free(c->N2); free(c->F1); free(c->F2); free(c->R1); free(c->R2);
}

int orc_seta_raw(const orc_geom* G, int mode, int32_t i1, int32_t i2, int32_t i3,
                 double ErrChi, double ErrEps, double* atemp) {
  seta_ctx c;
  double dist2, dphi;
  pair_setup(G, &c, mode, i1, i2, i3, &dist2, &dphi);
  c.ErrEps = ErrEps;
  ctx_alloc(&c);
  int rc = seta_raw_ctx(&c, ErrChi, atemp);
  ctx_free(&c);
  return rc;
}

int64_t orc_seta(const orc_geom* G, int mode, int32_t i1, double Trad, double ErrChi,
                 double ErrEps, double min_sens, int64_t cap, double* A, int32_t* A_xi,
                 int32_t* A_xi1, int nthreads) {
  int64_t ni = G->nnodes;
  /* per-i2 record lists, concatenated in i2 order = the serial order */
  int64_t* cnt = (int64_t*)calloc((size_t)ni + 1, sizeof(int64_t));
  double** la = (double**)calloc((size_t)ni, sizeof(double*));
  int32_t** lx = (int32_t**)calloc((size_t)ni, sizeof(int32_t*));
  int nt = set_threads(nthreads), err = 0;
#pragma omp parallel num_threads(nt)
  {
    seta_ctx c;
    ctx_alloc(&c);
#pragma omp for schedule(dynamic, 1)
    for (int64_t i2 = 1; i2 <= ni; i2++) {
      int64_t n = 0, capl = 16;
      double* a = (double*)malloc(sizeof(double) * capl);
      int32_t* x = (int32_t*)malloc(sizeof(int32_t) * capl);
      for (int64_t i3 = i2 + 1; i3 <= ni; i3++) {
        double dist2, dphi;
        pair_setup(G, &c, mode, i1, (int32_t)i2, (int32_t)i3, &dist2, &dphi);
        if ((dist2 > c.dsph + dphi) || (dist2 + dphi < c.dsph) || (c.dsph > Trad)) continue;
        c.ErrEps = ErrEps;
        double at;
        if (seta_raw_ctx(&c, ErrChi, &at)) { err = 1; break; }
        if (fabs(at) > min_sens) { /* lines 641-659 */
          at = at * (G->nodes_gwts[i3 - 1]);
          at = at * (G->nodes_gwts[i2 - 1]);
          double ds = c.dsph / 8.0;
          if (n == capl) {
            capl *= 2;
            a = (double*)realloc(a, sizeof(double) * capl);
            x = (int32_t*)realloc(x, sizeof(int32_t) * capl);
          }
          a[n] = at * ds;
          x[n] = (int32_t)i3;
          n++;
        }
      }
      cnt[i2] = n; la[i2 - 1] = a; lx[i2 - 1] = x;
    }
    ctx_free(&c);
  }
  int64_t tot = 0;
  for (int64_t i2 = 1; i2 <= ni; i2++) {
    for (int64_t k = 0; k < cnt[i2] && !err; k++) {
      if (tot >= cap) { err = 2; break; }
      A[tot] = la[i2 - 1][k]; A_xi[tot] = (int32_t)i2; A_xi1[tot] = lx[i2 - 1][k];
      tot++;
    }
    free(la[i2 - 1]); free(lx[i2 - 1]);
  }
  free(cnt); free(la); free(lx);
  if (err) return -(int64_t)err;
  return tot;
}

/* ---- A-array truncation ------------------------------------------------------
 * Both routines walk the basis functions in order and copy the surviving
 * records forward (the reference's compaction through scratch arrays,
 * dgvtools_mod.f90:1339-1359 and 1405-1426). */
int64_t orc_trunc_threshold(int64_t ncap, int32_t* capphi, double* A, int32_t* A_xi,
                            int32_t* A_xi1, double trhld) {
  int64_t src = 0, dst = 0;
  for (int64_t i = 0; i < ncap; ++i) {
    int32_t kept = 0;
    for (int32_t j = 0; j < capphi[i]; ++j, ++src) {
      if (fabs(A[src]) > trhld) {
        A[dst] = A[src];
        A_xi[dst] = A_xi[src];
        A_xi1[dst] = A_xi1[src];
        ++dst;
        ++kept;
      }
    }
    capphi[i] = kept;
  }
  return dst;
}

int64_t orc_trunc_radius(const orc_geom* G, int32_t* capphi, double* A, int32_t* A_xi,
                         int32_t* A_xi1, double rad) {
  const double radsq = rad * rad;
  int64_t src = 0, dst = 0;
  for (int64_t i = 0; i < G->nnodes; ++i) {
    int32_t kept = 0;
    for (int32_t j = 0; j < capphi[i]; ++j, ++src) {
      const int64_t a = A_xi[src] - 1, b = A_xi1[src] - 1;
      const double du = G->nodes_u[a] - G->nodes_u[b];
      const double dv = G->nodes_v[a] - G->nodes_v[b];
      const double dw = G->nodes_w[a] - G->nodes_w[b];
      const double dsq = du * du + dv * dv + dw * dw;
      if (dsq <= radsq) {
        A[dst] = A[src];
        A_xi[dst] = A_xi[src];
        A_xi1[dst] = A_xi1[src];
        ++dst;
        ++kept;
      }
    }
    capphi[i] = kept;
  }
  return dst;
}
