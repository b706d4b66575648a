/*
 * dgv_oracle.h -- CPU restatement of the reference's direct DGV collision path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle: only tests/, the
 * smoke() entry of __graft_entry__.py and the cpu_baseline leg of bench.py may
 * load it.  The product (dgv0d3v_direct_collision_amd) never links or calls it.
 *
 * Every function restates one routine of alexmalekseenko/DGV0D3V_direct_collision
 * (Fortran, 1-based arrays).  Index conventions of the reference are kept on
 * purpose: node ids inside A_xi / A_xi1 and cell ids are 1-based, exactly as the
 * commvar module stores them; C arrays themselves are 0-based (element k of a
 * Fortran array lives at [k-1]).
 */
#ifndef DGV_ORACLE_H
#define DGV_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* One uniform velocity grid (SetDGVblzmmesh / OldCreateUmesh, miscset.f90:210-525). */
typedef struct {
  int32_t Mu, Mv, Mw;        /* cells per dimension                          */
  int32_t su, sv, sw;        /* Gauss nodes per cell per dimension           */
  double u_L, u_R, v_L, v_R, w_L, w_R;
} orc_grid;

/* Geometry arrays of commvar (commvar.f90:57-86), filled by orc_geom_build. */
typedef struct {
  orc_grid g;
  int64_t nnodes, ncells;
  int32_t ccell;             /* canonical cell, 1-based (FindCellContainsPoint_DGV(0,0,0)) */
  /* cells (length ncells) */
  double *cells_lu, *cells_ru, *cells_lv, *cells_rv, *cells_lw, *cells_rw;
  int32_t *cells_ugi, *cells_vgi, *cells_wgi;           /* 1-based grid address */
  /* nodes (length nnodes) */
  double *nodes_u, *nodes_v, *nodes_w, *nodes_gwts;
  int32_t *nodes_pcell;                                 /* 1-based cell id      */
  int32_t *nodes_ui, *nodes_vi, *nodes_wi;              /* 1-based local index  */
  int32_t *nodes_dui, *nodes_dvi, *nodes_dwi;           /* shift to canonical   */
  int32_t *nodes_phican;                                /* 1-based canonical node */
  /* Gauss nodes per dimension (g_nds_all(:,s), miscset.f90:29-190) */
  double gn_u[16], gn_v[16], gn_w[16], gw_u[16], gw_v[16], gw_w[16];
} orc_geom;

int  orc_gauleg(int n, double x1, double x2, double* x, double* w);
int  orc_geom_build(const orc_grid* g, orc_geom** out);
void orc_geom_free(orc_geom* G);
/* nodes_Ashift (miscset.f90:1307), int64 because the running sum passes 2^31 at 33^3 */
int  orc_ashift(const orc_geom* G, const int32_t* capphi, int64_t* ashift);

/* a1: EvalCollisionPeriodicAPlus_DGV (collision_mod.f90:143-229), nodes
 * [node_begin, node_end) (0-based), OpenMP over nodes when nthreads > 1.       */
int orc_eval_aplus(const orc_geom* G, const int32_t* capphi, const double* A,
                   const int32_t* A_xi, const int32_t* A_xi1, const double* f,
                   double* fc, int64_t node_begin, int64_t node_end, int nthreads);
/* a1 on an explicit list of 0-based nodes (bounded samples of large grids) */
int orc_eval_aplus_list(const orc_geom* G, const int32_t* capphi, const double* A,
                        const int32_t* A_xi, const int32_t* A_xi1, const double* f,
                        double* fc, const int64_t* nodes, int64_t nsel, int nthreads);
/* a2: EvalCollisionPeriodicA_DGV (collision_mod.f90:29-129): compress, then sum */
int orc_eval_a(const orc_geom* G, const int32_t* capphi, const double* A,
               const int32_t* A_xi, const int32_t* A_xi1, const double* f,
               double* fc, int64_t node_begin, int64_t node_end, int nthreads);
/* a1 with Neumaier-compensated accumulation: estimates a1's own rounding.      */
/* a1 at listed nodes in double-double (exact products, two-sum accumulation) */
int orc_eval_aplus_dd_list(const orc_geom* G, const int32_t* capphi, const double* A,
                           const int32_t* A_xi, const int32_t* A_xi1, const double* f, double* fc,
                           double* abs_sum, const int64_t* nodes, int64_t nsel, int nthreads);
int orc_eval_aplus_comp(const orc_geom* G, const int32_t* capphi, const double* A,
                        const int32_t* A_xi, const int32_t* A_xi1, const double* f,
                        double* fc, double* abs_sum, int64_t node_begin,
                        int64_t node_end, int nthreads);
/* a8: IphiDGV (dgvtools_mod.f90:847-877), unshifted evaluator over a full A.
 * coef = (mol_diam/L_inf)^2 * N_inf.                                            */
int orc_iphi(int64_t nnodes, const int32_t* capphi, const double* A,
             const int32_t* A_xi, const int32_t* A_xi1, const double* gwts,
             const double* f, double coef, double* out);

/* a6: MassCheckRec (dgvtools_mod.f90:1437-1458): out = {n, ubar, vbar, wbar, T} */
void orc_masscheck(const orc_geom* G, const double* f, double out[5]);
void orc_masscheck_high(const orc_geom* G, const double* f, double out[20]);
/* maxwelveldist (distributions_mod.f90:83-94) */
double orc_maxwell(double T, double u0, double v0, double w0, double n,
                   double u, double v, double w);

/* Collision evaluator callback used by the steppers (run_mode 0). */
typedef int (*orc_q_fn)(void* user, const double* f, double* fc);
/* a6: SpatialOperSH_DGV run_mode 0 (spat_oper_mod.f90:28-68) */
int orc_spatial_oper(orc_q_fn q, void* user, int64_t n, const double* f,
                     double* fcol, double dt, double coef);
/* a7: TimeIntegratorORK_SH_DGV (time_integr_mod.f90:26-152), rkmts = 1..5.
 * frhs1 receives the first stage divided by dt (line 151).                     */
int orc_rk_step(orc_q_fn q, void* user, int64_t n, double* f, double dt,
                double coef, int rkmts, double* frhs1);
/* TimeIntegratorMTS_SH_DGV (time_integr_mod.f90:314-364): hist holds
 * f1..f4 then frhs1..frhs5 (9 arrays of length n), advanced in place.          */
int orc_ab_step(orc_q_fn q, void* user, int64_t n, double* f, double dt,
                double coef, int rkmts, double* hist);

/* ---- A-array generation: SetA_DGV (dgvtools_mod.f90:377-681) ---------------- */
/* basis mode 0: Lagrange basis restored (the "(rem to restore)" lines,
 * dgvtools_mod.f90:254-287); modes 1..5: the moment toggles y = 1, u, v, w,
 * u^2+v^2+w^2 (dgvtools_mod.f90:267-274), no cell restriction.                 */
double orc_basis(const orc_geom* G, int mode, double u, double v, double w, int32_t i1);
/* Prefilter of one pair (dgvtools_mod.f90:493-530): 1 = keep, 0 = skip.       */
int orc_prefilter(const orc_geom* G, int32_t i1, int32_t i2, int32_t i3, double Trad);
/* The raw Atemp of one pair (the chi/eps adaptive Simpson, lines 531-637)
 * before the min_sens test and the weights; status -1 on node overflow.        */
int orc_seta_raw(const orc_geom* G, int mode, int32_t i1, int32_t i2, int32_t i3,
                 double ErrChi, double ErrEps, double* atemp);
/* Full SetA_DGV for one basis function i1 over all pairs i2<i3 (tiny grids).
 * Records are written in the reference's serial order; returns count or <0.    */
int64_t orc_seta(const orc_geom* G, int mode, int32_t i1, double Trad,
                 double ErrChi, double ErrEps, double min_sens, int64_t cap,
                 double* A, int32_t* A_xi, int32_t* A_xi1, int nthreads);

/* EvalCollisionPeriodicMixedTermsA_DGV (collision_mod.f90:244-332)           */
int orc_eval_mixed(const orc_geom* G, const int32_t* capphi, const double* A,
                   const int32_t* A_xi, const int32_t* A_xi1, const double* f, const double* fm,
                   double* fc, int nthreads);
/* PrepareFMA_DGV (collision_mod.f90:377-464): dense fmA, column-major N x N   */
int orc_prepare_fma(const orc_geom* G, const int32_t* capphi, const double* A,
                    const int32_t* A_xi, const int32_t* A_xi1, const double* fm, double* fmA);
/* EvalCollisionLinear (collision_mod.f90:346-359): fc = 2 f . fmA            */
void orc_eval_linear(int64_t N, const double* fmA, const double* f, double* fc);
/* TruncAarrsTrhls_DGV (dgvtools_mod.f90:1313-1362): keep records with
 * |A| > trhld, in order; capphi updated in place.  Returns the new count.     */
int64_t orc_trunc_threshold(int64_t ncap, int32_t* capphi, double* A, int32_t* A_xi,
                            int32_t* A_xi1, double trhld);
/* TruncAarrsRadius_DGV (dgvtools_mod.f90:1375-1429): keep records whose nodes
 * (1-based ids) are at squared distance <= rad**2.  Returns the new count.    */
int64_t orc_trunc_radius(const orc_geom* G, int32_t* capphi, double* A, int32_t* A_xi,
                         int32_t* A_xi1, double rad);

#ifdef __cplusplus
}
#endif
#endif
