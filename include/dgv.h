/*
 * dgv.h -- C ABI of the MI355X-native direct DGV collision evaluator.
 *
 * Drop-in replacement for the collision_mod / spat_oper_mod / time_integr_mod
 * path of alexmalekseenko/DGV0D3V_direct_collision.  Every entry point returns
 * an int status (DGV_OK = 0, negative = error; text via dgv_last_error) where the
 * reference prints a message and STOPs.  Plain pointers and sizes only.
 *
 * Conventions
 *  - Node order is the reference's (cell u->v->w outer, local u->v->w inner,
 *    miscset.f90:915-927,1150-1176); arrays handed in or out use that order.
 *  - A-arrays handed in use the reference layout (commvar.f90:157-167): per
 *    node k the A_capphi(k) records of basis function k are contiguous,
 *    records hold A, A_xi, A_xi1 with node ids in the given index_base.
 *  - "_dev" entry points take device pointers and enqueue on the context's
 *    stream without synchronising; the others take host pointers and block.
 *  - A context is not reentrant: one host thread per context.
 */
#ifndef DGV_H
#define DGV_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
  DGV_OK = 0,
  DGV_ERR_ARG = -1,       /* bad argument / size mismatch                    */
  DGV_ERR_GRID = -2,      /* grid unsupported (origin outside, s > 8, ...)   */
  DGV_ERR_INDEX = -3,     /* A_xi / A_xi1 out of range for index_base        */
  DGV_ERR_NOMEM = -4,     /* device or host allocation failed                */
  DGV_ERR_HIP = -5,       /* HIP runtime error                               */
  DGV_ERR_STATE = -6,     /* call out of order (no A-array set, ...)         */
  DGV_ERR_IO = -7,        /* file missing, truncated or corrupt              */
  DGV_ERR_UNSUPPORTED = -8
};

enum { DGV_FP64 = 0, DGV_FP32_COMP = 1, DGV_FP32_F64ACC = 2, DGV_INT8_SPLIT = 3 };   /* precision */
enum { DGV_KERNEL_AUTO = 0, DGV_KERNEL_PAIRBLOCK = 1, DGV_KERNEL_RECORD = 2, DGV_KERNEL_MFMA = 3 };

typedef struct dgv_ctx dgv_ctx;

/* One uniform velocity grid: SetDGVblzmmesh/Set3DCellsR_DGVblzm inputs
 * (readwrite.f90:139-347: "left/right endpoint in u", "number of cells in u",
 * "degree of local Lagrange basis in u" -> su = nodes per cell). */
typedef struct {
  int32_t Mu, Mv, Mw;
  int32_t su, sv, sw;
  double u_L, u_R, v_L, v_R, w_L, w_R;
} dgv_grid_desc;

/* Parameters of the A-array generator (SetA_DGV, dgvtools_mod.f90:377-681;
 * keys of readwrite.f90:652-696). */
typedef struct {
  double Trad;        /* cutoff radius ("cutoff radius of A-array")          */
  double min_sens;    /* drop |A_raw| <= min_sens ("cutoff values for A")     */
  int32_t nquad;      /* Gauss points per smooth t-interval (0 = default 24)  */
  int32_t moment_mode;/* 0: Lagrange basis; 1..5: psi = 1,u,v,w,|u|^2 test    */
} dgv_agen_params;

/* ---- context, geometry ---------------------------------------------------- */
/* replaces SetDGblzmGnodes + SetDGVblzmmesh + Set3DCellsR_DGVblzm + SetNodesDGV
 * + SetCellsUGINdsSftArrs(FindCellContainsPoint_DGV(0,0,0))
 * (DGVblzmPL.f90:75-96, miscmpiset.f90:553) */
int dgv_ctx_create(const dgv_grid_desc* grid, int device, dgv_ctx** out);
void dgv_ctx_destroy(dgv_ctx* ctx);
const char* dgv_last_error(const dgv_ctx* ctx);   /* ctx may be NULL        */
int dgv_ctx_num_nodes(const dgv_ctx* ctx, int64_t* nnodes);
/* the same geometry without a context or GPU: node coordinates and weights
 * (arrays of (Mu*Mv*Mw)*(su*sv*sw) doubles, any may be NULL), 1-based
 * canonical cell */
int dgv_grid_geometry(const dgv_grid_desc* grid, double* u, double* v, double* w, double* gwts,
                      int32_t* ccell, int64_t* nnodes);
/* 1-based canonical cell, as FindCellContainsPoint_DGV (dgvtools_mod.f90:1175) */
int dgv_ctx_canonical_cell(const dgv_ctx* ctx, int32_t* ccell);
/* nodes_u/_v/_w/_gwts (commvar.f90:75-82, SetNodesDGV miscset.f90:1160-1167);
 * any pointer may be NULL */
int dgv_ctx_get_nodes(const dgv_ctx* ctx, double* u, double* v, double* w, double* gwts);
/* nodes_phican / nodes_dui,dvi,dwi (miscset.f90:1298-1305), 1-based phican */
int dgv_ctx_get_shifts(const dgv_ctx* ctx, int32_t* phican, int32_t* dui, int32_t* dvi,
                       int32_t* dwi);

/* ---- A-array ----------------------------------------------------------------- */
/* replaces ReadAarraysChnksOfstNrecDGV's in-memory result (readwrite.f90:1261-1449):
 * the caller's arrays are copied; records of every node k may be present, only
 * the canonical cell's are used by the evaluator (collision_mod.f90:187-188). */
int dgv_ctx_set_aarray(dgv_ctx* ctx, const int32_t* capphi, int64_t ncapphi,
                       const double* A, const int32_t* A_xi, const int32_t* A_xi1,
                       int64_t nrec, int index_base);
/* replaces SetA_DGV for all canonical basis functions (dgvtools_mod.f90:377-681),
 * run on the GPU; the result becomes the context's A-array. */
int dgv_ctx_generate_aarray(dgv_ctx* ctx, const dgv_agen_params* params);
/* Host-side evaluation of one pair (1-based node ids i2 < i3, as SetA_DGV's
 * loop variables, dgvtools_mod.f90:484-488) with the same arithmetic as the GPU
 * generator: raw[l] = Atemp_l, A[l] = record value (0 below min_sens); returns
 * 1 when the pair fails the prefilter.  Needs no GPU. */
int dgv_agen_pair(const dgv_grid_desc* grid, int64_t i2, int64_t i3, const dgv_agen_params* params,
                  double* raw, double* A);
/* record count of the current A-array (reference layout), per canonical basis
 * function when per_basis != NULL (length su*sv*sw) */
int dgv_ctx_aarray_size(const dgv_ctx* ctx, int64_t* nrec, int64_t* per_basis);
/* work of one Q evaluation with the current A-array: out[0] = records R
 * (reference layout), out[1] = node pairs, out[2] = in-grid record visits V
 * (the records that pass the bounds tests of collision_mod.f90:196-211, summed
 * over all nodes), out[3] = in-grid pair visits */
int dgv_ctx_workload(const dgv_ctx* ctx, int64_t out[4]);
/* export the records of one canonical basis function (0-based local index l)
 * in the reference's serial order (xi ascending, then xi1); index_base 0/1 */
int dgv_ctx_get_aarray_basis(const dgv_ctx* ctx, int32_t l, double* A, int32_t* A_xi,
                             int32_t* A_xi1, int64_t cap, int64_t* nrec, int index_base);
/* the records of the canonical basis functions ls[0..nl-1], each list in the
 * reference's serial order, back to back in A / A_xi / A_xi1 (counts[k] records
 * for ls[k]); one pass over the device layout (no host cache).  With A, A_xi and
 * A_xi1 all NULL only the counts are returned.  Same readers as above
 * (ReadAarraysDGV's A, A_xi, A_xi1 arrays, readwrite.f90:1329-1333). */
int dgv_ctx_get_aarray_bases(const dgv_ctx* ctx, const int32_t* ls, int32_t nl, double* A,
                             int32_t* A_xi, int32_t* A_xi1, int64_t cap, int64_t* counts,
                             int index_base);
/* ---- A-array files (readwrite.f90:1066-1484, 2224-2365) ---------------------
 * Fortran unformatted sequential: record 1 = (mm, nn) int32; record 2 =
 * A_capphi(mm) int32, A(nn) float64, A_xi(nn), A_xi1(nn), A_phi(nn) int32.
 * Records above 2^31 - 9 bytes (every grid from 21^3 on) are framed either as
 * Intel Fortran / gfortran subrecords or as LLVM Flang's single record with the
 * length modulo 2^32; readers recognise both, writers take the framing. */
enum { DGV_FRAMING_SUBRECORDS = 0, DGV_FRAMING_UNSIGNED32 = 1 };
/* WriteAarraysDGV / ReadAarraysDGV of the context's A-array (readwrite.f90:
 * 1459-1484, 1066-1118); the writer uses DGV_FRAMING_SUBRECORDS. */
int dgv_aarray_write_file(const dgv_ctx* ctx, const char* path);
int dgv_ctx_read_aarray_file(dgv_ctx* ctx, const char* path);
/* the same file on host arrays (no GPU): sizes of record 1, record 2 read into
 * caller arrays of exactly those sizes, and the writer */
int dgv_aarray_file_header(const char* path, int64_t* mm, int64_t* nn);
int dgv_aarray_file_read(const char* path, int32_t* capphi, int64_t mm, double* A, int32_t* A_xi,
                         int32_t* A_xi1, int32_t* A_phi, int64_t nn);
int dgv_aarray_file_write(const char* path, const int32_t* capphi, int64_t mm, const double* A,
                          const int32_t* A_xi, const int32_t* A_xi1, const int32_t* A_phi, int64_t nn,
                          int32_t framing);
/* test hook: largest subrecord payload of DGV_FRAMING_SUBRECORDS (default and
 * maximum 2^31 - 9 bytes), so that the framing is exercised on small files */
int dgv_set_fortran_subrecord_limit(int64_t bytes);
/* chunk file names: dgv_aarray_chunk_name below */
/* ScanAarrsChnks4Acapphi (readwrite.f90:2292-2365): records in each of chunks
 * 0 .. nchnks-1 (sum of the chunk's A_capphi) */
int dgv_aarray_chunks_scan(const dgv_grid_desc* grid, const char* dir, const char* base, int32_t k_c,
                           int32_t N_x, int32_t nchnks, int64_t* chunks_Act);
/* ReadAarraysChnksOfstNrecDGV(fstchnk, ofst, numrec, nchnks) (readwrite.f90:
 * 1261-1449): numrec records from chunk fstchnk (1-based) on, the first ofst
 * records of that chunk skipped; capphi[mm] gets the reference's per-node
 * counts of the records read; *nread = records read.  ReadAarraysChnksDGV
 * (readwrite.f90:1130-1243) is fstchnk = 1, ofst = 0, numrec = all.  Where the
 * reference prints and stops, the call returns DGV_ERR_INDEX or DGV_ERR_IO. */
int dgv_aarray_chunks_read(const dgv_grid_desc* grid, const char* dir, const char* base, int32_t k_c,
                           int32_t N_x, int32_t fstchnk, int64_t ofst, int64_t numrec, int32_t nchnks,
                           int32_t* capphi, int64_t mm_cap, double* A, int32_t* A_xi, int32_t* A_xi1,
                           int32_t* A_phi, int64_t nrec_cap, int64_t* nread);
/* the context's A-array as a chunk set of at most max_records records per chunk
 * (the reference's chunk files come from its MPI generator; no single routine) */
int dgv_ctx_write_aarray_chunks(const dgv_ctx* ctx, const char* dir, const char* base, int32_t k_c,
                                int32_t N_x, int64_t max_records, int32_t framing, int32_t* nchunks);
/* ReadAarraysChnksDGV(nchnks) into the context */
int dgv_ctx_read_aarray_chunks(dgv_ctx* ctx, const char* dir, const char* base, int32_t k_c, int32_t N_x,
                               int32_t nchnks);
/* TruncAarrsTrhls_DGV(trhld) (dgvtools_mod.f90:1313-1362): keeps the records with
 * |A| > trhld.  TruncAarrsRadius_DGV(rad) (dgvtools_mod.f90:1375-1429): keeps
 * the records whose nodes xi, xi1 lie at most rad apart.  Both work on the
 * device copy in place and keep the records' order; batches and surrogates
 * built on the context before the call must be created again. */
int dgv_ctx_trunc_aarray_threshold(dgv_ctx* ctx, double trhld);
/* Diagnostic, no counterpart in the reference: replaces every coefficient by its
 * absolute value.  For f >= 0 the evaluator then returns 2/w sum |A| f f, the
 * scale of the terms whose sum is Q (the conditioning of the sum). */
int dgv_ctx_aarray_abs(dgv_ctx* ctx);
int dgv_ctx_trunc_aarray_radius(dgv_ctx* ctx, double rad);
/* reference chunk-file base name (MakeBaseNameAoperChnksDGV, readwrite.f90:2224-2281) */
int dgv_aarray_chunk_name(const dgv_grid_desc* grid, const char* dir, const char* base,
                          int32_t chunk, int32_t k_c, int32_t N_x, char* out, size_t cap);

/* ---- collision operator --------------------------------------------------------- */
/* EvalCollisionPeriodicAPlus_DGV(f,fc) (collision_mod.f90:143-229):
 * fc(i) = (2/w_i) sum_r A_r f(xi_r + d_i) f(xi1_r + d_i) over in-grid records. */
int dgv_eval_q(dgv_ctx* ctx, const double* f, double* fc, int64_t n);
int dgv_eval_q_dev(dgv_ctx* ctx, const double* d_f, double* d_fc, int64_t n);
/* No counterpart in the reference: dgv_eval_q_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * It allocates nothing: the partial sums (and the projection's moments) go to
 * the library's static device array of 128 MiB per GPU, held from the library's
 * load to the end of the process.  A plan whose partial sums exceed it returns
 * DGV_ERR_UNSUPPORTED before any launch (raise the tile).  Calls that overlap on
 * one GPU (two streams or threads) share the array and race. */
int dgv_eval_q_dev_noalloc(dgv_ctx* ctx, const double* d_f, double* d_fc, int64_t n);
/* EvalCollisionPeriodicA_DGV(f,fc) (collision_mod.f90:29-129): same quantity */
int dgv_eval_q_a(dgv_ctx* ctx, const double* f, double* fc, int64_t n);
/* EvalCollisionPeriodicMixedTermsA_DGV(f,fm,fc) (collision_mod.f90:244-332):
 * fc_i = 2/w_i sum A (fm(xi) f(xi1) + f(xi) fm(xi1)) over the shifted records,
 * the cross term of the decomposition f = fm + df; equals twice Q(f) for fm = f */
int dgv_eval_mixed(dgv_ctx* ctx, const double* f, const double* fm, double* fc, int64_t n);
int dgv_eval_mixed_dev(dgv_ctx* ctx, const double* d_f, const double* d_fm, double* d_fc, int64_t n);
/* selects the kernel used by dgv_eval_q* (DGV_KERNEL_*) */
int dgv_ctx_set_kernel(dgv_ctx* ctx, int kernel);
/* tunables of the pair-block kernel (pairs per work item); 0 = keep */
int dgv_ctx_set_tile(dgv_ctx* ctx, int32_t pairs_per_item);
/* on != 0: every Q of the context (evaluations, spatial operator, RK/AB stages,
 * graph replays) is projected in FP64 onto the complement of the collision
 * invariants -- the conservation-moment correction.  Off by default: the
 * reference applies none. */
int dgv_ctx_set_projection(dgv_ctx* ctx, int32_t on);
int dgv_ctx_autotune(dgv_ctx* ctx, int32_t* best_pairs_per_item, double* best_ms);

/* ---- decomposition and linearised regimes ---------------------------------------- */
/* PrepareFMA_DGV (collision_mod.f90:377-464): builds the dense linearised
 * operator fmA(r, i) = (1/w_i) sum over node i's shifted records of
 * fm(xi1) A at r = xi and fm(xi) A at r = xi1, N x N doubles in device memory
 * (3.1 GB at 27^3).  fm becomes the context's Maxwellian of the regime check. */
int dgv_prepare_fma(dgv_ctx* ctx, const double* fm, int64_t n);
int dgv_prepare_fma_dev(dgv_ctx* ctx, const double* d_fm, int64_t n);
/* EvalCollisionLinear(f,fc) (collision_mod.f90:346-359): fc = 2 MATMUL(f, fmA).
 * The _dev form applies it to nvec vectors (f at d_f + k*ldf, fc at d_fc + k*ldq). */
int dgv_eval_linear(dgv_ctx* ctx, const double* f, double* fc, int64_t n);
int dgv_eval_linear_dev(dgv_ctx* ctx, const double* d_f, int64_t ldf, double* d_fc, int64_t ldq,
                        int64_t n, int32_t nvec);
/* fmA in the reference's column-major layout: fmA(r, i) at fmA[r + n*i] (n*n doubles) */
int dgv_ctx_get_fma(dgv_ctx* ctx, double* fmA, int64_t n);
/* the Maxwellian fmA was last built at (commvar fm) */
int dgv_ctx_get_fm(dgv_ctx* ctx, double* fm, int64_t n);
/* run_mode, decomp_lev, linear_lev of commvar (parameters.dat).  decomp_lev > 0
 * turns on CheckSolutionMode_DGV before every spatial operator, as the
 * reference's SpatialOperSH_DGV calls it; decomp_lev <= 0 keeps run_mode 0 (the
 * reference's check then always lands in run_mode 0).  run_mode 1 or 2 needs a
 * prepared fmA.  With switching on, RK runs step eagerly (no graph replay).
 * decomp_lev > 0 with decomp_lev <= linear_lev is DGV_ERR_ARG, where the
 * reference prints and stops (spat_oper_mod.f90:331-334). */
int dgv_ctx_set_regime(dgv_ctx* ctx, int32_t run_mode, double decomp_lev, double linear_lev);
/* current run_mode, the last relative L1 deviation from the local Maxwellian
 * (the value KeepTrackL1_err_DGV records), the last |fm - f_M|_1 / n, and
 * counts = {regime checks, fmA builds} */
int dgv_ctx_get_regime(dgv_ctx* ctx, int32_t* run_mode, double* l1_err, double* l1_err_fm,
                       int64_t counts[2]);
/* IphiDGV(f) (dgvtools_mod.f90:847-877): for every node (one basis function per
 * node) 2/w_i sum over its records of A f(xi) f(xi1), times coef (the
 * reference's (mol_diam/L_inf)^2 N_inf) -- on the shifted A-array, Q(f) scaled.
 * IphiDGV_decomp(f) (dgvtools_mod.f90:896-927): the same with f split at its
 * local Maxwellian fM (MassCheckRec, maxwelveldist):
 * 2/w_i sum A (fM Df + Df fM + Df Df), Df = f - fM.  Both are the plain record
 * sums, as in the reference: the context's projection setting does not apply. */
int dgv_iphi(dgv_ctx* ctx, const double* f, double* iphi, int64_t n, double coef);
int dgv_iphi_decomp(dgv_ctx* ctx, const double* f, double* iphi, int64_t n, double coef);
/* CheckSolutionMode_DGV(f,Df) (spat_oper_mod.f90:195-272): moments, local
 * Maxwellian, L1 deviation, regime switch (rebuilding fmA as the reference
 * does), Df = f - fm (zeros in run_mode 0; df may be NULL). */
int dgv_check_solution_mode(dgv_ctx* ctx, const double* f, double* df, int64_t n);

/* SpatialOperSH_DGV(f,fcol,dt) (spat_oper_mod.f90:28-68): fcol = C(f) * coef * dt
 * with coef = (mol_diam/L_inf)^2 * N_inf and C = Q(f) (run_mode 0),
 * L(Df) + Q(Df) (run_mode 1) or L(Df) (run_mode 2), after the regime check when
 * it is enabled (dgv_ctx_set_regime). */
int dgv_spatial_oper(dgv_ctx* ctx, const double* f, double* fcol, int64_t n, double dt,
                     double coef);
int dgv_spatial_oper_dev(dgv_ctx* ctx, const double* d_f, double* d_fcol, int64_t n,
                         double dt, double coef);
/* TimeIntegratorORK_SH_DGV(f,dt), rkmts 1..5 (time_integr_mod.f90:26-152);
 * frhs1 (may be NULL) receives the first stage / dt (line 151). */
int dgv_rk_step(dgv_ctx* ctx, double* f, int64_t n, double dt, double coef, int32_t rkmts,
                double* frhs1);
int dgv_rk_step_dev(dgv_ctx* ctx, double* d_f, int64_t n, double dt, double coef,
                    int32_t rkmts, double* d_frhs1);
/* nsteps RK steps replayed from one captured hipGraph of a step */
int dgv_rk_run_dev(dgv_ctx* ctx, double* d_f, int64_t n, double dt, double coef,
                   int32_t rkmts, int64_t nsteps, int use_graph);
/* TimeIntegratorMTS_SH_DGV(f,dt), Adams-Bashforth 1..5 (time_integr_mod.f90:314-364);
 * hist = f1..f4, frhs1..frhs5 (9*n doubles), updated in place. */
int dgv_ab_step(dgv_ctx* ctx, double* f, int64_t n, double dt, double coef, int32_t rkmts,
                double* hist);

/* MassCheckRec (dgvtools_mod.f90:1437-1458): out = {n, ubar, vbar, wbar, T} */
int dgv_moments(dgv_ctx* ctx, const double* f, int64_t n, double out[5]);
/* MassCheckRecHighCMoments (dgvtools_mod.f90:1501-1549): out = {n, ubar, vbar,
 * wbar, T, T_u, T_v, T_w, mom3_u, mom3_v, mom3_w, mom4_*, mom5_*, mom6_*} with
 * mom_k_d = sum f w (u_d - ubar_d)^k / n and T_d = sum f w (u_d - ubar_d)^2 / n * 2/3 */
int dgv_moments_high(dgv_ctx* ctx, const double* f, int64_t n, double out[20]);
int dgv_moments_high_dev(dgv_ctx* ctx, const double* d_f, int64_t n, double out[20]);

/* ---- solution checkpoint and driver files (host only) ----------------------------- */
/* MakeBaseNameDGV (readwrite.f90:1505-1560, uniform meshes) + "_time" + trim(adjustl(suff))
 * + "_sol.dat", the name WriteSol_DGV / RestartSolutionDGV use */
int dgv_sol_file_name(const dgv_grid_desc* grid, const char* dir, const char* base, int32_t k_c,
                      int32_t N_x, const char* suff, char* out, size_t cap);
/* MakeBaseNameDGV + suffix ("_macroerr.txt" of WriteErrHighCMoments_DGV,
 * "_L1err.txt" of KeepTrackL1_err_DGV) */
int dgv_log_file_name(const dgv_grid_desc* grid, const char* dir, const char* base, int32_t k_c,
                      int32_t N_x, const char* suffix, char* out, size_t cap);
/* WriteSol_DGV (readwrite.f90:1756-1796): Fortran-unformatted records
 * (e_time, dt), (rkmts), (size(f)), (f, f1..f_{k-1}, frhs1..frhs_{k-1});
 * hist = f1..f4, frhs1..frhs5 (9 n doubles, may be NULL for rkmts 1) */
int dgv_sol_write_file(const char* path, double e_time, double dt, int32_t rkmts, int64_t n,
                       const double* f, const double* hist);
/* WriteSol_init_DGV (readwrite.f90:1810-1850): the same records with f alone
 * in the last, whatever rkmts (the driver's initial-data file, suffix "i") */
int dgv_sol_write_init_file(const char* path, double e_time, double dt, int32_t rkmts, int64_t n,
                            const double* f);
/* RestartSolutionDGV (readwrite.f90:2389-2513): the reverse; fails when the
 * stored size differs from n.  rkmts receives the file's order. */
int dgv_sol_read_file(const char* path, double* e_time, double* dt, int32_t* rkmts, int64_t n,
                      double* f, double* hist);
/* sum_i w_i psi_k(v_i) q_i for psi = 1,u,v,w,|u|^2 (RecordMomDerivRelaxTimes_DGV's
 * conservation check, dgvtools_mod.f90:1680-1700), FP64 on the device */
int dgv_collision_moments(dgv_ctx* ctx, const double* q, int64_t n, double out[5]);
/* FP64 projection of q onto the complement of the collision invariants in the
 * w-weighted inner product: afterwards sum_i w_i psi_k q_i = 0 for all k.
 * No counterpart in the reference (its conservation is only as good as A). */
int dgv_project_invariants(dgv_ctx* ctx, double* q, int64_t n);
int dgv_project_invariants_dev(dgv_ctx* ctx, double* d_q, int64_t n);

/* ---- batched evaluation: many VDFs, one or several grids ----------------------- */
/* No counterpart in the reference (it evaluates one VDF per MPI job).  An
 * instance is one VDF on one of the given grid contexts (each with its A-array).
 * f of all instances is packed per grid in the [cell][local][instance] order. */
typedef struct dgv_batch dgv_batch;
/* precision: DGV_FP64; DGV_FP32_COMP (FP32 products and matrix-core sums over
 * 32 pairs, two-sum compensated FP32 accumulation, FP64 final sums);
 * DGV_FP32_F64ACC (the same FP32 products and sums, accumulated in FP64);
 * DGV_INT8_SPLIT (coefficients as 38-bit and products f f as 30-bit fixed-point
 * integers, per block and per cell scaled, split into signed base-256 digits;
 * the digit products with weight >= 256^-4 of the leading one are summed exactly
 * on the int8 matrix cores in int32, per block, then scaled and added in FP64) */
int dgv_batch_create(dgv_ctx* const* grids, int32_t ngrids, const int32_t* grid_of_instance,
                     int32_t ninstances, int32_t precision, dgv_batch** out);
/* A batch reads its grids' A-arrays: once one of them is replaced, truncated or
 * changed in place, every call that would evaluate returns DGV_ERR_STATE (create
 * a new batch).  Destroy may come before or after the grids' contexts are
 * destroyed; no other call may follow a grid context's destruction. */
void dgv_batch_destroy(dgv_batch* b);
/* offset (in doubles) of instance i's first value inside the packed buffers and
 * the stride between its consecutive nodes */
int dgv_batch_layout(const dgv_batch* b, int32_t instance, int64_t* offset, int64_t* stride,
                     int64_t* total);
/* pack / unpack one instance (host arrays in reference node order) */
int dgv_batch_set_f(dgv_batch* b, int32_t instance, const double* f, int64_t n);
int dgv_batch_get_q(dgv_batch* b, int32_t instance, double* q, int64_t n);
/* the same for `count` consecutive instances of one grid at once, from / to
 * device rows (row r = instance first + r, leading dimension ld doubles) */
int dgv_batch_set_f_rows_dev(dgv_batch* b, const double* d_rows, int64_t ld, int32_t first, int32_t count);
int dgv_batch_get_q_rows_dev(dgv_batch* b, double* d_rows, int64_t ld, int32_t first, int32_t count);
int dgv_batch_get_f_rows_dev(dgv_batch* b, double* d_rows, int64_t ld, int32_t first, int32_t count);
/* Q of every instance, one persistent launch over the (instance, cell) queue */
int dgv_batch_eval_q(dgv_batch* b);
/* one RK step (rkmts 1..3) of every instance, optionally hipGraph-replayed */
int dgv_batch_rk_run(dgv_batch* b, double dt, double coef, int32_t rkmts, int64_t nsteps,
                     int use_graph);
/* times segment sizes 2^14..2^21 pairs and 4 or 8 waves per CU, keeps the
 * fastest (best = its segment size, best_ms = its time per evaluation) */
int dgv_batch_autotune(dgv_batch* b, int32_t* best, double* best_ms);
/* segment size (pairs; 0 = default) and persistent-grid size (waves per CU;
 * 0 = keep) of the batched evaluation */
int dgv_batch_set_tile(dgv_batch* b, int32_t pairs_per_segment, int32_t waves_per_cu);
/* on != 0: every evaluation of the batch (dgv_batch_eval_q and the RK stages
 * of dgv_batch_rk_run) projects each instance's Q in FP64 onto the complement
 * of the collision invariants -- the conservation-moment correction.  Off by
 * default (the reference applies none). */
int dgv_batch_set_projection(dgv_batch* b, int32_t on);
/* device pointers of the packed f and q buffers */
int dgv_batch_device_buffers(dgv_batch* b, double** d_f, double** d_q);
/* FP64 projection of every instance's q (packed buffer) onto the complement of
 * the collision invariants, as dgv_project_invariants does for one VDF */
int dgv_batch_project_invariants(dgv_batch* b);
/* Q of every instance on a set of probe cells only (0-based cell ids, valid on
 * every grid): items of the other cells are not planned, so the cost is the
 * probe's share of a full evaluation.  q at the probe cells' nodes is written
 * into the packed q buffer, other nodes are left as they were. */
int dgv_batch_set_probe_cells(dgv_batch* b, const int32_t* cells, int32_t ncells);
int dgv_batch_eval_q_probe(dgv_batch* b);

/* ---- learned surrogate ------------------------------------------------------------ */
/* No counterpart in the reference.  An MLP widths[0] -> ... -> widths[nlayers]
 * with GELU between layers (widths[0] = widths[nlayers] = nodes of the context's
 * grid) maps f / f_scale to Q(f) / q_scale.  BF16 weights and activations on the
 * matrix cores with FP32 accumulation, FP32 master weights, Adam (0.9, 0.999,
 * 1e-8).  The output is scaled back to FP64 and projected onto the complement
 * of the collision invariants, so that sum_i w_i psi_k q_i = 0 to rounding.
 * Rows are VDFs in reference node order, row-major with leading dimension ld
 * (doubles); every row pointer is a device pointer. */
typedef struct dgv_surrogate dgv_surrogate;
int dgv_sur_create(dgv_ctx* ctx, const int32_t* widths, int32_t nlayers, int32_t batch, uint64_t seed,
                   dgv_surrogate** out);
/* may come before or after dgv_ctx_destroy of the surrogate's context */
void dgv_sur_destroy(dgv_surrogate* s);
int dgv_sur_set_scales(dgv_surrogate* s, double f_scale, double q_scale);
int dgv_sur_layer_shape(const dgv_surrogate* s, int32_t layer, int32_t* in, int32_t* out);
/* FP32 master weights W[out][in] and bias[out] of one layer (host arrays) */
int dgv_sur_get_layer(dgv_surrogate* s, int32_t layer, float* W, float* b);
int dgv_sur_set_layer(dgv_surrogate* s, int32_t layer, const float* W, const float* b);
/* gradients of the last training step's loss */
int dgv_sur_get_grad(dgv_surrogate* s, int32_t layer, float* dW, float* db);
/* one forward and backward pass over nrows <= batch rows of (f, Q); the loss is
 * the mean squared error of the scaled output over the nrows x N entries; lr > 0
 * applies one Adam step, lr = 0 leaves the weights unchanged */
int dgv_sur_train_step(dgv_surrogate* s, const double* d_f, const double* d_q, int64_t ld, int32_t nrows,
                       double lr, double* loss);
/* surrogate Q of nrows rows (batch rows per pass); project = 1 applies the FP64 projection */
int dgv_sur_predict(dgv_surrogate* s, const double* d_f, int64_t ld, int64_t nrows, double* d_q,
                    int32_t project);
/* The direct evaluator as judge and fallback: on the probe cells the batched
 * FP64 engine gives each row's residual ||q_sur - q_direct|| / ||q_direct||,
 * with q_sur the surrogate's Q before the projection; rows within tol receive
 * the projected surrogate Q, the others the direct FP64 Q on every node.
 * residual and fallback (host, nrows entries) may be NULL. */
int dgv_sur_set_probe_cells(dgv_surrogate* s, const int32_t* cells, int32_t ncells);
int dgv_sur_eval(dgv_surrogate* s, const double* d_f, int64_t ld, int64_t nrows, double* d_q, double tol,
                 double* residual, int32_t* fallback, int64_t* nfallback);
int dgv_sur_time_predict(dgv_surrogate* s, const double* d_f, int64_t ld, int64_t nrows, double* d_q,
                         int32_t niter, double* ms_per_call);
/* Decomposition mode (on = 1): Q(f) = Q(fm) + L(Df) + Q(Df) exactly, with
 * Df = f - fm, fm and L = 2 MATMUL(., fmA) the context's prepared linearised
 * operator (dgv_prepare_fma; collision_mod.f90:346-359) and Q(fm) evaluated
 * once by this call.  This is the reference's run_mode 1 split
 * (spat_oper_mod.f90:28-68) plus Q(fm), which run_mode 1 leaves out.  The
 * network gives the quadratic part from the direction of Df: with s the root
 * mean square of a row's Df, Q(Df) = s^2 net(Df / s).  dgv_sur_predict,
 * dgv_sur_eval and dgv_sur_time_predict then take rows of f; the judge and the
 * fallback still evaluate Q(f) directly.  Training takes the rows
 * (Df / s, (Q(f) - Q(fm) - L(Df)) / s^2).  DGV_ERR_STATE without a prepared
 * operator, or (at predict / eval) when it was rebuilt or dropped since this
 * call.  on = 0 turns the mode off. */
int dgv_sur_set_decomposition(dgv_surrogate* s, int32_t on);

/* ---- stream / timing helpers ----------------------------------------------------- */
int dgv_ctx_synchronize(dgv_ctx* ctx);
/* times niter calls of dgv_eval_q_dev on the context's stream with HIP events;
 * returns the mean milliseconds per call and the mean of the dominant kernel */
int dgv_time_eval_q_dev(dgv_ctx* ctx, const double* d_f, double* d_fc, int32_t niter,
                        double* ms_per_call, double* ms_main_kernel);
int dgv_batch_time_eval(dgv_batch* b, int32_t niter, double* ms_per_call,
                        double* ms_main_kernel);
/* device memory helpers for callers without a runtime of their own */
int dgv_device_alloc(dgv_ctx* ctx, size_t bytes, void** dptr);
int dgv_device_free(dgv_ctx* ctx, void* dptr);
int dgv_memcpy_htod(dgv_ctx* ctx, void* dst, const void* src, size_t bytes);
int dgv_memcpy_dtoh(dgv_ctx* ctx, void* dst, const void* src, size_t bytes);
/* device-to-device copy by a plain kernel (16-byte aligned pointers and size) */
int dgv_device_copy(dgv_ctx* ctx, void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
