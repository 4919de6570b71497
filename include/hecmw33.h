/*
 * hecmw33 — MI355X (gfx950) solver for hecmw1's 3-DOF-per-node linear systems.
 *
 * C ABI of libhecmw33.so.  Plain pointers and sizes only; arrays follow the
 * reference's hecmwST_matrix layout verbatim (hecmw1/src/common/hecmw_util_f.F90:458-470):
 *   D(9*NP), AL(9*NPL), AU(9*NPU): 3x3 blocks, row-major inside a block
 *   indexL/indexU(0:NP): cumulative counts (0-based offsets), NP+1 entries
 *   itemL/itemU: 1-based node ids (itemL < row, itemU > row, itemU may be > N)
 *   B, X (3*NP): xyz interleaved per node
 *   iarray[100], rarray[100]: hecMAT%Iarray / hecMAT%Rarray, 1-based slot k is
 *   element k-1 (index map: hecmw1/src/solver/matrix/hecmw_matrix_misc.F90:107-150)
 * Host arrays are never freed or resized by the library; B and the matrix are
 * left bitwise unchanged.
 *
 * Return codes are the reference's (hecmw1/src/solver/init/hecmw_solve_error.f90:9-16);
 * negative codes are library failures (see hecmw33_last_error()).
 */
#ifndef HECMW33_H
#define HECMW33_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  HECMW33_OK = 0,
  HECMW33_ERROR_INCONS_PC = 1001,
  HECMW33_ERROR_ZERO_DIAG = 2001,
  HECMW33_ERROR_ZERO_RHS = 2002,
  HECMW33_ERROR_NOCONV_MAXIT = 3001,
  HECMW33_ERROR_DIVERGE_MAT = 3002,
  HECMW33_ERROR_DIVERGE_PC = 3003,
  HECMW33_ERROR_DIVERGE_NAN = 3004,
  HECMW33_ERROR_ARGS = -1,     /* invalid sizes or pointers */
  HECMW33_ERROR_DEVICE = -2,   /* HIP runtime failure or no GPU */
  HECMW33_ERROR_UNSUPPORTED = -3
};

/* The system: pointers into the caller's hecmwST_matrix arrays. */
typedef struct hecmw33_sys {
  int32_t N, NP, NPL, NPU;
  const double* D;
  const double* AL;
  const double* AU;
  const int32_t* indexL;
  const int32_t* indexU;
  const int32_t* itemL;
  const int32_t* itemU;
} hecmw33_sys;

/* What hecmw_solve_CG / hecmw_solve_BiCGSTAB return through their arguments,
 * plus the dispatcher's post-solve figures. */
typedef struct hecmw33_out {
  int32_t iter;          /* ITER (the Fortran DO variable: MAXIT+1 when not converged) */
  int32_t error;         /* 0 or a reference error code */
  int32_t method;        /* METHOD actually used (after a METHOD2 fall-back) */
  int32_t ncolor;        /* colours of the SSOR ordering, 0 otherwise */
  double resid;          /* RESID of the Krylov loop */
  double resid_true;     /* hecmw_rel_resid_L2 after the solve (hecmw33_solve only) */
  double Tset, Tsol, Tcomm;
  double time_Ax, time_precond;  /* filled when TIMELOG (Iarray 22) >= 1 */
  double t_upload;       /* host->device copies of this call (not part of Tset/Tsol) */
} hecmw33_out;

/* ---- the drop-in entry points --------------------------------------------- */

/* Replaces hecmw_solve_iterative for NDOF == 3
 * (hecmw1/src/solver/iterative/hecmw_solver_Iterative.f90:13-187): parameter
 * read-out, zero-diagonal check (returned as 2001 instead of aborting), the
 * preconditioner recycle setting (writes iarray 96/97/98), the METHOD2
 * fall-back loop, hecmw_rel_resid_L2 and the converged/diverged flags
 * (iarray 81/82).  METHOD 1 (CG) and 2 (BiCGSTAB), PRECOND 1/2 (multicolour
 * SSOR) and 3 (DIAG), ITERPREMAX >= 0.  X: x0 in, solution out (host memory). */
int32_t hecmw33_solve(const hecmw33_sys* A, const double* B, double* X, int32_t iarray[100], double rarray[100],
                      hecmw33_out* out);

/* Replaces one call of hecmw_solve_CG (METHOD 1,
 * hecmw1/src/solver/iterative/hecmw_solver_CG.f90:19-320) or
 * hecmw_solve_BiCGSTAB (METHOD 2, hecmw_solver_BiCGSTAB.f90:16-308), same
 * control flow, breakdown codes and residual recomputation cadence.  The
 * preconditioner is set up as hecmw_precond_setup does (reuse unless
 * iarray 97/98 ask for a new one; both are cleared afterwards). */
int32_t hecmw33_solve_krylov(const hecmw33_sys* A, const double* B, double* X, int32_t iarray[100],
                             double rarray[100], hecmw33_out* out);

/* Mixed precision (no counterpart in the reference's CG; selected through the
 * SOLVER_OPT block that CG/BiCGSTAB with PRECOND 1-3 never read,
 * hecmw_matrix_misc.F90:138):
 *   iarray[40] (Iarray 41) = 1 with METHOD 1, ITERPREMAX <= 1: flexible CG
 *   (Notay's FCG(1)) whose SSOR/DIAG sweeps and work vectors are FP32,
 *   with FP64 inner products, scalars and x; the residual is replaced by
 *   b - A x computed in FP64 whenever the recurrence residual has dropped by
 *   rarray[40] (Rarray 41, default 1e-2), every 50 iterations and before any
 *   convergence is declared, so RESID is always an FP64 true residual.
 *   iarray[41] (Iarray 42) picks the matrix values of its q = A p: 1 FP32,
 *   2 FP64 (FP32 vectors, FP64 accumulation), 0 (default) FP32 with the
 *   two-level correction below, otherwise FP64 when the stiffness contrast
 *   max/min |tr D_ii| exceeds 1e4 (over the long solves of high-contrast
 *   systems an FP32-rounded operator loses the search directions'
 *   conjugacy), else FP32.
 *   iarray[42] (Iarray 43) > 0 adds a two-level correction to the flexible CG's
 *   preconditioner, z = M^-1 r + Z E^-1 Z^T r: Z holds the six rigid-body modes
 *   (three translations, three rotations about the centroid) of each aggregate;
 *   E = Z^T A Z is inverted densely (FP64 Cholesky, FP32 copy for the
 *   iterations).  Aggregates: the nodes of each cell of a uniform grid over the
 *   node coordinates given by hecmw33_set_coords / hecmw33_ctx_set_coords.  1 aims
 *   at 1000 aggregates, larger values at that many (at most 2048).  If E is not
 *   positive definite the solve runs without it (hecmw33_ctx_info [11] = -1).
 *   iarray[43] (Iarray 44) picks its form: 0 (default) the symmetric step
 *   z1 = Q r, z2 = z1 + M^-1 (r - A z1), z = z2 + Q (r - A z2) with Q = Z E^-1 Z^T
 *   (A Z is kept per node and neighbouring aggregate, and q = A p is carried as
 *   A z + beta q, so an iteration still has one SpMV); 1 the additive
 *   z = M^-1 r + Q r; 2 the deflation form y = M^-1 r, z = y + Q (r - A y)
 *   (A-DEF2: one coarse solve per iteration, the symmetric step's spectrum).
 *   iarray[46] (Iarray 47) picks the L'/U' values the FP32 SSOR sweeps read: 0
 *   (default) BF16 copies rounded to nearest even from the FP32 image (22 B per
 *   off-diagonal block instead of 40; D^-1 stays FP32, and the rounding is
 *   entrywise, so M stays symmetric positive definite), 1 FP32.
 *   Accepted by hecmw33_solve, hecmw33_solve_krylov and hecmw33_ctx_solve_device. */

/* Node coordinates for the two-level correction: xyz interleaved, 3*NP doubles
 * (hecMESH%node, hecmw1/src/common/hecmw_util_f.F90:264; the reference's ML
 * preconditioner reads the same array for its rigid-body modes,
 * hecmw1/src/solver/precond/ml/hecmw_ML_helper.f90:35-37).  Copied; a
 * new call replaces them and rebuilds the aggregation at the next solve. */
int32_t hecmw33_set_coords(int32_t NP, const double* xyz);

/* y = A x for the internal rows (hecmw_matvec_33, hecmw1/src/solver/las/hecmw_solver_las_33.F90:34-59).
 * Host pointers, x has 3*NP entries, y receives 3*N. */
int32_t hecmw33_matvec(const hecmw33_sys* A, const double* x, double* y);

/* Multipoint constraints of the MPC-CG method (hecMESH%mpc, hecmw1/src/common/hecmw_util_f.F90:97-104):
 * n_mpc constraints over N internal nodes, mpc_index[0..n_mpc] cumulative (0-based), and per term
 * mpc_item (1-based node), mpc_dof (1..3) and mpc_val; the first term of a constraint is its slave.
 * Copied; n_mpc = 0 clears.  hecmw33_solve then multiplies by T^T A T when MPC_METHOD (iarray[12])
 * is 2, as hecmw_solve_iterative does (hecmw_solver_Iterative.f90:79-84,163-166), and
 * hecmw33_solve_krylov whenever iarray[82] (the MPC matvec flag) is set.  Constraints with a dof
 * above 3 are skipped as the reference skips them; chained constraints (a slave that is another
 * constraint's slave or master) and constraints on external nodes return HECMW33_ERROR_UNSUPPORTED. */
int32_t hecmw33_set_mpc(int32_t N, int32_t n_mpc, const int32_t* mpc_index, const int32_t* mpc_item,
                        const int32_t* mpc_dof, const double* mpc_val);

/* y = T^T A T x with the constraints of hecmw33_set_mpc (hecmw_TtmatTvec_33,
 * hecmw1/src/solver/las/hecmw_solver_las_33.F90:572-586; hecmw_Tvec_33 / hecmw_Ttvec_33 :442-565).
 * Host pointers, x has 3*NP entries, y receives 3*N. */
int32_t hecmw33_matvec_mpc(const hecmw33_sys* A, const double* x, double* y);

/* z = M^-1 r with the preconditioner selected by iarray/rarray
 * (hecmw_precond_setup + hecmw_precond_apply, hecmw1/src/solver/precond/hecmw_precond.f90:30-172).
 * A new setup is always made.  Host pointers, 3*NP entries. */
int32_t hecmw33_precond_apply(const hecmw33_sys* A, const int32_t iarray[100], const double rarray[100],
                              const double* r, double* z);

/* sqrt(|b - A x|^2 / |b|^2) (hecmw_rel_resid_L2_33, hecmw_solver_las_33.F90:406-435). */
int32_t hecmw33_rel_resid_L2(const hecmw33_sys* A, const double* B, const double* X, double* resid);

/* hecmw33_rel_resid_L2 for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  Every array of A, B, X (3*NP each) and resid
 * (one double, written) is in device memory; the work is enqueued on `stream`
 * (a hipStream_t; NULL: the null stream) and nothing is allocated: the partial
 * sums live in a fixed device array of HECMW33_NOALLOC_MAX_PARTS pairs, one per
 * 256 rows, so N above 256 * HECMW33_NOALLOC_MAX_PARTS rows is refused
 * (HECMW33_ERROR_ARGS).  Each GPU holds one such array (1 MiB) until the process
 * ends; calls that overlap on one GPU share it and race. */
#define HECMW33_NOALLOC_MAX_PARTS 65536
int32_t hecmw33_rel_resid_L2_device_noalloc(const hecmw33_sys* A, const double* B, const double* X, double* resid,
                                           void* stream);

/* Per-iteration relative residuals of the last hecmw33_solve / hecmw33_solve_krylov
 * call, the values the reference prints with ITERLOG (hecmw_solver_CG.f90:253);
 * returns the number written (at most cap). */
int32_t hecmw33_last_history(double* hist, int32_t cap);

/* RCM + capped greedy multicolouring of the internal rows, as
 * hecmw_precond_SSOR_33_setup builds it (hecmw_matrix_ordering_CM.f90:65-78 then
 * hecmw_matrix_ordering_MC.f90:15-72).  Host only.  perm_rcm, perm, iperm:
 * N entries of 1-based ids; colorindex: N+1 entries, *ncolor colours. */
int32_t hecmw33_ordering(int32_t N, const int32_t* indexL, const int32_t* itemL, const int32_t* indexU,
                         const int32_t* itemU, int32_t ncolor_in, int32_t* perm_rcm, int32_t* perm,
                         int32_t* iperm, int32_t* ncolor, int32_t* colorindex);

/* ---- device-resident use (no reference counterpart: for callers and benchmarks
 *      that keep the system in HBM between solves) ------------------------------ */

typedef struct hecmw33_ctx hecmw33_ctx;

hecmw33_ctx* hecmw33_ctx_create(void);
void hecmw33_ctx_destroy(hecmw33_ctx* ctx);
/* Upload the matrix (structure when new, values always) and set up the
 * preconditioner that iarray/rarray select. */
int32_t hecmw33_ctx_load(hecmw33_ctx* ctx, const hecmw33_sys* A, const int32_t iarray[100], const double rarray[100]);
/* Krylov solve with B, X already in device memory (3*NP doubles each). */
int32_t hecmw33_ctx_solve_device(hecmw33_ctx* ctx, const double* B_dev, double* X_dev, const int32_t iarray[100],
                                 const double rarray[100], hecmw33_out* out);
/* `count` back-to-back products y = A x on device vectors; returns the mean
 * kernel time in *ms (HIP events on the context's stream). */
int32_t hecmw33_ctx_matvec_device(hecmw33_ctx* ctx, const double* x_dev, double* y_dev, int32_t count, double* ms);
int32_t hecmw33_ctx_matvec_device_f32(hecmw33_ctx* ctx, const float* x_dev, float* y_dev, int32_t count, double* ms);
/* `count` back-to-back multicolour SSOR applications z = M^-1 r (the context's SSOR set-up,
 * hecmw_precond_SSOR_33_apply's sweeps) on device vectors of 3*NP entries: FP64 (fp32 = 0),
 * or the FP32 image the flexible CG uses on float vectors, with its L'/U' values in BF16
 * (fp32 = 1, Iarray 47 = 0) or FP32 (fp32 = 2, Iarray 47 = 1); mean time per application in
 * *ms (HIP events on the context's stream). */
int32_t hecmw33_ctx_ssor_device(hecmw33_ctx* ctx, const void* r_dev, void* z_dev, int32_t fp32, int32_t count,
                                double* ms);
/* No counterpart in the reference (a measurement aid for the benchmark): the
 * HBM read bandwidth a plain streaming kernel reaches on this GPU, in GB/s,
 * best of a few launch shapes, each timed over `reps` reads of the n doubles
 * at buf_dev (device memory, n even, 16-byte aligned).  part_dev: device
 * scratch of at least 65536 doubles, overwritten. */
int32_t hecmw33_stream_read(const double* buf_dev, int64_t n, double* part_dev, int32_t reps, double* gbs);
/* Residual history of the last solve (ITERLOG values), up to cap entries. */
int32_t hecmw33_ctx_history(hecmw33_ctx* ctx, double* hist, int32_t cap);
/* Context figures: [0] SELL-64 slots of A (incl. padding), [1] N, [2] NPL, [3] NPU, [4] colours,
 * [5] SSOR slots of L' + U', [6] FP64 residual replacements of the last flexible CG solve,
 * [7] the iteration at which the last learned solve dropped its two-level step (0: kept), [8] 1 if
 * that solve ran without the learned step (or was not a learned solve), [9] 1 if the last flexible
 * CG used the FP64 matrix values in q = A p, [10] learned coarse modes of the current fit (over all
 * aggregates; 0: none fitted), [11] aggregates of the last flexible CG's two-level
 * correction (0: none, -1: E was not positive definite and the solve ran without it), [12] coupled
 * aggregate pairs (blocks of E), [13] A Z entries the symmetric two-level step streams (after
 * dropping the rows whose A Z is zero in exact arithmetic), [14] A Z entries before that (node and
 * touched aggregate); 0 and 0 when the symmetric step has not been set up; [15] modes per
 * aggregate of the last two-level solve (6, or 12 with the learned ones), [16] error samples and
 * [17] FCG iterations per sample of the last learned fit, [18] its harvest and [19] its fit time
 * (microseconds), [20] coarse unknowns (modes per aggregate x aggregates). */
int32_t hecmw33_ctx_info(hecmw33_ctx* ctx, int64_t* info, int32_t n);
/* the multipoint constraints of a device context (as hecmw33_set_mpc) */
int32_t hecmw33_ctx_set_mpc(hecmw33_ctx* ctx, int32_t N, int32_t n_mpc, const int32_t* mpc_index,
                            const int32_t* mpc_item, const int32_t* mpc_dof, const double* mpc_val);
/* the two-level correction's node coordinates for a device context (as hecmw33_set_coords) */
int32_t hecmw33_ctx_set_coords(hecmw33_ctx* ctx, int32_t NP, const double* xyz);
/* Inspection of the last two-level set-up: *n_agg aggregates (0: none built) and *n_pair coupled
 * aggregate pairs; any non-null array receives agg (N: aggregate of each node), dv (3N: the
 * scaled offsets from the aggregate centroid that define Z), pair_a / pair_b (n_pair: the pairs
 * (a, b) coupled by a matrix block, b ascending per a) and Eblk (36*n_pair: E_(a,b) row-major 6x6). */
int32_t hecmw33_ctx_coarse_get(hecmw33_ctx* ctx, int32_t* n_agg, int32_t* n_pair, int32_t* agg, float* dv,
                               int32_t* pair_a, int32_t* pair_b, double* Eblk);

/* ---- batched solves (no counterpart in the reference: hecmw_solve_CG handles one system and
 * one right-hand side per call; these solve many systems with up to 16 right-hand sides each
 * as one block-diagonal SpMM) ---------------------------------------------------------------
 * DIAG-preconditioned flexible CG per (system, rhs) with FP32 matrix/SpMM/vectors, FP64 x,
 * inner products and scalars, FP64 residual replacement (Rarray 41 = delta, default 1e-2);
 * RESID (Rarray 1) is met by an FP64 true residual.  Systems are one domain each (NP == N).
 * B[s]: nrhs columns of 3*N_s (column k at B[s] + 3*N_s*k); X[s] likewise.
 * iter/error/resid: nsys*nrhs entries, system-major; error codes as hecmw_solve_CG's. */
typedef struct hecmw33_batch hecmw33_batch;
hecmw33_batch* hecmw33_batch_create(void);
void hecmw33_batch_destroy(hecmw33_batch* b);
/* upload, FP32 copies, DIAG factors (sigma = Rarray 2), tiles, SpMM autotuning */
int32_t hecmw33_batch_load(hecmw33_batch* b, int32_t nsys, const hecmw33_sys* sys, int32_t nrhs,
                           const double* const* B, const int32_t iarray[100], const double rarray[100]);
/* x0 = 0, solve every instance; ITER (Iarray 1), PRECOND must be 3 (DIAG), ITERPREMAX 1.
 * times (optional, 4): set-up s, solve s, tuned SpMM ms, iterations enqueued */
int32_t hecmw33_batch_solve(hecmw33_batch* b, const int32_t iarray[100], const double rarray[100], int32_t* iter,
                            int32_t* error, double* resid, double* times);
int32_t hecmw33_batch_get_x(hecmw33_batch* b, int32_t s, double* X);
/* Y_s = A_s X_s, host arrays.  fp64 = 0: the tuned FP32 SpMM; 1: the tuned FP64
 * shape; 2 + v: FP64 shape v.  Every FP64 shape sums each row's blocks in
 * hecmw_matvec_33's order. */
int32_t hecmw33_batch_spmm(hecmw33_batch* b, int32_t fp64, const double* const* X, double* const* Y);
/* [0] rows, [1] blocks, [2] tiles, [3] tuned FP32 variant, [4] n32 FP32 variants,
 * [5 .. 5+n32) their SpMM times (us, with the fused dots), [5+n32] tuned FP64
 * shape, [6+n32] n64 FP64 shapes, [7+n32 .. 7+n32+n64) their residual-SpMM
 * times (us).  Returns the number of entries. */
int32_t hecmw33_batch_info(hecmw33_batch* b, int64_t* info, int32_t n);
const char* hecmw33_batch_variant_name(int32_t v);
const char* hecmw33_batch_variant64_name(int32_t v);
const char* hecmw33_batch_last_error(void);
/* one shot: load + solve + copy back */
int32_t hecmw33_solve_batch(int32_t nsys, const hecmw33_sys* sys, int32_t nrhs, const double* const* B,
                            double* const* X, const int32_t iarray[100], const double rarray[100], int32_t* iter,
                            int32_t* error, double* resid);

/* ---- learned coarse modes (no counterpart in the reference) ---------------------------------
 * Inside the flexible CG (iarray[40] = Iarray 41 = 2, ITERPREMAX >= 1, the symmetric two-level
 * step: Iarray 43 (1 when not set) > 0, Iarray 44 = 0 or 2) each aggregate gets six learned coarse
 * modes beside its six rigid-body modes.  They are the decoders of per-aggregate linear
 * autoencoders fitted to slow errors the solver generates on the system itself: for seeded random
 * v_s (Iarray 45 samples, default 8, at most 16), b_s = A v_s and Iarray 46 iterations (default 8)
 * of the flexible CG with the rigid-body two-level step from x = 0, e_s = v_s - x_s = A^-1 r_s;
 * per aggregate the part of the e_s beyond the rigid-body modes, its Gram matrix (BF16 MFMA, FP32
 * accumulation) and that matrix's top eigenvectors give the autoencoder's optimum.  The fit is
 * part of the numeric set-up (redone when Iarray 97 asks for it).  E = Z^T A Z keeps the
 * preconditioner symmetric positive definite; if the fit or E fails the rigid-body modes stay.
 * Plain SSOR stays the fallback: the solve runs the plain SSOR flexible CG from the iteration at
 * which its FP64 residual at a replacement exceeds rarray[41] (Rarray 42, default 1000) times the
 * smallest one seen at a replacement (hecmw33_ctx_info [7], [8]).
 *
 * A per-node classifier network (hecmw33_net_*): a 16 -> 64 -> 64 -> 3 MLP (ReLU; BF16 MFMA
 * forward and backward kernels accumulating in FP32; Adam) over static node features, trained
 * on labels built from the solver's errors: an edge of the matrix graph whose axial strain stays
 * below tau times the sample's median in every error sample lies in a rigidly moving piece, and
 * its nodes are labelled 1 (the logistic loss of the network's logit against the labels).  The
 * solve does not use it.  A network must outlive the contexts it is attached to; contexts must
 * outlive its samples. */
/* The learned coarse modes of a device context's last fit: xm (18*N floats: node i, mode j,
 * component d at 18 i + 3 j + d); when the twelve-mode E is assembled, Eblk (144 per coupled
 * aggregate pair, in hecmw33_ctx_coarse_get's pair order: E_(a,b) row-major 12x12, rows and columns
 * 0-5 the rigid-body modes, 6-11 the learned ones); R (48*N floats: the error samples' part beyond
 * their aggregate's rigid-body modes that the fit used, entry 16 (3 i + d) + s for node i,
 * component d, sample s; unused samples zero).  Any may be NULL. */
int32_t hecmw33_ctx_modes_get(hecmw33_ctx* ctx, float* xm, double* Eblk, float* R);
/* The coarse solve operator the iterations of a device context's last two-level set-up read (no
 * counterpart in the reference; for tests): *nc coarse unknowns, *ldw the row stride of W,
 * *fp64 = 0 with six modes per aggregate (W: E^-1 in FP32, nc rows of ldw floats), 1 with twelve
 * (W: FP64, L^-1 on and below the diagonal and L^-T above it, E = L L^T); columns nc..ldw-1 zero.
 * W may be NULL to query the sizes. */
int32_t hecmw33_ctx_coarse_inverse_get(hecmw33_ctx* ctx, int32_t* nc, int32_t* ldw, int32_t* fp64, void* W);

typedef struct hecmw33_net hecmw33_net;
hecmw33_net* hecmw33_net_create(uint64_t seed);
void hecmw33_net_destroy(hecmw33_net* net);
/* FP32 master weights: W1 [16][64], W2 [64][64], W3 [64][16] row-major (W3 columns 3..15 zero) */
int32_t hecmw33_net_weight_count(void);
int32_t hecmw33_net_get_weights(hecmw33_net* net, float* w, int32_t n);
int32_t hecmw33_net_set_weights(hecmw33_net* net, const float* w, int32_t n);
/* Y = N(X) for n x 16 features (host); with dY (n x 3) also dW = dL/dW for L = sum dY.Y */
int32_t hecmw33_net_eval(hecmw33_net* net, int64_t n, const float* X, const float* dY, float* Y, float* dW);
/* an error sample e = A^-1 r (device, FP64, 3*NP) on a loaded context with node coordinates;
 * returns that context's sample count */
int32_t hecmw33_net_add_sample(hecmw33_net* net, hecmw33_ctx* ctx, const double* e_dev, double tau);
int32_t hecmw33_net_clear_samples(hecmw33_net* net);
/* loss[0] = mean logistic loss before training, loss[1..epochs] = mean per epoch */
int32_t hecmw33_net_train(hecmw33_net* net, int32_t epochs, double lr, uint64_t seed, double* loss);
/* mean logistic loss and fraction of nodes classified as their label, over the current samples */
int32_t hecmw33_net_score(hecmw33_net* net, double* mean, double* accuracy);
int32_t hecmw33_ctx_set_net(hecmw33_ctx* ctx, hecmw33_net* net);
int32_t hecmw33_set_net(hecmw33_net* net);
/* The training labels alone, host only: nsamp error samples e (nsamp x 3N, xyz interleaved) of an
 * N-node system with node coordinates xyz (3N); labels (N) receive 1 for the nodes of every edge
 * whose axial strain stays below tau times its sample's median in every sample, else 0.  Returns
 * the number of nodes labelled 1. */
int32_t hecmw33_rigid_labels(int32_t N, const int32_t* indexL, const int32_t* itemL, const int32_t* indexU,
                             const int32_t* itemU, const double* xyz, int32_t nsamp, const double* e, double tau,
                             float* labels);

/* ---- synthetic systems (fistr1's assembly for a structured HEX8 cube) ------- */
int32_t hecmw33_gen_hex8_count(int32_t n, int32_t* N, int32_t* NPL, int32_t* NPU);
int32_t hecmw33_gen_hex8(int32_t n, int32_t mat_kind, const double* params, uint64_t seed, double* D, double* AL,
                         double* AU, int32_t* indexL, int32_t* itemL, int32_t* indexU, int32_t* itemU, double* B,
                         float* elem_E);

/* TET10 (fistr1 type 342) box [0,box[0]]x[0,box[1]]x[0,box[2]] on an nx x ny x nz cell grid
 * (6 tetrahedra per cell), uniform (E, nu), bottom face fixed, nrhs columns of N(0,1) nodal
 * loads on the top face in B (column k at B + 3*N*k). */
int32_t hecmw33_gen_tet10_count(int32_t nx, int32_t ny, int32_t nz, int32_t* N, int32_t* NPL, int32_t* NPU);
int32_t hecmw33_gen_tet10(int32_t nx, int32_t ny, int32_t nz, const double* box, double E, double nu, int32_t nrhs,
                          uint64_t seed, double* D, double* AL, double* AU, int32_t* indexL, int32_t* itemL,
                          int32_t* indexU, int32_t* itemU, double* B);

/* ---- misc -------------------------------------------------------------------- */
const char* hecmw33_last_error(void);
int32_t hecmw33_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* HECMW33_H */
