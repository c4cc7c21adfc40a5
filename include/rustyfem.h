/*
 * rustyfem-mi355x: the C solver ABI of RustyFEM's ffi/ library, served by
 * hand-written HIP kernels on one MI355X.
 *
 * The eight rustyfem_* entry points below keep the reference's signatures and
 * semantics exactly (ffi/src/lib.rs); a caller built against the reference's
 * librustyfem_ffi binds them unchanged.  The functions marked "additive" are
 * new symbols; no reference signature changed.
 *
 * Input of every factorization: a symmetric (or, for _unsym, unsymmetric)
 * matrix in CSC form in the caller's numbering (both triangles, or for the
 * symmetric modes also only the lower one), the node of every DOF (loc, n
 * entries) and the node coordinates (coords, 3*nn entries).  Ordering is
 * geometric nested dissection at node level with 16 nodes per leaf.
 */
#ifndef RUSTYFEM_H
#define RUSTYFEM_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct Handle Handle; /* opaque; owned by the library until rustyfem_free */

/* Positive definite multifrontal LDL^T.  Returns NULL on error: *err_dof is the
 * caller DOF of the first non-positive pivot, or -1 for invalid input.  d_out
 * (n entries, may be NULL) receives the pivots D in caller order.
 * Replaces ffi/src/lib.rs:64-97 (rustyfem_factor). */
Handle *rustyfem_factor(size_t n, size_t nn, const int64_t *loc, const double *coords, const int64_t *indptr,
                        const int64_t *indices, const double *data, double *d_out, int64_t *err_dof);

/* Symmetric indefinite LDL^T with 1x1/2x2 pivots and delayed pivots.
 * inertia_out (3 entries, may be NULL) = [negative, zero, positive].
 * Replaces ffi/src/lib.rs:272-290 (rustyfem_factor_indef). */
Handle *rustyfem_factor_indef(size_t n, size_t nn, const int64_t *loc, const double *coords, const int64_t *indptr,
                              const int64_t *indices, const double *data, int64_t *inertia_out, int64_t *err_dof);

/* Unsymmetric multifrontal LU on the symmetrized pattern (both triangles given).
 * Replaces ffi/src/lib.rs:299-311 (rustyfem_factor_unsym). */
Handle *rustyfem_factor_unsym(size_t n, size_t nn, const int64_t *loc, const double *coords, const int64_t *indptr,
                              const int64_t *indices, const double *data, int64_t *err_dof);

/* Inertia [negative, zero, positive]; 0, or -1 for an LU handle.
 * Replaces ffi/src/lib.rs:319-329 (rustyfem_inertia). */
int32_t rustyfem_inertia(const Handle *h, int64_t *out);

/* Solves K X = B in place for nrhs right-hand sides stored column by column
 * (n each).  0 on success, -1 on an internal error.
 * Replaces ffi/src/lib.rs:336-371 (rustyfem_solve). */
int32_t rustyfem_solve(const Handle *h, size_t nrhs, double *b);

/* out[0] = estimated flops of the factorization, out[1] = stored entries of L.
 * Replaces ffi/src/lib.rs:377-383 (rustyfem_info). */
void rustyfem_info(const Handle *h, double *out);

/* Internal equation of every caller DOF; uncoupled DOFs numbered last.
 * Replaces ffi/src/lib.rs:390-398 (rustyfem_order). */
void rustyfem_order(const Handle *h, int64_t *out);

/* Frees a handle (NULL-safe, at most once).  Replaces ffi/src/lib.rs:402-407. */
void rustyfem_free(Handle *h);

/* ---------------------------------------------------------------- additive */

/* Options of the additive entry points. */
typedef struct rustyfem_options {
  int32_t precision;  /* 0: FP64 factor (the reference's arithmetic class);
                         1: mixed -- FP32 Schur updates on MFMA with FP32
                            accumulation, FP64 pivot blocks, FP64 refinement */
  int32_t leaf;       /* nodes per nested-dissection leaf; 0 = 16 (ffi/src/lib.rs:17) */
  int32_t max_refine; /* refinement corrections per solve; -1 = default
                         (0 for precision 0, 30 for precision 1) */
  int32_t reserved;
  double refine_tol;  /* refinement target ||b - A x|| / ||b|| (0 = 1e-10) */
} rustyfem_options;

/* rustyfem_factor with options (additive; no reference counterpart). */
Handle *rustyfem_factor_ex(size_t n, size_t nn, const int64_t *loc, const double *coords, const int64_t *indptr,
                           const int64_t *indices, const double *data, const rustyfem_options *opt,
                           double *d_out, int64_t *err_dof);

/* New values on the same pattern (the caller's indptr/indices of the first
 * factorization), refactored on the GPU.  data == NULL refactors the values
 * already resident in device memory.  0 on success, 1 on a failing pivot
 * (*err_dof set), -1 on an internal error.  Mirrors the reference's
 * SystemMatrix zero/add/factor cycle (src/solver/api.rs:12-60) (additive). */
int32_t rustyfem_refactor(Handle *h, const double *data, int64_t *err_dof);

/* Solve with device pointers (caller numbering, column-major n x nrhs, in
 * place); runs refinement when the handle's options ask for it (additive). */
int32_t rustyfem_solve_device(const Handle *h, size_t nrhs, double *d_b);

/* rustyfem_solve_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The solve's work vectors come from one fixed scratch array per device
 * (64 MiB, held until the process ends); the kernels run eagerly (no graph
 * is instantiated).  Limits: positive definite single-problem handles, at
 * most 64 right-hand sides, and work vectors (2 sum(m) + 4 n + the partial
 * sums, in doubles, per right-hand side) that fit the 64 MiB; larger calls
 * are rejected with an error (-1, rustyfem_last_error).  Calls that overlap
 * on one device race on its scratch array (calls on one handle are
 * serialized).  The refinement runs as in rustyfem_solve_device, but a mixed
 * handle that misses its refinement target is not repeated in FP64 (that
 * would allocate): the call returns 1 instead.  0 on success. */
int32_t rustyfem_solve_device_noalloc(const Handle *h, size_t nrhs, double *d_b);

/* Shift-invert block subspace iteration with Rayleigh-Ritz (additive; the
 * reference's src/solver/eigen.rs:229-375 `subspace`, a consumer of the
 * factorization).  h holds the positive definite factorization of
 * Op = K - sigma M; B (= M) is symmetric n x n, given in CSC with both
 * triangles in the caller's numbering.  Finds the nev values theta of largest
 * magnitude of Op x = (1/theta) B x, i.e. the eigenvalues
 * lambda = sigma + 1/theta of K x = lambda M x closest to sigma, with
 * nvec iteration vectors (0 = min(2 nev, nev + 8)), at most max_iter
 * iterations (0 = 2000), until every requested Ritz vector has
 * ||Op^-1 B x - theta x||_Op <= tol |theta| (tol 0 = 1e-10).  theta_out
 * (nev entries, largest |theta| first); x_out (n x nev, column-major, may be
 * NULL): vectors with x^T Op x = 1; iterations_out may be NULL.  Returns the
 * number of pairs found (fewer than nev when B has lower rank), -1 on an
 * error (rustyfem_last_error), including no convergence. */
int64_t rustyfem_eigen(const Handle *h, const int64_t *b_indptr, const int64_t *b_indices, const double *b_data,
                       size_t nev, size_t nvec, size_t max_iter, double tol, double *theta_out, double *x_out,
                       int64_t *iterations_out);

/* Statistics of the last factorization / solve (additive):
 * out[0] factor ms, [1] solve ms, [2] refined-solve ms, [3] refinement steps,
 * [4] relative residual after refinement (-1 if not refined), [5] flops,
 * [6] device bytes held, [7] precision, [8] number of fronts, [9] levels,
 * [10] summed duration (ms) of the Schur-update GEMM launches, [11] their
 * count and [12] their algorithmic flops -- [10] and [11] are filled when the
 * environment variable RUSTYFEM_KERNEL_TIMING=1 was set at factorization;
 * [13] time (ms) of the GEMM autotuning done before the first factorization
 * (RUSTYFEM_AUTOTUNE=0 turns it off), [14] the GEMM launches it timed and
 * [15] how many of them it moved to the AGPR-accumulator kernel,
 * [16..19] the worst relative residual after 0, 1, 2 and 3 refinement
 * corrections of the last solve (-1 where not reached), [20] how many GEMM
 * launches the autotuning moved to the kernels with double-width K tiles,
 * [21] 1 when the last refinement took back a correction that raised the
 * worst residual (the iterate before it is the result), [22] 1 when the last
 * solve of a mixed-precision handle missed the refinement target and was
 * repeated with an FP64 factorization of the same values (built at the first
 * such solve, kept until the next refactorization), [23] the relative
 * residual that repeat reached (-1 if none ran), [24] how many GEMM launches
 * the autotuning moved to the kernel with 16-byte accesses (FP32 only),
 * [25] how many tree levels the triangular solves take through the fused
 * per-front kernels (levels whose fronts all have at most 384 rows and 128
 * pivots; with more than four right-hand sides only those of at most 192
 * rows run them; RUSTYFEM_NO_SMALL_SOLVE=1 in the environment at
 * factorization: none).
 * At most 26 entries are written. */
void rustyfem_stats(const Handle *h, double *out, size_t len);

/* Last error message of the calling thread ("" if none) (additive). */
const char *rustyfem_last_error(void);

/* Pivot sequence of a factored handle (additive, for parity checks), in the
 * front postorder of src/sparse/indef.rs:236-253 and src/sparse/lu.rs:216-241:
 * internal equation, d, e and kind of each pivot.  LDL^T: kind 1 = 1x1 pivot
 * d, kind 2 = first of a 2x2 block [d_i e_i; e_i d_i+1], kind 0 = its second.
 * LU: d = the U diagonal, kind 1, or kind 3 with e = the row permutation entry
 * inside the partial-pivoting block of a root.  Positive definite: equation
 * order, D, kind 1.  Returns the number of pivots (n_sym), or -1. */
int64_t rustyfem_pivots(const Handle *h, int64_t *eq, double *d, double *e, int32_t *kind);

/* ------------------------------------------------ batches of problems (additive) */

typedef struct BatchHandle BatchHandle;

/* Factors nprob independent positive definite problems together (arrays of
 * per-problem inputs with the meaning of rustyfem_factor's arguments).  Each
 * problem is ordered as rustyfem_factor orders it alone; the elimination trees
 * form one forest, so every GPU launch covers the fronts of all problems at
 * the same tree level.  err_dof (nprob entries) is -1 or the caller DOF of
 * the failing pivot / invalid input (-1 with NULL for invalid loc/indices).
 * NULL on any failure.  No reference counterpart: the reference factors one
 * problem per call. */
BatchHandle *rustyfem_factor_batch(size_t nprob, const size_t *n, const size_t *nn, const int64_t *const *loc,
                                   const double *const *coords, const int64_t *const *indptr,
                                   const int64_t *const *indices, const double *const *data,
                                   const rustyfem_options *opt, int64_t *err_dof);
/* New values (data != NULL) or the values in device memory (data == NULL);
 * 0, 1 on a failing pivot (err_dof set), -1 on an internal error. */
int32_t rustyfem_refactor_batch(BatchHandle *h, const double *const *data, int64_t *err_dof);
/* Solves every problem in place: b[i] holds nrhs columns of problem i. */
int32_t rustyfem_solve_batch(const BatchHandle *h, size_t nrhs, double *const *b);
/* Device variant: d_b is the concatenation of all problems' DOFs, column-major
 * (sum n) x nrhs. */
int32_t rustyfem_solve_batch_device(const BatchHandle *h, size_t nrhs, double *d_b);
/* [problems, total DOFs, fronts, levels, work flops] */
void rustyfem_batch_info(const BatchHandle *h, double *out5);
/* as rustyfem_stats */
void rustyfem_batch_stats(const BatchHandle *h, double *out, size_t len);
/* rustyfem_order of problem i */
void rustyfem_batch_order(const BatchHandle *h, size_t i, int64_t *out);
void rustyfem_free_batch(BatchHandle *h);

/* ------------------------------------------------ iterative path (additive) */

/* Options of the GPU preconditioned Krylov solver. */
typedef struct rustyfem_pcg_options {
  int32_t krylov;       /* 0: PCG as src/solver/iterative/csr.rs:341-380; 1: flexible CG (FCG(1)) */
  int32_t vcycle_fp32;  /* AMG V-cycle arithmetic: 0 FP64, 1 FP32 (the Krylov loop stays FP64) */
  int32_t smoother;     /* 0: Chebyshev on [rho/ratio, 1.1 rho] of D^-1 A; 1: damped Jacobi,
                           omega = 4/(3 rho) (src/solver/iterative/amg.rs:300) */
  int32_t degree;       /* smoothing steps before and after the coarse correction; 0 = 2 */
  double cheb_ratio;    /* 0 = 20 */
  int32_t flags;        /* RUSTYFEM_PCG_* bits below; 0 = IterMatrix's behaviour */
  int32_t reserved;
} rustyfem_pcg_options;

/* flags: keep the host AMG hierarchy for rustyfem_pcg_level */
#define RUSTYFEM_PCG_KEEP_HIERARCHY 1
/* flags: no direct fallback (solve_lower's behaviour, src/solver/iterative.rs:275-292:
 * a non-converged solve is reported, not repeated directly) */
#define RUSTYFEM_PCG_NO_DIRECT_FALLBACK 2
/* flags, test hook: the AMG setup reports a breakdown as if the coarse LDL^T had
 * failed, so that the direct fallback can be exercised on a healthy matrix */
#define RUSTYFEM_PCG_FORCE_AMG_BREAKDOWN 4

typedef struct PcgHandle PcgHandle;

/* Preconditioner setup for the lower triangle (CSC, natural numbering: node v
 * owns ndof[v] consecutive equations) -- the cached-preconditioner half of
 * IterMatrix (src/solver/iterative.rs:208-243).  kind: 0 AMG (smoothed
 * aggregation with rigid body modes, src/solver/iterative/amg.rs:282-341;
 * the hierarchy is built on the host, the spectral radii on the GPU),
 * 1 IC(0) (src/solver/iterative/csr.rs:219-324: factored and applied on the
 * GPU level by level, with the diagonal-shift retry of csr.rs:250-256),
 * 2 Jacobi (csr.rs:198-217).  NULL on error (rustyfem_last_error). */
PcgHandle *rustyfem_pcg_setup(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val,
                              int32_t kind, size_t nn, const double *coords, const int64_t *ndof,
                              const rustyfem_pcg_options *opt);

/* rustyfem_pcg_setup with IterMatrix::factor's error report
 * (src/solver/iterative.rs:208-243): a non-positive diagonal entry sends the
 * matrix to the direct solver for diagnosis, and so does an AMG hierarchy
 * that cannot be built.  When the direct factorization (positive definite
 * multifrontal LDL^T, nested dissection with 48-node leaves as
 * src/sparse.rs:14) succeeds the handle solves directly; when it fails the
 * result is NULL with *err_eq = the natural equation of the failing pivot
 * (-1 for other errors).  Without RUSTYFEM_PCG_NO_DIRECT_FALLBACK a solve
 * that does not converge is repeated with the direct solver, as
 * IterMatrix::solve does (src/solver/iterative.rs:249-269);
 * rustyfem_pcg_stats out[40..42] report which fallback ran. */
PcgHandle *rustyfem_pcg_setup_ex(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val,
                                 int32_t kind, size_t nn, const double *coords, const int64_t *ndof,
                                 const rustyfem_pcg_options *opt, int64_t *err_eq);

/* Solves from x0 = 0 until the recursive residual satisfies ||r|| <= tol ||b||
 * (src/solver/iterative/csr.rs:341-380).  b and x are host pointers, or device
 * pointers when on_device != 0.  maxit <= 0 means max(20000, n).  Returns the
 * iterations (-1 on an internal error). */
int64_t rustyfem_pcg_solve(PcgHandle *h, const double *b, double *x, double tol, int64_t maxit, int32_t on_device,
                           double *rel_residual, int32_t *converged);

/* [0] setup ms, [1] last solve ms, [2] levels, [3] operator complexity,
 * [4] ms of one fine-matrix product, [5] its bytes, [6] coarsest size,
 * [7] iterations of the last solve, [8..19] rows per level, [20] IC(0)
 * diagonal shift (Ic0::shift, src/solver/iterative/csr.rs:228), [21] IC(0)
 * forward levels, [22] IC(0) backward levels, [23] IC(0) factor attempts,
 * [24] 1 when the last solve fell back from the learned smoother to plain
 * AMG-PCG, [25] its true relative residual ||b - Ax|| / ||b|| (learned
 * smoother set; else -1), [26] iterations of the abandoned learned-smoother
 * attempt, [27] learned-smoother sweeps (0 = none), [28..39] DOFs per
 * block of the level matrices, finest first (6 x 6 blocks for the aggregates
 * of a 3D hierarchy; 1 = scalar CSR), [40] the handle's direct solver: 0 none,
 * 1 set up because a diagonal entry is not positive, 2 because the AMG
 * hierarchy broke down, 3 built when a solve did not converge, [41] 1 when the
 * last solve's result came from the direct solver, [42] the true relative
 * residual ||b - A x|| / ||b|| of that direct solution (-1 otherwise), [43] 1
 * when the AMG hierarchy was built on the GPU (0: on the host, as with
 * RUSTYFEM_AMG_HOST_SETUP=1 or a matrix row too long for the device setup's
 * LDS tables).  At most 44 entries. */
void rustyfem_pcg_stats(const PcgHandle *h, double *out, size_t len);

/* Host copy of AMG level `level` of a handle set up with
 * RUSTYFEM_PCG_KEEP_HIERARCHY (additive, for parity checks; the arrays of
 * src/solver/iterative/amg.rs:282-341).  Levels 0 .. L-1 are the smoothed
 * levels, level L the coarsest matrix (A and B only).  what:
 *   -1: iout[0..7] = rows, nnz(A), blocks, aggregates, near-null-space columns,
 *       coarse columns (rows of level + 1), nnz(P~), nnz(P)
 *    0: A (CSR: iout = row pointers, idx = columns, dout = values)
 *    1: P~, the tentative prolongator (CSR), 2: P, the smoothed one (CSR)
 *    3: the aggregate of every block (iout), 4: block boundaries (iout, blocks + 1)
 *    5: the near-null space B of this level (dout, rows x columns, row-major)
 * Returns 0, or -1 when the level or the hierarchy does not exist. */
int32_t rustyfem_pcg_level(const PcgHandle *h, int32_t level, int32_t what, int64_t *iout, int32_t *idx,
                           double *dout);

/* Vanek's three-pass aggregation of one level alone (additive, host only, for
 * parity checks: src/solver/iterative/amg.rs:108-181), as the AMG setup runs
 * it: a full symmetric CSR matrix (ptr n + 1, idx, val), DOF blocks
 * [bptr[b], bptr[b + 1]) and the strength threshold theta.  agg receives the
 * aggregate of every block; returns the number of aggregates (-1 on error). */
int64_t rustyfem_amg_aggregate(size_t n, const int64_t *ptr, const int32_t *idx, const double *val, size_t nblk,
                               const int64_t *bptr, double theta, int64_t *agg);

/* Coefficients of the V-cycle's Chebyshev smoother (additive, no counterpart
 * in the reference, whose smoother is Gauss-Seidel): degree steps on
 * [hi / ratio, hi] of D^-1 A, d_0 = c0 D^-1 b, then d_k = c1[k-1] d_{k-1} +
 * c2[k-1] D^-1 (b - A x_k) and x += d_k; the V-cycle uses hi = 1.1 x the
 * power-method radius.  c1, c2 hold degree - 1 entries.  0, or -1 for invalid
 * arguments. */
int32_t rustyfem_chebyshev(double hi, double ratio, int32_t degree, double *c0, double *c1, double *c2);

/* z = M r: one application of the handle's preconditioner (Precond::apply,
 * src/solver/iterative/csr.rs:193-196; for IC(0) csr.rs:296-316, for Jacobi
 * csr.rs:209-215).  r and z are host pointers, or device pointers when
 * on_device != 0.  0 on success, -1 on error. */
int32_t rustyfem_pcg_precondition(PcgHandle *h, const double *r, double *z, int32_t on_device);
void rustyfem_pcg_free(PcgHandle *h);

/* Learned smoother (additive, no counterpart in the reference): the AMG
 * hierarchy's finest smoother becomes x = M b, then x += M (b - A x), `sweeps`
 * steps before and after the coarse correction, with M computed on the GPU by
 * the message-passing network of the learned-smoother parameters (nparams =
 * rustyfem_mpnn_nparams()).  AMG handles with 3 DOFs per node only; use with
 * krylov = 1 (flexible CG).  Plain AMG-PCG is the fallback and the judge:
 * each solve checks the true FP64 residual of its result, and a breakdown,
 * a residual that has not halved over 40 iterations or a true residual above
 * tol restarts the solve with the Chebyshev V-cycle and standard PCG
 * (rustyfem_pcg_stats out[24..26]).  0 on success, -1 on error. */
int32_t rustyfem_pcg_set_learned_smoother(PcgHandle *h, const double *params, size_t nparams, int32_t sweeps);

/* Training interface of the learned smoother (additive).  The network maps a
 * 3x3-block stiffness matrix (lower triangle in CSC, natural node-major
 * numbering, 3 DOFs per node) to a symmetric block matrix M on its block
 * pattern; the layout of the parameters is documented in csrc/mpnn.hpp. */
typedef struct MpnnHandle MpnnHandle;
int64_t rustyfem_mpnn_nparams(void);
MpnnHandle *rustyfem_mpnn_create(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val);
/* L = mean over the nvec columns of e0 (n x nvec, column-major) of
 * e_nu^T A e_nu / e_0^T A e_0 with e_{k+1} = (I - M A) e_k; grad (nparams,
 * may be NULL) receives dL/dparams.  NaN on error. */
double rustyfem_mpnn_loss(MpnnHandle *h, const double *params, const double *e0, size_t nvec, int32_t nu,
                          double *grad);
/* M's blocks (nnzb x 9, row-major 3x3) and block pattern (n/3 + 1 row
 * pointers, nnzb sorted block columns); any output may be NULL.  Returns nnzb
 * (-1 on error). */
int64_t rustyfem_mpnn_smoother(MpnnHandle *h, const double *params, double *blocks, int64_t *ptr, int32_t *idx);
void rustyfem_mpnn_free(MpnnHandle *h);

/* One-shot setup + solve with the default options: mirrors
 * solve_lower(k, b, kind, coords, ndof, tol) -> (x, PcgReport)
 * (src/solver/iterative.rs:275-292). */
int64_t rustyfem_solve_lower(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val,
                             const double *b, int32_t kind, size_t nn, const double *coords, const int64_t *ndof,
                             double tol, int64_t maxit, double *x, double *rel_residual, int32_t *converged);

/* [neg, zero, pos, delayed pivots, 2x2 blocks, factor entries] of a factored
 * handle (additive); the inertia here excludes uncoupled DOFs (the fields of
 * IndefFactor and LuFactor in src/sparse/indef.rs:57-70, src/sparse/lu.rs:100-106). */
void rustyfem_factor_stats(const Handle *h, int64_t *out6);

/* Host-only symbolic analysis of the positive definite path (additive): the
 * ordering rustyfem_order would report (n entries) and rustyfem_info's
 * [flops, factor entries], without touching the GPU.  0 on success, 1 when
 * the input is rejected before the factorization (*err_dof as for
 * rustyfem_factor), -1 on an internal error.  leaf <= 0 means 16. */
int32_t rustyfem_analyse(size_t n, size_t nn, const int64_t *loc, const double *coords, const int64_t *indptr,
                         const int64_t *indices, const double *data, int32_t leaf, int64_t *order_out,
                         double *info_out, int64_t *err_dof);

/* Static hazard check of the GPU factorization schedule for this input
 * (additive, host only): 0 when no tile of any launch writes an arena entry
 * that another tile of the same launch reads or writes, every extend-add
 * launch covers its children's contribution blocks exactly, and no launch that
 * runs on the second stream (the rest of an outer block's update) conflicts
 * with the pivot-block launches that run beside it; 1 with a description in
 * msg otherwise, -1 on an internal error.  plant = 1 adds a known race (an
 * in-place write of each pivot block by its diagonal tile), plant = 2 a known
 * conflict between the two streams (a write of the next outer block's first
 * column by the rest of the update), plant = 3 a gap (an extend-add launch's
 * first tall tile left out), so that tests can check that the check fires;
 * plant = 0 checks the real schedule. */
int32_t rustyfem_check_plan(size_t n, size_t nn, const int64_t *loc, const double *coords, const int64_t *indptr,
                            const int64_t *indices, const double *data, int32_t leaf, int32_t plant, char *msg,
                            size_t msg_len);

/* ------------------------------------------------------------ assembly */
/* Element assembly of linear elastic stiffness matrices on the GPU
 * (additive): replaces the element loop and the serial SystemMatrix::add
 * scatter of src/fem/assembly.rs:315-401 (stiffness() of each element,
 * symmetric part 0.5 (Ke + Ke^T), zero terms skipped) for
 *   RUSTYFEM_C3D8 (1) hexahedron, 2x2x2 Gauss, volumetric B-bar
 *                     (src/elements/continuum3d.rs:39-121),
 *   RUSTYFEM_C3D4 (2) tetrahedron, one point (src/elements/solid_higher.rs:109),
 *   RUSTYFEM_CPS4 (3) quadrilateral, plane stress, 2x2 Gauss,
 *   RUSTYFEM_CPE4 (4) quadrilateral, plane strain, planar B-bar
 *                     (src/elements/continuum2d.rs:24-27),
 * isotropic elastic material per element (src/materials/elastic.rs:55-75).
 * The values equal the reference's sequential evaluation bit for bit. */
#define RUSTYFEM_C3D8 1
#define RUSTYFEM_C3D4 2
#define RUSTYFEM_CPS4 3
#define RUSTYFEM_CPE4 4
typedef struct AssemblyHandle AssemblyHandle;

/* The pattern and the per-entry contribution lists of a mesh (host work, done
 * once): conn nelem x nodes per element (8 or 4) node indices, coords nn x 3
 * (plane elements use x, y), eq nn x DOFs per node (3 or 2) equation numbers
 * in 0..n-1, -1 for a DOF without equation (constrained); every equation
 * belongs to one DOF.  NULL on invalid input (rustyfem_last_error). */
AssemblyHandle *rustyfem_assembly_create(int32_t etype, size_t nelem, const int64_t *conn, size_t nn,
                                         const double *coords, const int64_t *eq, size_t n);
/* stored entries of the assembled matrix (CSC, both triangles) */
int64_t rustyfem_assembly_nnz(const AssemblyHandle *a);
/* indptr (n + 1) and indices (nnz) of that CSC, rows sorted per column */
void rustyfem_assembly_pattern(const AssemblyHandle *a, int64_t *indptr, int64_t *indices);
/* Values (nnz, on the host, or on the device with on_device = 1) for Young's
 * modulus young[e], Poisson's ratio nu[e] and thickness[e] (plane elements;
 * NULL = 1) of every element.  -1 on success, the lowest element whose
 * Jacobian determinant fails the reference's test (an element that is
 * "distorted or inverted", src/elements/common.rs:15), or -2 on an error. */
int64_t rustyfem_assembly_values(AssemblyHandle *a, const double *young, const double *nu, const double *thickness,
                                 double *data, int32_t on_device);
/* Assembles straight into the factorization's own value layout (the lower
 * triangle in its nested-dissection numbering, on the device) and refactors
 * h, whose pattern must be this assembly's (a positive definite handle from
 * rustyfem_factor or rustyfem_factor_ex with rustyfem_assembly_pattern's
 * indptr/indices and every DOF coupled): the reference's zero / add / factor
 * cycle without a host copy of the matrix.  0 on success, 1 on a failing pivot
 * (*err_dof, caller numbering), 2 on a distorted element (*err_elem), -1 on
 * an error. */
int32_t rustyfem_assembly_refactor(AssemblyHandle *a, Handle *h, const double *young, const double *nu,
                                   const double *thickness, int64_t *err_dof, int64_t *err_elem);
/* [element kernel ms, gather kernel ms, host setup s, contributions,
 * doubles per stored element matrix, bound-layout gather ms]; at most 6 */
void rustyfem_assembly_stats(const AssemblyHandle *a, double *out, size_t len);
void rustyfem_assembly_free(AssemblyHandle *a);

#ifdef __cplusplus
}
#endif
#endif
