/*
 * CPU oracle: a plain-C FP64 restatement of RustyFEM's sparse solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product library never links it.
 *
 * Every function names the reference file:line it restates.  The arithmetic
 * follows the reference's operation order: `dot` keeps the 8-lane reduction
 * tree of src/sparse.rs:59-76, the dense kernels use fma() exactly where
 * src/sparse.rs:281-287 uses mul_add (the AVX2+FMA path), and the whole
 * library is compiled with -ffp-contract=off so that no other contraction
 * happens.
 *
 * The reference sorts node sets by coordinate with Rust's unstable sort
 * (src/sparse.rs:108), whose order among equal coordinates decides the labels
 * inside a supernode.  ora_rsort.c restates that sort (ipnsort of Rust's core
 * library, Rust >= 1.81) from its published algorithm; the restatement is
 * pinned by properties and by the algorithm's fixed cases
 * (tests/test_oracle_rsort.py), not by a run of the Rust binary, which cannot
 * be built here.
 */
#ifndef RUSTYFEM_ORACLE_H
#define RUSTYFEM_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- symbolic analysis (src/sparse.rs:18-36, 161-275) ---- */
typedef struct ora_sym ora_sym;

/* coords: nn x 3; adjacency in CSR form (adj_ptr nn+1, adj_idx); ndof per node. */
ora_sym *ora_analyse(size_t nn, const double *coords, const int64_t *adj_ptr, const int64_t *adj_idx,
                     const int64_t *ndof, size_t leaf_max);
void ora_sym_free(ora_sym *s);
size_t ora_sym_n(const ora_sym *s);
size_t ora_sym_nfronts(const ora_sym *s);
double ora_sym_work(const ora_sym *s);
size_t ora_sym_factor_size(const ora_sym *s);
/* per front in postorder: start, k, u, number of children, parent (-1 root) */
void ora_sym_fronts(const ora_sym *s, int64_t *start, int64_t *k, int64_t *u, int64_t *nchild, int64_t *parent);
/* row list of front f (length u) */
void ora_sym_rows(const ora_sym *s, size_t f, int64_t *rows);
void ora_sym_first_dof(const ora_sym *s, int64_t *first_dof);
/* node list of front f in labelling order; returns count */
size_t ora_sym_nodes(const ora_sym *s, size_t f, int64_t *nodes);
/* lower pattern (src/sparse.rs:256-275): returns nnz, fills colptr (n+1) when non-null, rowidx (nnz) when non-null */
size_t ora_sym_pattern(const ora_sym *s, const int64_t *adj_ptr, const int64_t *adj_idx, const int64_t *ndof,
                       int64_t *colptr, int64_t *rowidx);

/* ---- PD multifrontal LDL^T (src/sparse.rs:445-653) ---- */
typedef struct ora_factor ora_factor;
/* lower CSC in internal numbering; returns NULL and *err = failing equation on a bad pivot */
ora_factor *ora_factor_pd(const ora_sym *s, const int64_t *colptr, const int64_t *rowidx, const double *val,
                          int64_t *err);
void ora_factor_free(ora_factor *f);
void ora_factor_pivots(const ora_factor *f, double *d);
void ora_factor_solve(const ora_factor *f, double *x);
/* dense L of front f: l11 packed (k(k-1)/2), l21 (u*k row-major) */
void ora_factor_front(const ora_factor *f, size_t fr, double *l11, double *d, double *l21);
/* update matrix that front fr passes to its parent (u*u row-major lower); recomputed on request */

/* ---- indefinite LDL^T (src/sparse/indef.rs) ---- */
typedef struct ora_ifactor ora_ifactor;
ora_ifactor *ora_factor_indef(const ora_sym *s, const int64_t *colptr, const int64_t *rowidx, const double *val,
                              double threshold, int64_t *err);
void ora_ifactor_free(ora_ifactor *f);
void ora_ifactor_solve(const ora_ifactor *f, double *x);
/* out: neg, zero, pos, delayed, two_by_two, factor_size */
void ora_ifactor_stats(const ora_ifactor *f, int64_t *out6);
/* pivot sequence in front postorder: equation, d, e (2x2 off-diagonal on the
   first of a pair), kind (1 = 1x1, 2 = first of a 2x2, 0 = second); returns the count */
/* candidates the roots left to Bunch-Kaufman in the last factorization (test probe) */
int64_t ora_indef_last_forced(void);
size_t ora_ifactor_pivots(const ora_ifactor *f, int64_t *eq, double *d, double *e, int32_t *kind);

/* ---- multifrontal LU (src/sparse/lu.rs) ---- */
typedef struct ora_lufactor ora_lufactor;
/* lower values and strictly-upper-transposed values on the same lower pattern */
ora_lufactor *ora_factor_lu(const ora_sym *s, const int64_t *colptr, const int64_t *rowidx, const double *lower,
                            const double *upper_t, double threshold, int64_t *err);
void ora_lufactor_free(ora_lufactor *f);
void ora_lufactor_solve(const ora_lufactor *f, double *x);
/* out: delayed, factor_size */
void ora_lufactor_stats(const ora_lufactor *f, int64_t *out2);
/* pivot sequence in front postorder: equation, U diagonal, kind 1 (threshold
   pivot) or 3 (partial pivoting at a root; e = the row permutation entry) */
size_t ora_lufactor_pivots(const ora_lufactor *f, int64_t *eq, double *d, double *e, int32_t *kind);

/* ---- the C ABI of ffi/src/lib.rs, restated with an ora_ prefix ---- */
typedef struct ora_handle ora_handle;
ora_handle *ora_ffi_factor(size_t n, size_t nn, const int64_t *loc, const double *coords, const int64_t *indptr,
                           const int64_t *indices, const double *data, double *d_out, int64_t *err_dof);
ora_handle *ora_ffi_factor_indef(size_t n, size_t nn, const int64_t *loc, const double *coords,
                                 const int64_t *indptr, const int64_t *indices, const double *data,
                                 int64_t *inertia_out, int64_t *err_dof);
ora_handle *ora_ffi_factor_unsym(size_t n, size_t nn, const int64_t *loc, const double *coords,
                                 const int64_t *indptr, const int64_t *indices, const double *data,
                                 int64_t *err_dof);
int32_t ora_ffi_inertia(const ora_handle *h, int64_t *out);
int32_t ora_ffi_solve(const ora_handle *h, size_t nrhs, double *b);
void ora_ffi_info(const ora_handle *h, double *out);
void ora_ffi_order(const ora_handle *h, int64_t *out);
void ora_ffi_free(ora_handle *h);
/* pivot sequence of the handle's factor (PD: equation order, d, kind 1); returns the count */
size_t ora_ffi_pivots(const ora_handle *h, int64_t *eq, double *d, double *e, int32_t *kind);
/* neg, zero, pos, delayed, two_by_two, factor_size (PD: 0, 0, n, 0, 0, size) */
void ora_ffi_stats(const ora_handle *h, int64_t *out6);
/* the symbolic analysis a handle holds (NULL-safe), for structure checks */
const ora_sym *ora_ffi_sym(const ora_handle *h);
/* leaf size used by the C ABI (ffi/src/lib.rs:17) */
size_t ora_ffi_leaf(void);
void ora_set_ffi_leaf(size_t leaf);

/* ---- eigenvalues (src/solver/eigen.rs) ---- */
/* shift-invert block subspace iteration on a PD handle holding Op; B in CSC
 * with both triangles; returns the pairs found, -1 (no convergence) or -2 */
int64_t ora_eigen(const ora_handle *h, size_t n, const int64_t *bp, const int64_t *bi, const double *bv, size_t nev,
                  size_t vectors, size_t max_iter, double tol, double *theta_out, double *x_out,
                  int64_t *iterations);
/* cyclic Jacobi of a dense symmetric n x n matrix (row-major); v[i n + j] =
 * component i of eigenvector j */
void ora_jacobi_eigen(size_t n, const double *a, double *d, double *v);

/* ---- element stiffness and assembly (ora_elem.c) ---- */
/* etype 1 C3D8 (B-bar), 2 C3D4, 3 CPS4, 4 CPE4; x: nodes x 3; ke: nd x nd
 * row-major (src/elements/).  0, -1 (distorted), -2 (unknown type) */
int32_t ora_element_stiffness(int32_t etype, const double *x, double young, double nu, double thickness,
                              double *ke);
/* The element loop and scatter of src/fem/assembly.rs:378-398 into a given
 * both-triangle CSC pattern (eq: nn x dofs per node, -1 = no equation).
 * Returns -1, the first distorted element, -2 (entry missing) or -3. */
int64_t ora_assemble(int32_t etype, int64_t nelem, const int64_t *conn, int64_t nn, const double *coords,
                     const int64_t *eq, int64_t n, const double *young, const double *nu, const double *thickness,
                     const int64_t *indptr, const int64_t *indices, double *data);

/* ---- iterative path (src/solver/iterative*.rs) ---- */
/* kind: 0 AMG, 1 IC(0), 2 Jacobi.  Returns iterations; rel (relative recursive residual), converged. */
int64_t ora_solve_lower(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val,
                        const double *b, int kind, size_t nn, const double *coords, const int64_t *ndof,
                        double tol, int64_t maxit, double *x, double *rel, int32_t *converged,
                        int64_t *amg_sizes /* 16 or NULL */, double *complexity);
/* IC(0) factor (src/solver/iterative/csr.rs:229-257): strictly lower L
 * row-wise and D, shift of the last attempt; returns nnz(L) (outputs may be NULL) */
int64_t ora_ic0_factor(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val, int64_t *ptr,
                       int64_t *idx, double *lval, double *d, double *shift);
/* z = M r, kind 1 IC(0) or 2 Jacobi (src/solver/iterative/csr.rs:193-217, 296-316) */
void ora_precond_apply(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val, int kind,
                       const double *r, double *z);
/* natural pattern (src/solver/iterative.rs:74-97): returns nnz; first (nn) */
size_t ora_natural_pattern(size_t nn, const int64_t *adj_ptr, const int64_t *adj_idx, const int64_t *ndof,
                           int64_t *first, int64_t *colptr, int64_t *rowidx);
/* y = A x for a lower triangle (src/sparse/indef.rs:600-611) */
void ora_sym_matvec(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val, const double *x,
                    double *y);
/* rigid body modes (src/solver/iterative/amg.rs:39-95): returns nb, b is n*6 capacity */
size_t ora_rigid_body_modes(size_t nn, const double *coords, const int64_t *ndof, double *b);

/* threads used by the CPU-parallel parts (default hardware concurrency) */
void ora_set_threads(int t);
int ora_threads(void);

/* ---- AMG hierarchy and aggregation hooks (ora_iter.c) ---- */
typedef struct ora_amg ora_amg;
int64_t ora_amg_aggregate(size_t n, const int64_t *ptr, const int64_t *idx, const double *val, size_t nblk,
                          const int64_t *bptr, double theta, int64_t *agg_out);
ora_amg *ora_amg_build(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val, size_t nn,
                       const double *coords, const int64_t *ndof);
size_t ora_amg_nlev(const ora_amg *h);
int32_t ora_amg_level(const ora_amg *h, size_t l, int32_t what, int64_t *iout, int64_t *idx, double *dout);
void ora_amg_free(ora_amg *h);

/* ---- test hooks of the restated Rust unstable sort (ora_rsort.c) ---- */
/* items sorted by ikey[item] (ikey NULL: by the item value) */
void ora_rsort_by_ikey(int64_t *items, size_t n, const int64_t *key);
/* the 9- and 13-input sorting networks of its small sort, alone */
void ora_rsort_net9(int64_t *v, const int64_t *ikey);
void ora_rsort_net13(int64_t *v, const int64_t *ikey);

#ifdef __cplusplus
}
#endif
#endif
