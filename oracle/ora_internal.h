/* Internal structures shared by the oracle's translation units.
 * Test infrastructure only (see ora.h). */
#ifndef ORA_INTERNAL_H
#define ORA_INTERNAL_H
#include "ora.h"
#include "ora_util.h"

/* One node of the dissection tree (src/sparse.rs:18-28). */
typedef struct {
  vi nodes;     /* mesh nodes of the separator / leaf, in labelling order */
  vi children;  /* tree indices, at most two */
  int64_t start, k;
  vi rows;      /* later equations of the front, sorted */
  vi unodes;    /* row nodes, sorted by first_dof */
  double work;  /* flops of the whole subtree */
  int64_t parent;
} ora_tnode;

struct ora_sym {
  ora_tnode *tree;
  size_t ntree, cap;
  vi roots;
  int64_t *first_dof;
  size_t nn, n;
};

/* src/sparse.rs:59-76 */
double ora_dot(const double *a, const double *b, size_t n);

/* src/sparse.rs:289-303, 305-385 (FMA variants) */
void ora_kern_forward(double *w, const double *lt, size_t n);
void ora_kern_gemm(double *c, size_t ldc, size_t nrows, size_t ncols, const double *a, size_t lda, const double *b,
                   size_t ldb, size_t kk);
void ora_kern_syrk(double *chunk, size_t ldc, size_t r0, size_t nr, size_t coloff, size_t rows, size_t kk,
                   const double *w, const double *lt);
/* row-block parallel syrk over f[b1*m..] as src/sparse.rs:569-573 / indef.rs:527-531 */
void ora_par_syrk(double *fb, size_t m, size_t rows, size_t b1, size_t kk, const double *w, const double *lt,
                  int threads);
/* row-block parallel general update (src/sparse/lu.rs:337-353) */
void ora_par_gemm_rows(double *fb, size_t m, size_t rows, size_t b1, size_t npb, const double *a, const double *b,
                       int threads);

#define ORA_PAR_FLOPS 4e6
#endif
