/*
 * Oracle: nested-dissection symbolic analysis and the positive definite
 * multifrontal LDL^T with its solve.  Restates src/sparse.rs.
 * Test infrastructure only (see ora.h).
 */
#include "ora_internal.h"
#include <math.h>

/* ------------------------------------------------------------------ */
/* dot with 8 partial sums and a fixed reduction tree (src/sparse.rs:59-76) */
double ora_dot(const double *a, const double *b, size_t n) {
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  size_t n8 = n / 8 * 8;
  for (size_t i = 0; i < n8; i += 8)
    for (int l = 0; l < 8; l++) s[l] += a[i + l] * b[i + l];
  double t = (s[0] + s[4]) + (s[1] + s[5]) + ((s[2] + s[6]) + (s[3] + s[7]));
  for (size_t i = n8; i < n; i++) t += a[i] * b[i];
  return t;
}

/* ------------------------------------------------------------------ */
/* dense kernels, FMA variant (src/sparse.rs:279-443) */

/* w <- w L^-T for one row; lt is the transposed unit triangle, n x n (src/sparse.rs:291-303) */
void ora_kern_forward(double *w, const double *lt, size_t n) {
  for (size_t p = 0; p < n; p++) {
    double wp = -w[p];
    if (wp == 0.0) continue;
    const double *col = lt + p * n;
    for (size_t q = p + 1; q < n; q++) w[q] = fma(wp, col[q], w[q]);
  }
}

/* C -= A B with a 4x8 register block and 512-wide column blocks (src/sparse.rs:309-369).
 * Every element gets s = fma-chain over p from 0, then c -= s, in every branch. */
void ora_kern_gemm(double *c, size_t ldc, size_t nrows, size_t ncols, const double *a, size_t lda, const double *b,
                   size_t ldb, size_t kk) {
  const size_t YB = 512;
  size_t j0 = 0;
  while (j0 < ncols) {
    size_t j1 = j0 + YB < ncols ? j0 + YB : ncols;
    size_t i = 0;
    while (i + 4 <= nrows) {
      size_t j = j0;
      while (j + 8 <= j1) {
        double acc[4][8];
        for (int r = 0; r < 4; r++)
          for (int l = 0; l < 8; l++) acc[r][l] = 0.0;
        for (size_t p = 0; p < kk; p++) {
          const double *bp = b + p * ldb + j;
          for (int r = 0; r < 4; r++) {
            double ar = a[(i + r) * lda + p];
            for (int l = 0; l < 8; l++) acc[r][l] = fma(ar, bp[l], acc[r][l]);
          }
        }
        for (int r = 0; r < 4; r++) {
          double *cr = c + (i + r) * ldc + j;
          for (int l = 0; l < 8; l++) cr[l] -= acc[r][l];
        }
        j += 8;
      }
      for (int r = 0; r < 4; r++)
        for (size_t jj = j; jj < j1; jj++) {
          double s = 0.0;
          for (size_t p = 0; p < kk; p++) s = fma(a[(i + r) * lda + p], b[p * ldb + jj], s);
          c[(i + r) * ldc + jj] -= s;
        }
      i += 4;
    }
    while (i < nrows) {
      for (size_t jj = j0; jj < j1; jj++) {
        double s = 0.0;
        for (size_t p = 0; p < kk; p++) s = fma(a[i * lda + p], b[p * ldb + jj], s);
        c[i * ldc + jj] -= s;
      }
      i++;
    }
    j0 = j1;
  }
}

/* lower-triangle update of rows r0..r0+nr in a row block (src/sparse.rs:375-385) */
void ora_kern_syrk(double *chunk, size_t ldc, size_t r0, size_t nr, size_t coloff, size_t rows, size_t kk,
                   const double *w, const double *lt) {
  const size_t YB = 512;
  size_t xend = r0 + nr;
  size_t yb0 = 0;
  while (yb0 < xend) {
    size_t yb1 = yb0 + YB < xend ? yb0 + YB : xend;
    size_t rs = r0 > yb0 ? r0 : yb0;
    ora_kern_gemm(chunk + (rs - r0) * ldc + coloff + yb0, ldc, xend - rs, yb1 - yb0, w + rs * kk, kk, lt + yb0,
                  rows, kk);
    yb0 = yb1;
  }
}

typedef struct {
  double *fb;
  size_t m, rows, b1, kk, r0, nr;
  const double *w, *lt;
} syrk_task;

static void syrk_task_run(void *p) {
  syrk_task *t = (syrk_task *)p;
  ora_kern_syrk(t->fb + t->r0 * t->m, t->m, t->r0, t->nr, t->b1, t->rows, t->kk, t->w, t->lt);
}

/* par::blocks_mut with cost r+1 around kern::syrk (src/sparse.rs:569-573) */
void ora_par_syrk(double *fb, size_t m, size_t rows, size_t b1, size_t kk, const double *w, const double *lt,
                  int threads) {
  int tp = ((double)(rows * rows * kk) / 2.0 > ORA_PAR_FLOPS) ? threads : 1;
  size_t *bounds = (size_t *)ora_xmalloc(sizeof(size_t) * ((size_t)tp + 2));
  size_t nb = ora_cost_bounds(rows, tp, 1, bounds);
  int ntask = (int)nb - 1;
  syrk_task *tasks = (syrk_task *)ora_xmalloc(sizeof(syrk_task) * (size_t)(ntask > 0 ? ntask : 1));
  void **args = (void **)ora_xmalloc(sizeof(void *) * (size_t)(ntask > 0 ? ntask : 1));
  for (int i = 0; i < ntask; i++) {
    syrk_task t = {fb, m, rows, b1, kk, bounds[i], bounds[i + 1] - bounds[i], w, lt};
    tasks[i] = t;
    args[i] = &tasks[i];
  }
  ora_run_parallel(syrk_task_run, args, ntask);
  free(bounds);
  free(tasks);
  free(args);
}

typedef struct {
  double *fb;
  size_t m, rows, b1, npb, r0, nr;
  const double *a, *b;
} gemm_task;

static void gemm_task_run(void *p) {
  gemm_task *t = (gemm_task *)p;
  ora_kern_gemm(t->fb + t->r0 * t->m + t->b1, t->m, t->nr, t->rows, t->a + t->r0 * t->npb, t->npb, t->b, t->rows,
                t->npb);
}

/* src/sparse/lu.rs:348-353: par::blocks_mut with cost 1 around kern::gemm */
void ora_par_gemm_rows(double *fb, size_t m, size_t rows, size_t b1, size_t npb, const double *a, const double *b,
                       int threads) {
  int tp = ((double)(rows * rows * npb) > ORA_PAR_FLOPS) ? threads : 1;
  size_t *bounds = (size_t *)ora_xmalloc(sizeof(size_t) * ((size_t)tp + 2));
  size_t nb = ora_cost_bounds(rows, tp, 0, bounds);
  int ntask = (int)nb - 1;
  gemm_task *tasks = (gemm_task *)ora_xmalloc(sizeof(gemm_task) * (size_t)(ntask > 0 ? ntask : 1));
  void **args = (void **)ora_xmalloc(sizeof(void *) * (size_t)(ntask > 0 ? ntask : 1));
  for (int i = 0; i < ntask; i++) {
    gemm_task t = {fb, m, rows, b1, npb, bounds[i], bounds[i + 1] - bounds[i], a, b};
    tasks[i] = t;
    args[i] = &tasks[i];
  }
  ora_run_parallel(gemm_task_run, args, ntask);
  free(bounds);
  free(tasks);
  free(args);
}

/* ------------------------------------------------------------------ */
/* nested dissection (src/sparse.rs:78-157) */

typedef struct {
  ora_sym *S;
  const double *coords;
  const int64_t *aptr, *aidx;
  uint8_t *side;
  size_t leaf_max;
} dctx;

static int64_t push_tnode(ora_sym *S, const int64_t *nodes, size_t nnodes, const int64_t *children, size_t nch) {
  if (S->ntree == S->cap) {
    S->cap = S->cap ? 2 * S->cap : 64;
    S->tree = (ora_tnode *)ora_xrealloc(S->tree, S->cap * sizeof(ora_tnode));
  }
  ora_tnode *t = &S->tree[S->ntree];
  memset(t, 0, sizeof(*t));
  for (size_t i = 0; i < nnodes; i++) vi_push(&t->nodes, nodes[i]);
  for (size_t i = 0; i < nch; i++) vi_push(&t->children, children[i]);
  t->parent = -1;
  return (int64_t)S->ntree++;
}

static int any_side(const dctx *c, int64_t v, uint8_t s) {
  for (int64_t p = c->aptr[v]; p < c->aptr[v + 1]; p++)
    if (c->side[c->aidx[p]] == s) return 1;
  return 0;
}

/* `set` is owned (freed) by this call. Returns the tree index or -1 for an empty set. */
static int64_t dissect(dctx *c, int64_t *set, size_t n) {
  if (n == 0) {
    free(set);
    return -1;
  }
  if (n <= c->leaf_max) {
    int64_t id = push_tnode(c->S, set, n, NULL, 0);
    free(set);
    return id;
  }
  const double *X = c->coords;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) {
      double x = X[set[i] * 3 + a];
      lo[a] = fmin(lo[a], x);
      hi[a] = fmax(hi[a], x);
    }
  /* longest axis first, stable on ties (src/sparse.rs:100-101) */
  int axes[3] = {0, 1, 2};
  for (int i = 1; i < 3; i++) {
    int x = axes[i];
    int j = i;
    while (j > 0 && (hi[axes[j - 1]] - lo[axes[j - 1]]) < (hi[x] - lo[x])) {
      axes[j] = axes[j - 1];
      j--;
    }
    axes[j] = x;
  }
  int have_best = 0, best_ax = 0, best_in_left = 1;
  double best_score = 0.0, best_t = 0.0;
  for (int ai = 0; ai < 3; ai++) {
    int ax = axes[ai];
    if (!(hi[ax] > lo[ax])) continue;
    ora_rsort_by_dkey(set, n, X, 3, (size_t)ax); /* sort_unstable_by, src/sparse.rs:108 */
    size_t qs[3] = {n / 2, 2 * n / 5, 3 * n / 5};
    for (int qi = 0; qi < 3; qi++) {
      double t = X[set[qs[qi]] * 3 + ax];
      /* partition_point of coord < t on the sorted set */
      size_t l = 0, h = n;
      while (l < h) {
        size_t mid = (l + h) / 2;
        if (X[set[mid] * 3 + ax] < t) l = mid + 1; else h = mid;
      }
      size_t nl = l;
      if (nl == 0 || nl == n) continue;
      for (size_t i = 0; i < n; i++) c->side[set[i]] = (uint8_t)(i >= nl);
      size_t sep_l = 0, sep_r = 0;
      for (size_t i = 0; i < nl; i++) sep_l += (size_t)any_side(c, set[i], 1);
      for (size_t i = nl; i < n; i++) sep_r += (size_t)any_side(c, set[i], 0);
      size_t s;
      int in_left;
      if (sep_r < sep_l) { s = sep_r; in_left = 0; } else { s = sep_l; in_left = 1; }
      size_t mn = nl < n - nl ? nl : n - nl;
      double score = (double)s / ((double)mn / (double)n);
      if (!have_best || score < best_score) {
        have_best = 1;
        best_score = score;
        best_ax = ax;
        best_t = t;
        best_in_left = in_left;
      }
    }
  }
  int64_t *left = (int64_t *)ora_xmalloc(n * sizeof(int64_t));
  int64_t *right = (int64_t *)ora_xmalloc(n * sizeof(int64_t));
  size_t nl = 0, nr = 0;
  if (have_best) {
    for (size_t i = 0; i < n; i++) {
      if (X[set[i] * 3 + best_ax] < best_t) left[nl++] = set[i]; else right[nr++] = set[i];
    }
  } else { /* all nodes at the same point: split on index */
    for (size_t i = 0; i < n; i++) {
      if (i < n / 2) left[nl++] = set[i]; else right[nr++] = set[i];
    }
  }
  free(set);
  int in_left = have_best ? best_in_left : 1;
  for (size_t i = 0; i < nl; i++) c->side[left[i]] = 0;
  for (size_t i = 0; i < nr; i++) c->side[right[i]] = 1;
  int64_t *sep = (int64_t *)ora_xmalloc(n * sizeof(int64_t));
  int64_t *a = (int64_t *)ora_xmalloc(n * sizeof(int64_t));
  int64_t *b = (int64_t *)ora_xmalloc(n * sizeof(int64_t));
  size_t ns = 0, na = 0, nb = 0;
  if (in_left) {
    for (size_t i = 0; i < nl; i++) {
      if (any_side(c, left[i], 1)) sep[ns++] = left[i]; else a[na++] = left[i];
    }
    for (size_t i = 0; i < nr; i++) b[nb++] = right[i];
  } else {
    for (size_t i = 0; i < nr; i++) {
      if (any_side(c, right[i], 0)) sep[ns++] = right[i]; else b[nb++] = right[i];
    }
    for (size_t i = 0; i < nl; i++) a[na++] = left[i];
  }
  for (size_t i = 0; i < nl; i++) c->side[left[i]] = 2;
  for (size_t i = 0; i < nr; i++) c->side[right[i]] = 2;
  free(left);
  free(right);
  int64_t ch[2];
  size_t nch = 0;
  int64_t ca = dissect(c, a, na);
  int64_t cb = dissect(c, b, nb);
  if (ca >= 0) ch[nch++] = ca;
  if (cb >= 0) ch[nch++] = cb;
  int64_t id = push_tnode(c->S, sep, ns, ch, nch);
  free(sep);
  return id;
}

/* src/sparse.rs:167-232 */
ora_sym *ora_analyse(size_t nn, const double *coords, const int64_t *adj_ptr, const int64_t *adj_idx,
                     const int64_t *ndof, size_t leaf_max) {
  if (leaf_max < 1) leaf_max = 1;
  ora_sym *S = (ora_sym *)ora_xcalloc(1, sizeof(ora_sym));
  S->nn = nn;
  uint8_t *side = (uint8_t *)ora_xmalloc(nn ? nn : 1);
  memset(side, 2, nn);
  dctx c = {S, coords, adj_ptr, adj_idx, side, leaf_max};
  int64_t *set = (int64_t *)ora_xmalloc(nn * sizeof(int64_t));
  for (size_t i = 0; i < nn; i++) set[i] = (int64_t)i;
  int64_t r = dissect(&c, set, nn);
  free(side);
  if (r >= 0) vi_push(&S->roots, r);
  /* postorder with an explicit stack (children before parent, child 0 first) */
  int64_t *post = (int64_t *)ora_xmalloc((S->ntree + 1) * sizeof(int64_t));
  size_t npost = 0;
  vi stk = {0}, dn = {0};
  for (size_t i = 0; i < S->roots.n; i++) { vi_push(&stk, S->roots.a[i]); vi_push(&dn, 0); }
  while (stk.n) {
    int64_t t = stk.a[--stk.n];
    int64_t done = dn.a[--dn.n];
    if (done) {
      post[npost++] = t;
    } else {
      vi_push(&stk, t);
      vi_push(&dn, 1);
      ora_tnode *tn = &S->tree[t];
      for (size_t j = tn->children.n; j-- > 0;) { vi_push(&stk, tn->children.a[j]); vi_push(&dn, 0); }
    }
  }
  vi_free(&stk);
  vi_free(&dn);
  /* The construction order of `dissect` is already a postorder; the fronts are
   * addressed by tree index, which the check below confirms. */
  for (size_t i = 0; i < npost; i++)
    if (post[i] != (int64_t)i) { fprintf(stderr, "oracle: tree order is not a postorder\n"); abort(); }
  S->first_dof = (int64_t *)ora_xcalloc(nn ? nn : 1, sizeof(int64_t));
  int64_t next = 0;
  for (size_t pi = 0; pi < npost; pi++) {
    ora_tnode *t = &S->tree[post[pi]];
    t->start = next;
    for (size_t i = 0; i < t->nodes.n; i++) {
      int64_t v = t->nodes.a[i];
      S->first_dof[v] = next;
      next += ndof[v];
    }
    t->k = next - t->start;
    for (size_t j = 0; j < t->children.n; j++) S->tree[t->children.a[j]].parent = post[pi];
  }
  S->n = (size_t)next;
  /* front pattern at node level (src/sparse.rs:201-230) */
  int64_t *stamp = (int64_t *)ora_xmalloc((nn ? nn : 1) * sizeof(int64_t));
  for (size_t i = 0; i < nn; i++) stamp[i] = -1;
  for (size_t pi = 0; pi < npost; pi++) {
    int64_t ti = post[pi];
    ora_tnode *t = &S->tree[ti];
    int64_t end = t->start + t->k;
    vi list = {0};
    for (size_t i = 0; i < t->nodes.n; i++) {
      int64_t v = t->nodes.a[i];
      for (int64_t p = adj_ptr[v]; p < adj_ptr[v + 1]; p++) {
        int64_t w = adj_idx[p];
        if (S->first_dof[w] >= end && stamp[w] != ti) { stamp[w] = ti; vi_push(&list, w); }
      }
    }
    for (size_t j = 0; j < t->children.n; j++) {
      ora_tnode *ch = &S->tree[t->children.a[j]];
      for (size_t i = 0; i < ch->unodes.n; i++) {
        int64_t w = ch->unodes.a[i];
        if (S->first_dof[w] >= end && stamp[w] != ti) { stamp[w] = ti; vi_push(&list, w); }
      }
    }
    ora_rsort_by_ikey(list.a, list.n, S->first_dof); /* sort_unstable_by_key, src/sparse.rs:224 */
    for (size_t i = 0; i < list.n; i++) {
      int64_t w = list.a[i];
      for (int64_t d = 0; d < ndof[w]; d++) vi_push(&t->rows, S->first_dof[w] + d);
    }
    double k = (double)t->k, u = (double)t->rows.n;
    double own = k * k * k / 3.0 + u * k * k + u * u * k;
    double cs = 0.0;
    for (size_t j = 0; j < t->children.n; j++) cs += S->tree[t->children.a[j]].work;
    t->work = own + cs;
    t->unodes = list;
  }
  free(stamp);
  free(post);
  return S;
}

void ora_sym_free(ora_sym *S) {
  if (!S) return;
  for (size_t i = 0; i < S->ntree; i++) {
    vi_free(&S->tree[i].nodes);
    vi_free(&S->tree[i].children);
    vi_free(&S->tree[i].rows);
    vi_free(&S->tree[i].unodes);
  }
  free(S->tree);
  vi_free(&S->roots);
  free(S->first_dof);
  free(S);
}

size_t ora_sym_n(const ora_sym *s) { return s->n; }
size_t ora_sym_nfronts(const ora_sym *s) { return s->ntree; }

double ora_sym_work(const ora_sym *s) {
  double w = 0.0;
  for (size_t i = 0; i < s->roots.n; i++) w += s->tree[s->roots.a[i]].work;
  return w;
}

size_t ora_sym_factor_size(const ora_sym *s) {
  size_t tot = 0;
  for (size_t i = 0; i < s->ntree; i++) {
    size_t k = (size_t)s->tree[i].k, u = s->tree[i].rows.n;
    tot += k * (k ? k - 1 : 0) / 2 + k * u;
  }
  return tot;
}

void ora_sym_fronts(const ora_sym *s, int64_t *start, int64_t *k, int64_t *u, int64_t *nchild, int64_t *parent) {
  for (size_t i = 0; i < s->ntree; i++) {
    if (start) start[i] = s->tree[i].start;
    if (k) k[i] = s->tree[i].k;
    if (u) u[i] = (int64_t)s->tree[i].rows.n;
    if (nchild) nchild[i] = (int64_t)s->tree[i].children.n;
    if (parent) parent[i] = s->tree[i].parent;
  }
}

void ora_sym_rows(const ora_sym *s, size_t f, int64_t *rows) {
  memcpy(rows, s->tree[f].rows.a, s->tree[f].rows.n * sizeof(int64_t));
}

void ora_sym_first_dof(const ora_sym *s, int64_t *fd) { memcpy(fd, s->first_dof, s->nn * sizeof(int64_t)); }

size_t ora_sym_nodes(const ora_sym *s, size_t f, int64_t *nodes) {
  if (nodes) memcpy(nodes, s->tree[f].nodes.a, s->tree[f].nodes.n * sizeof(int64_t));
  return s->tree[f].nodes.n;
}

/* src/sparse.rs:256-275 */
size_t ora_sym_pattern(const ora_sym *s, const int64_t *adj_ptr, const int64_t *adj_idx, const int64_t *ndof,
                       int64_t *colptr, int64_t *rowidx) {
  size_t nn = s->nn;
  int64_t *order = (int64_t *)ora_xmalloc((nn ? nn : 1) * sizeof(int64_t));
  for (size_t i = 0; i < nn; i++) order[i] = (int64_t)i;
  ora_rsort_by_ikey(order, nn, s->first_dof); /* src/sparse.rs:260 */
  size_t nnz = 0;
  if (colptr) colptr[0] = 0;
  vi nb = {0};
  for (size_t oi = 0; oi < nn; oi++) {
    int64_t v = order[oi];
    nb.n = 0;
    for (int64_t p = adj_ptr[v]; p < adj_ptr[v + 1]; p++) vi_push(&nb, adj_idx[p]);
    vi_push(&nb, v);
    ora_rsort_by_ikey(nb.a, nb.n, s->first_dof); /* src/sparse.rs:263 */
    int64_t f = s->first_dof[v];
    for (int64_t c = f; c < f + ndof[v]; c++) {
      for (size_t j = 0; j < nb.n; j++) {
        int64_t w = nb.a[j], fw = s->first_dof[w];
        for (int64_t r = fw; r < fw + ndof[w]; r++)
          if (r >= c) {
            if (rowidx) rowidx[nnz] = r;
            nnz++;
          }
      }
      if (colptr) colptr[c + 1] = (int64_t)nnz;
    }
  }
  vi_free(&nb);
  free(order);
  return nnz;
}

/* ------------------------------------------------------------------ */
/* PD multifrontal LDL^T (src/sparse.rs:445-602) */

typedef struct {
  int64_t start, k, u;
  double *l11, *d, *l21;
} pd_front;

struct ora_factor {
  const ora_sym *S;
  size_t nf;
  pd_front *fr;
};

typedef struct {
  const ora_sym *S;
  const int64_t *colptr, *rowidx;
  const double *val;
  pd_front *fr;
} pd_ctx;

#define PD_KB 96

typedef struct {
  pd_ctx *c;
  int64_t t;
  int threads;
  double *upd;
  int64_t err; /* -1 ok */
} pd_job;

static void pd_node(pd_job *job);
static void pd_job_run(void *p) { pd_node((pd_job *)p); }

typedef struct {
  double *fb;
  size_t m, j0, jb, r0, nr;
  const double *lbt, *dj;
} panel_task;

static void panel_run(void *p) {
  panel_task *t = (panel_task *)p;
  for (size_t r = t->r0; r < t->r0 + t->nr; r++) {
    double *wr = t->fb + r * t->m + t->j0;
    ora_kern_forward(wr, t->lbt, t->jb);
    for (size_t q = 0; q < t->jb; q++) wr[q] /= t->dj[q];
  }
}

static void pd_node(pd_job *job) {
  pd_ctx *c = job->c;
  const ora_sym *S = c->S;
  const ora_tnode *node = &S->tree[job->t];
  size_t nch = node->children.n;
  pd_job cj[2];
  for (size_t i = 0; i < nch; i++) {
    cj[i].c = c;
    cj[i].t = node->children.a[i];
    cj[i].upd = NULL;
    cj[i].err = -1;
  }
  if (nch == 2 && job->threads > 1 && fmin(S->tree[cj[0].t].work, S->tree[cj[1].t].work) > 2e7) {
    int half = job->threads / 2;
    cj[0].threads = half;
    cj[1].threads = job->threads - half;
    void *args[2] = {&cj[1], &cj[0]};
    ora_run_parallel(pd_job_run, args, 2);
  } else {
    for (size_t i = 0; i < nch; i++) {
      cj[i].threads = job->threads;
      pd_node(&cj[i]);
    }
  }
  job->upd = NULL;
  job->err = -1;
  for (size_t i = 0; i < nch; i++)
    if (cj[i].err >= 0) {
      job->err = cj[i].err;
      for (size_t j = 0; j < nch; j++) free(cj[j].upd);
      return;
    }
  size_t start = (size_t)node->start, k = (size_t)node->k, u = node->rows.n, m = k + u;
  const int64_t *rows = node->rows.a;
#define LOCAL(dof) (((size_t)(dof) >= start && (size_t)(dof) < start + k) ? (size_t)(dof) - start \
                    : k + (size_t)ora_bsearch(rows, u, (dof)))
  double *f = (double *)ora_xcalloc(m * m, sizeof(double));
  for (size_t cc = 0; cc < k; cc++) {
    size_t col = start + cc;
    for (int64_t p = c->colptr[col]; p < c->colptr[col + 1]; p++) f[LOCAL(c->rowidx[p]) * m + cc] += c->val[p];
  }
  for (size_t i = 0; i < nch; i++) {
    const ora_tnode *ch = &S->tree[cj[i].t];
    size_t uc = ch->rows.n;
    size_t *map = (size_t *)ora_xmalloc((uc ? uc : 1) * sizeof(size_t));
    for (size_t x = 0; x < uc; x++) map[x] = LOCAL(ch->rows.a[x]);
    const double *upd = cj[i].upd;
    for (size_t x = 0; x < uc; x++) {
      size_t fx = map[x] * m;
      for (size_t y = 0; y <= x; y++) f[fx + map[y]] += upd[x * uc + y];
    }
    free(map);
    free(cj[i].upd);
  }
#undef LOCAL
  double *d = (double *)ora_xcalloc(k ? k : 1, sizeof(double));
  size_t j0 = 0;
  while (j0 < k) {
    size_t j1 = j0 + PD_KB < k ? j0 + PD_KB : k;
    size_t jb = j1 - j0;
    double *lb = (double *)ora_xcalloc(jb * jb, sizeof(double));
    double *w = (double *)ora_xcalloc(jb, sizeof(double));
    for (size_t i = 0; i < jb; i++) {
      size_t row = (j0 + i) * m + j0;
      for (size_t j = 0; j < i; j++) w[j] = f[row + j] - ora_dot(w, lb + j * jb, j);
      double di = f[row + i];
      for (size_t p = 0; p < i; p++) di -= w[p] * w[p] / d[j0 + p];
      double orig = c->val[c->colptr[start + j0 + i]];
      if (!(orig > 0.0) || !(di > 1e-12 * orig)) {
        job->err = (int64_t)(start + j0 + i);
        free(lb);
        free(w);
        free(d);
        free(f);
        return;
      }
      d[j0 + i] = di;
      for (size_t p = 0; p < i; p++) {
        double l = w[p] / d[j0 + p];
        lb[i * jb + p] = l;
        f[row + p] = l;
      }
    }
    if (j1 < m) {
      size_t nrows = m - j1;
      double *lbt = (double *)ora_xcalloc(jb * jb, sizeof(double));
      for (size_t i = 0; i < jb; i++)
        for (size_t p = 0; p < i; p++) lbt[p * jb + i] = lb[i * jb + p];
      int tp = ((double)(nrows * jb * jb) > ORA_PAR_FLOPS) ? job->threads : 1;
      size_t *bounds = (size_t *)ora_xmalloc(sizeof(size_t) * ((size_t)tp + 2));
      size_t nb = ora_cost_bounds(nrows, tp, 0, bounds);
      int nt = (int)nb - 1;
      panel_task *pt = (panel_task *)ora_xmalloc(sizeof(panel_task) * (size_t)nt);
      void **args = (void **)ora_xmalloc(sizeof(void *) * (size_t)nt);
      for (int i = 0; i < nt; i++) {
        panel_task t = {f + j1 * m, m, j0, jb, bounds[i], bounds[i + 1] - bounds[i], lbt, d + j0};
        pt[i] = t;
        args[i] = &pt[i];
      }
      ora_run_parallel(panel_run, args, nt);
      free(pt);
      free(args);
      free(bounds);
      double *wt = (double *)ora_xmalloc(nrows * jb * sizeof(double));
      double *lt = (double *)ora_xmalloc(jb * nrows * sizeof(double));
      for (size_t r = 0; r < nrows; r++) {
        const double *fr = f + (j1 + r) * m + j0;
        for (size_t p = 0; p < jb; p++) {
          wt[r * jb + p] = fr[p] * d[j0 + p];
          lt[p * nrows + r] = fr[p];
        }
      }
      ora_par_syrk(f + j1 * m, m, nrows, j1, jb, wt, lt, job->threads);
      free(wt);
      free(lt);
      free(lbt);
    }
    free(lb);
    free(w);
    j0 = j1;
  }
  pd_front *F = &c->fr[job->t];
  F->start = (int64_t)start;
  F->k = (int64_t)k;
  F->u = (int64_t)u;
  F->l11 = (double *)ora_xmalloc((k * (k ? k - 1 : 0) / 2 + 1) * sizeof(double));
  size_t off = 0;
  for (size_t i = 0; i < k; i++) {
    memcpy(F->l11 + off, f + i * m, i * sizeof(double));
    off += i;
  }
  F->d = d;
  F->l21 = (double *)ora_xmalloc((u * k + 1) * sizeof(double));
  double *upd = (double *)ora_xcalloc(u * u + 1, sizeof(double));
  for (size_t r = 0; r < u; r++) {
    const double *fr = f + (k + r) * m;
    memcpy(F->l21 + r * k, fr, k * sizeof(double));
    memcpy(upd + r * u, fr + k, (r + 1) * sizeof(double));
  }
  free(f);
  job->upd = upd;
}

ora_factor *ora_factor_pd(const ora_sym *S, const int64_t *colptr, const int64_t *rowidx, const double *val,
                          int64_t *err) {
  *err = -1;
  ora_factor *F = (ora_factor *)ora_xcalloc(1, sizeof(ora_factor));
  F->S = S;
  F->nf = S->ntree;
  F->fr = (pd_front *)ora_xcalloc(S->ntree ? S->ntree : 1, sizeof(pd_front));
  pd_ctx c = {S, colptr, rowidx, val, F->fr};
  for (size_t i = 0; i < S->roots.n; i++) {
    pd_job job = {&c, S->roots.a[i], ora_threads(), NULL, -1};
    pd_node(&job);
    free(job.upd);
    if (job.err >= 0) {
      *err = job.err;
      ora_factor_free(F);
      return NULL;
    }
  }
  return F;
}

void ora_factor_free(ora_factor *F) {
  if (!F) return;
  for (size_t i = 0; i < F->nf; i++) {
    free(F->fr[i].l11);
    free(F->fr[i].d);
    free(F->fr[i].l21);
  }
  free(F->fr);
  free(F);
}

void ora_factor_pivots(const ora_factor *F, double *d) {
  memset(d, 0, F->S->n * sizeof(double));
  for (size_t i = 0; i < F->nf; i++)
    if (F->fr[i].k) memcpy(d + F->fr[i].start, F->fr[i].d, (size_t)F->fr[i].k * sizeof(double));
}

void ora_factor_front(const ora_factor *F, size_t i, double *l11, double *d, double *l21) {
  size_t k = (size_t)F->fr[i].k, u = (size_t)F->fr[i].u;
  if (l11) memcpy(l11, F->fr[i].l11, k * (k ? k - 1 : 0) / 2 * sizeof(double));
  if (d) memcpy(d, F->fr[i].d, k * sizeof(double));
  if (l21) memcpy(l21, F->fr[i].l21, u * k * sizeof(double));
}

/* src/sparse.rs:615-652 */
void ora_factor_solve(const ora_factor *F, double *x) {
  size_t nf = F->nf;
  double *xp = NULL;
  size_t xcap = 0;
  for (size_t fi = 0; fi < nf; fi++) {
    const pd_front *fr = &F->fr[fi];
    size_t k = (size_t)fr->k;
    const int64_t *rows = F->S->tree[fi].rows.a;
    double *xs = x + fr->start;
    for (size_t i = 0; i < k; i++) {
      size_t r = i * (i ? i - 1 : 0) / 2;
      double s = ora_dot(fr->l11 + r, xs, i);
      xs[i] -= s;
    }
    for (size_t r = 0; r < (size_t)fr->u; r++) x[rows[r]] -= ora_dot(fr->l21 + r * k, xs, k);
  }
  for (size_t fi = 0; fi < nf; fi++) {
    const pd_front *fr = &F->fr[fi];
    for (size_t i = 0; i < (size_t)fr->k; i++) x[fr->start + i] /= fr->d[i];
  }
  for (size_t fi = nf; fi-- > 0;) {
    const pd_front *fr = &F->fr[fi];
    size_t k = (size_t)fr->k;
    const int64_t *rows = F->S->tree[fi].rows.a;
    if (k > xcap) {
      xcap = k;
      xp = (double *)ora_xrealloc(xp, xcap * sizeof(double));
    }
    memcpy(xp, x + fr->start, k * sizeof(double));
    for (size_t r = 0; r < (size_t)fr->u; r++) {
      double xr = x[rows[r]];
      const double *l = fr->l21 + r * k;
      for (size_t p = 0; p < k; p++) xp[p] -= l[p] * xr;
    }
    for (size_t i = k; i-- > 0;) {
      double xi = xp[i];
      size_t r = i * (i ? i - 1 : 0) / 2;
      for (size_t q = 0; q < i; q++) xp[q] -= fr->l11[r + q] * xi;
    }
    memcpy(x + fr->start, xp, k * sizeof(double));
  }
  free(xp);
}
