/*
 * Oracle: the iterative path -- CSR from the lower triangle, PCG, Jacobi,
 * IC(0) and smoothed-aggregation AMG with a symmetric Gauss-Seidel V-cycle.
 * Restates src/solver/iterative.rs, src/solver/iterative/csr.rs and
 * src/solver/iterative/amg.rs.  Test infrastructure only (see ora.h).
 */
#include "ora_internal.h"
#include <math.h>

typedef struct {
  size_t nrows, ncols;
  int64_t *ptr, *idx;
  double *val;
} csr;

static void csr_free(csr *a) {
  free(a->ptr);
  free(a->idx);
  free(a->val);
  memset(a, 0, sizeof(*a));
}
static size_t csr_nnz(const csr *a) { return (size_t)a->ptr[a->nrows]; }

static csr csr_copy(const csr *a) {
  csr m;
  size_t nnz = csr_nnz(a);
  m.nrows = a->nrows;
  m.ncols = a->ncols;
  m.ptr = (int64_t *)ora_xmalloc((a->nrows + 1) * sizeof(int64_t));
  m.idx = (int64_t *)ora_xmalloc((nnz + 1) * sizeof(int64_t));
  m.val = (double *)ora_xmalloc((nnz + 1) * sizeof(double));
  memcpy(m.ptr, a->ptr, (a->nrows + 1) * sizeof(int64_t));
  memcpy(m.idx, a->idx, nnz * sizeof(int64_t));
  memcpy(m.val, a->val, nnz * sizeof(double));
  return m;
}

/* what amg_new built, level by level (for the hierarchy tests): A of levels
 * 0..nlev (nlev = the coarsest), the aggregates, P~ and P of the smoothed
 * levels, the block bounds and near-null spaces of every level */
#define ORA_MAXLEV 13
struct ora_amg {
  size_t nlev, nbm;
  csr A[ORA_MAXLEV], Pt[ORA_MAXLEV], P[ORA_MAXLEV];
  int64_t *agg[ORA_MAXLEV], *bptr[ORA_MAXLEV];
  size_t nblk[ORA_MAXLEV], nagg[ORA_MAXLEV];
  double *B[ORA_MAXLEV];
};

/* src/solver/iterative/csr.rs:20-62 */
static csr csr_from_lower(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val) {
  csr m;
  m.nrows = m.ncols = n;
  int64_t *cnt = (int64_t *)ora_xcalloc(n + 1, sizeof(int64_t));
  for (size_t c = 0; c < n; c++)
    for (int64_t p = colptr[c]; p < colptr[c + 1]; p++) {
      int64_t r = rowidx[p];
      cnt[r + 1]++;
      if ((size_t)r != c) cnt[c + 1]++;
    }
  for (size_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  size_t nnz = (size_t)cnt[n];
  int64_t *pos = (int64_t *)ora_xmalloc((n + 1) * sizeof(int64_t));
  memcpy(pos, cnt, (n + 1) * sizeof(int64_t));
  m.idx = (int64_t *)ora_xcalloc(nnz + 1, sizeof(int64_t));
  m.val = (double *)ora_xcalloc(nnz + 1, sizeof(double));
  for (size_t c = 0; c < n; c++)
    for (int64_t p = colptr[c]; p < colptr[c + 1]; p++) {
      int64_t r = rowidx[p];
      m.idx[pos[r]] = (int64_t)c;
      m.val[pos[r]] = val[p];
      pos[r]++;
    }
  for (size_t c = 0; c < n; c++)
    for (int64_t p = colptr[c]; p < colptr[c + 1]; p++) {
      int64_t r = rowidx[p];
      if ((size_t)r != c) {
        m.idx[pos[c]] = r;
        m.val[pos[c]] = val[p];
        pos[c]++;
      }
    }
  free(pos);
  m.ptr = cnt;
  return m;
}

/* src/solver/iterative/csr.rs:69-97 (the per-row arithmetic does not depend on threading) */
static void csr_matvec(const csr *a, const double *x, double *y) {
  for (size_t i = 0; i < a->nrows; i++) {
    double s = 0.0;
    for (int64_t p = a->ptr[i]; p < a->ptr[i + 1]; p++) s += a->val[p] * x[a->idx[p]];
    y[i] = s;
  }
}

/* src/solver/iterative/csr.rs:99-109 */
static double *csr_diag(const csr *a) {
  double *d = (double *)ora_xcalloc(a->nrows + 1, sizeof(double));
  for (size_t i = 0; i < a->nrows; i++) {
    int64_t lo = a->ptr[i], hi = a->ptr[i + 1];
    int64_t q = ora_bsearch(a->idx + lo, (size_t)(hi - lo), (int64_t)i);
    if (q >= 0) d[i] = a->val[lo + q];
    else {
      d[i] = 0.0;
      for (int64_t p = lo; p < hi; p++)
        if (a->idx[p] == (int64_t)i) {
          d[i] = a->val[p];
          break;
        }
    }
  }
  return d;
}

/* src/solver/iterative/csr.rs:111-131 */
static csr csr_transpose(const csr *a) {
  csr t;
  t.nrows = a->ncols;
  t.ncols = a->nrows;
  size_t nnz = csr_nnz(a);
  int64_t *cnt = (int64_t *)ora_xcalloc(a->ncols + 1, sizeof(int64_t));
  for (size_t p = 0; p < nnz; p++) cnt[a->idx[p] + 1]++;
  for (size_t i = 0; i < a->ncols; i++) cnt[i + 1] += cnt[i];
  int64_t *pos = (int64_t *)ora_xmalloc((a->ncols + 1) * sizeof(int64_t));
  memcpy(pos, cnt, (a->ncols + 1) * sizeof(int64_t));
  t.idx = (int64_t *)ora_xmalloc((nnz + 1) * sizeof(int64_t));
  t.val = (double *)ora_xmalloc((nnz + 1) * sizeof(double));
  for (size_t r = 0; r < a->nrows; r++)
    for (int64_t p = a->ptr[r]; p < a->ptr[r + 1]; p++) {
      int64_t c = a->idx[p];
      t.idx[pos[c]] = (int64_t)r;
      t.val[pos[c]] = a->val[p];
      pos[c]++;
    }
  free(pos);
  t.ptr = cnt;
  return t;
}

/* Gustavson C = A B with sorted column indices (src/solver/iterative/csr.rs:134-189) */
typedef struct {
  const csr *a, *b;
  size_t r0, r1;
  vi cnt, idx;
  vd val;
} mul_part;

static void mul_work(void *p) {
  mul_part *w = (mul_part *)p;
  const csr *a = w->a, *b = w->b;
  int64_t *mark = (int64_t *)ora_xmalloc((b->ncols + 1) * sizeof(int64_t));
  double *acc = (double *)ora_xcalloc(b->ncols + 1, sizeof(double));
  for (size_t i = 0; i < b->ncols; i++) mark[i] = -1;
  vi cols = {0};
  for (size_t i = w->r0; i < w->r1; i++) {
    cols.n = 0;
    for (int64_t p = a->ptr[i]; p < a->ptr[i + 1]; p++) {
      int64_t k = a->idx[p];
      double av = a->val[p];
      for (int64_t q = b->ptr[k]; q < b->ptr[k + 1]; q++) {
        int64_t c = b->idx[q];
        if (mark[c] != (int64_t)i) {
          mark[c] = (int64_t)i;
          acc[c] = 0.0;
          vi_push(&cols, c);
        }
        acc[c] += av * b->val[q];
      }
    }
    ora_sort_i64(cols.a, cols.n);
    for (size_t j = 0; j < cols.n; j++) {
      vi_push(&w->idx, cols.a[j]);
      vd_push(&w->val, acc[cols.a[j]]);
    }
    vi_push(&w->cnt, (int64_t)cols.n);
  }
  vi_free(&cols);
  free(mark);
  free(acc);
}

static csr csr_mul(const csr *a, const csr *b) {
  size_t n = a->nrows;
  size_t t = (size_t)ora_threads();
  if (t > n / 2000) t = n / 2000;
  if (t < 1) t = 1;
  size_t chunk = (n + t - 1) / t;
  if (chunk < 1) chunk = 1;
  mul_part *parts = (mul_part *)ora_xcalloc(t, sizeof(mul_part));
  void **args = (void **)ora_xmalloc(t * sizeof(void *));
  for (size_t c = 0; c < t; c++) {
    size_t r0 = c * chunk < n ? c * chunk : n;
    size_t r1 = r0 + chunk < n ? r0 + chunk : n;
    parts[c].a = a;
    parts[c].b = b;
    parts[c].r0 = r0;
    parts[c].r1 = r1;
    args[c] = &parts[c];
  }
  ora_run_parallel(mul_work, args, (int)t);
  csr m;
  m.nrows = n;
  m.ncols = b->ncols;
  m.ptr = (int64_t *)ora_xmalloc((n + 1) * sizeof(int64_t));
  size_t tot = 0;
  for (size_t c = 0; c < t; c++) tot += parts[c].idx.n;
  m.idx = (int64_t *)ora_xmalloc((tot + 1) * sizeof(int64_t));
  m.val = (double *)ora_xmalloc((tot + 1) * sizeof(double));
  size_t row = 0, off = 0;
  m.ptr[0] = 0;
  for (size_t c = 0; c < t; c++) {
    for (size_t i = 0; i < parts[c].cnt.n; i++) {
      m.ptr[row + 1] = m.ptr[row] + parts[c].cnt.a[i];
      row++;
    }
    if (parts[c].idx.n) {
      memcpy(m.idx + off, parts[c].idx.a, parts[c].idx.n * sizeof(int64_t));
      memcpy(m.val + off, parts[c].val.a, parts[c].val.n * sizeof(double));
    }
    off += parts[c].idx.n;
    vi_free(&parts[c].cnt);
    vi_free(&parts[c].idx);
    vd_free(&parts[c].val);
  }
  free(parts);
  free(args);
  return m;
}

/* ---- preconditioners ---- */
typedef struct {
  int kind; /* 0 AMG, 1 IC0, 2 Jacobi */
  /* Jacobi */
  double *inv;
  /* IC0 */
  size_t n;
  int64_t *ptr, *idx;
  double *lval, *d;
  double shift;
  /* AMG */
  size_t nlev;
  csr *A, *P, *R;
  double **diag;
  size_t coarse_n;
  double *cl, *cd;
  int64_t sizes[16];
  double complexity;
} precond;

/* src/solver/iterative/csr.rs:259-293 */
static int ic0_try(size_t n, const int64_t *ptr, const int64_t *idx, const double *a, const double *diag,
                   double shift, double *val, double *d) {
  memcpy(val, a, (size_t)ptr[n] * sizeof(double));
  int64_t *mark = (int64_t *)ora_xmalloc((n + 1) * sizeof(int64_t));
  for (size_t i = 0; i < n; i++) mark[i] = -1;
  for (size_t i = 0; i < n; i++) d[i] = 0.0;
  for (size_t i = 0; i < n; i++) {
    int64_t lo = ptr[i], hi = ptr[i + 1];
    for (int64_t p = lo; p < hi; p++) mark[idx[p]] = p;
    for (int64_t p = lo; p < hi; p++) {
      int64_t j = idx[p];
      double s = val[p];
      for (int64_t q = ptr[j]; q < ptr[j + 1]; q++) {
        int64_t k = idx[q];
        int64_t pk = mark[k];
        if (pk != -1 && k < j) s -= val[pk] * d[k] * val[q];
      }
      val[p] = s / d[j];
    }
    double di = diag[i] * (1.0 + shift);
    for (int64_t p = lo; p < hi; p++) di -= val[p] * val[p] * d[idx[p]];
    for (int64_t p = lo; p < hi; p++) mark[idx[p]] = -1;
    if (!(di > 1e-12 * fabs(diag[i]))) {
      free(mark);
      return 0;
    }
    d[i] = di;
  }
  free(mark);
  return 1;
}

static void ic0_new(precond *pc, const csr *a) {
  size_t n = a->nrows;
  pc->n = n;
  pc->ptr = (int64_t *)ora_xcalloc(n + 1, sizeof(int64_t));
  vi idx = {0};
  vd av = {0};
  double *diag = (double *)ora_xcalloc(n + 1, sizeof(double));
  for (size_t i = 0; i < n; i++) {
    for (int64_t p = a->ptr[i]; p < a->ptr[i + 1]; p++) {
      int64_t j = a->idx[p];
      if ((size_t)j < i) {
        vi_push(&idx, j);
        vd_push(&av, a->val[p]);
      } else if ((size_t)j == i) {
        diag[i] += a->val[p];
      }
    }
    pc->ptr[i + 1] = (int64_t)idx.n;
  }
  pc->idx = idx.a ? idx.a : (int64_t *)ora_xmalloc(8);
  pc->lval = (double *)ora_xmalloc((idx.n + 1) * sizeof(double));
  pc->d = (double *)ora_xmalloc((n + 1) * sizeof(double));
  double shift = 0.0;
  for (;;) {
    if (ic0_try(n, pc->ptr, pc->idx, av.a, diag, shift, pc->lval, pc->d)) break;
    shift = shift == 0.0 ? 1e-3 : shift * 4.0;
  }
  pc->shift = shift;
  vd_free(&av);
  free(diag);
}

static void ic0_apply(const precond *pc, const double *r, double *z) {
  size_t n = pc->n;
  memcpy(z, r, n * sizeof(double));
  for (size_t i = 0; i < n; i++) {
    double s = z[i];
    for (int64_t p = pc->ptr[i]; p < pc->ptr[i + 1]; p++) s -= pc->lval[p] * z[pc->idx[p]];
    z[i] = s;
  }
  for (size_t i = 0; i < n; i++) z[i] /= pc->d[i];
  for (size_t i = n; i-- > 0;) {
    double zi = z[i];
    for (int64_t p = pc->ptr[i]; p < pc->ptr[i + 1]; p++) z[pc->idx[p]] -= pc->lval[p] * zi;
  }
}

/* ---- AMG (src/solver/iterative/amg.rs) ---- */

/* src/solver/iterative/amg.rs:39-95 */
size_t ora_rigid_body_modes(size_t nn, const double *coords, const int64_t *ndof, double *b) {
  int planar = 1;
  for (size_t v = 0; v < nn; v++)
    if (coords[3 * v + 2] != 0.0) {
      planar = 0;
      break;
    }
  size_t nb = planar ? 3 : 6;
  size_t n = 0;
  for (size_t v = 0; v < nn; v++) n += (size_t)ndof[v];
  double c0[3] = {0, 0, 0};
  double cnt = (double)(nn ? nn : 1);
  for (size_t v = 0; v < nn; v++)
    for (int a = 0; a < 3; a++) c0[a] += coords[3 * v + a] / cnt;
  for (size_t i = 0; i < n * nb; i++) b[i] = 0.0;
  size_t row = 0;
  for (size_t v = 0; v < nn; v++) {
    size_t d = (size_t)ndof[v];
    double x[3] = {coords[3 * v] - c0[0], coords[3 * v + 1] - c0[1], coords[3 * v + 2] - c0[2]};
    for (size_t l = 0; l < d; l++) {
      double *br = b + (row + l) * nb;
      if (planar) {
        if (l == 0) {
          br[0] = 1.0;
          br[2] = -x[1];
        } else if (l == 1) {
          br[1] = 1.0;
          br[2] = x[0];
        } else if (d == 3 && l == 2) {
          br[2] = 1.0;
        } else {
          br[l % nb] = 1.0;
        }
      } else {
        if (l == 0) {
          br[0] = 1.0;
          br[4] = x[2];
          br[5] = -x[1];
        } else if (l == 1) {
          br[1] = 1.0;
          br[3] = -x[2];
          br[5] = x[0];
        } else if (l == 2) {
          br[2] = 1.0;
          br[3] = x[1];
          br[4] = -x[0];
        } else if (l >= 3 && l <= 5 && d >= 6) {
          br[l] = 1.0;
        } else {
          br[l % nb] = 1.0;
        }
      }
    }
    row += d;
  }
  return nb;
}

/* src/solver/iterative/amg.rs:108-181 */
static int64_t *aggregate(const csr *a, const int64_t *bptr, size_t nblk, double theta, size_t *nagg_out) {
  size_t nb = nblk;
  int64_t *blk = (int64_t *)ora_xmalloc(((size_t)bptr[nb] + 1) * sizeof(int64_t));
  for (size_t b = 0; b < nb; b++)
    for (int64_t i = bptr[b]; i < bptr[b + 1]; i++) blk[i] = (int64_t)b;
  double *dnorm = (double *)ora_xcalloc(nb + 1, sizeof(double));
  for (size_t i = 0; i < a->nrows; i++)
    for (int64_t p = a->ptr[i]; p < a->ptr[i + 1]; p++)
      if (blk[a->idx[p]] == blk[i]) dnorm[blk[i]] += a->val[p] * a->val[p];
  vi *strong = (vi *)ora_xcalloc(nb + 1, sizeof(vi));
  double *acc = (double *)ora_xcalloc(nb + 1, sizeof(double));
  int64_t *mark = (int64_t *)ora_xmalloc((nb + 1) * sizeof(int64_t));
  for (size_t b = 0; b < nb; b++) mark[b] = -1;
  vi list = {0};
  for (size_t b = 0; b < nb; b++) {
    list.n = 0;
    for (int64_t i = bptr[b]; i < bptr[b + 1]; i++)
      for (int64_t p = a->ptr[i]; p < a->ptr[i + 1]; p++) {
        int64_t c = blk[a->idx[p]];
        if ((size_t)c == b) continue;
        if (mark[c] != (int64_t)b) {
          mark[c] = (int64_t)b;
          acc[c] = 0.0;
          vi_push(&list, c);
        }
        acc[c] += a->val[p] * a->val[p];
      }
    for (size_t q = 0; q < list.n; q++) {
      int64_t c = list.a[q];
      if (acc[c] > theta * theta * sqrt(dnorm[b] * dnorm[c])) vi_push(&strong[b], c);
    }
  }
  vi_free(&list);
  free(acc);
  free(mark);
  free(dnorm);
  free(blk);
  int64_t *agg = (int64_t *)ora_xmalloc((nb + 1) * sizeof(int64_t));
  for (size_t b = 0; b < nb; b++) agg[b] = -1;
  int64_t n = 0;
  for (size_t b = 0; b < nb; b++) {
    if (agg[b] != -1) continue;
    int all = 1;
    for (size_t q = 0; q < strong[b].n; q++)
      if (agg[strong[b].a[q]] != -1) {
        all = 0;
        break;
      }
    if (all) {
      agg[b] = n;
      for (size_t q = 0; q < strong[b].n; q++) agg[strong[b].a[q]] = n;
      n++;
    }
  }
  int64_t *snap = (int64_t *)ora_xmalloc((nb + 1) * sizeof(int64_t));
  memcpy(snap, agg, nb * sizeof(int64_t));
  for (size_t b = 0; b < nb; b++) {
    if (agg[b] != -1) continue;
    for (size_t q = 0; q < strong[b].n; q++)
      if (snap[strong[b].a[q]] != -1) {
        agg[b] = snap[strong[b].a[q]];
        break;
      }
  }
  free(snap);
  for (size_t b = 0; b < nb; b++) {
    if (agg[b] != -1) continue;
    agg[b] = n;
    for (size_t q = 0; q < strong[b].n; q++)
      if (agg[strong[b].a[q]] == -1) agg[strong[b].a[q]] = n;
    n++;
  }
  for (size_t b = 0; b < nb; b++) vi_free(&strong[b]);
  free(strong);
  *nagg_out = (size_t)n;
  return agg;
}

typedef struct { int64_t r, c; double v; } trip;
static int trip_cmp(const void *x, const void *y) {
  const trip *a = (const trip *)x, *b = (const trip *)y;
  if (a->r != b->r) return a->r < b->r ? -1 : 1;
  if (a->c != b->c) return a->c < b->c ? -1 : 1;
  return 0;
}

/* src/solver/iterative/amg.rs:184-257; returns P~, coarse null space bc (nc x nbm), coarse block bounds */
static csr tentative(const int64_t *bptr, size_t nblk, const int64_t *agg, size_t nagg, const double *b, size_t nbm,
                     double **bc_out, int64_t **cptr_out, size_t *ncb_out) {
  size_t n = (size_t)bptr[nblk];
  vi *members = (vi *)ora_xcalloc(nagg + 1, sizeof(vi));
  for (size_t blk = 0; blk < nblk; blk++) vi_push(&members[agg[blk]], (int64_t)blk);
  size_t ecap = 1024, ne = 0;
  trip *ent = (trip *)ora_xmalloc(ecap * sizeof(trip));
  vd bc = {0};
  int64_t *cptr = (int64_t *)ora_xmalloc((nagg + 2) * sizeof(int64_t));
  cptr[0] = 0;
  size_t nc = 0;
  for (size_t g = 0; g < nagg; g++) {
    vi dofs = {0};
    for (size_t q = 0; q < members[g].n; q++) {
      int64_t blk = members[g].a[q];
      for (int64_t i = bptr[blk]; i < bptr[blk + 1]; i++) vi_push(&dofs, i);
    }
    size_t nd = dofs.n;
    double *q = (double *)ora_xmalloc((nbm * nd + 1) * sizeof(double)); /* column c: q[c*nd + i] */
    for (size_t c = 0; c < nbm; c++)
      for (size_t i = 0; i < nd; i++) q[c * nd + i] = b[(size_t)dofs.a[i] * nbm + c];
    size_t rr = nbm < nd ? nbm : nd;
    double *r = (double *)ora_xcalloc(rr * nbm + 1, sizeof(double));
    double *kept = (double *)ora_xmalloc((nbm * nd + nd + 1) * sizeof(double));
    size_t nk = 0;
    double scale = 0.0;
    for (size_t c = 0; c < nbm; c++) {
      double s = 0.0;
      for (size_t i = 0; i < nd; i++) s += q[c * nd + i] * q[c * nd + i];
      scale = fmax(scale, s);
    }
    scale = sqrt(scale);
    for (size_t c = 0; c < nbm; c++) {
      double *v = q + c * nd;
      for (int pass = 0; pass < 2; pass++)
        for (size_t k = 0; k < nk; k++) {
          const double *qk = kept + k * nd;
          double s = 0.0;
          for (size_t i = 0; i < nd; i++) s += qk[i] * v[i];
          if (pass == 0) r[k * nbm + c] = s; else r[k * nbm + c] += s;
          for (size_t i = 0; i < nd; i++) v[i] -= s * qk[i];
        }
      double norm = 0.0;
      for (size_t i = 0; i < nd; i++) norm += v[i] * v[i];
      norm = sqrt(norm);
      if (nk < nd && norm > 1e-10 * fmax(scale, 2.2250738585072014e-308)) {
        r[nk * nbm + c] = norm;
        for (size_t i = 0; i < nd; i++) kept[nk * nd + i] = v[i] / norm;
        nk++;
      }
    }
    if (nk == 0) {
      double s = 1.0 / sqrt((double)nd);
      for (size_t i = 0; i < nd; i++) kept[i] = s;
      nk = 1;
    }
    for (size_t k = 0; k < nk; k++)
      for (size_t li = 0; li < nd; li++) {
        double x = kept[k * nd + li];
        if (x != 0.0) {
          if (ne == ecap) {
            ecap *= 2;
            ent = (trip *)ora_xrealloc(ent, ecap * sizeof(trip));
          }
          trip t = {dofs.a[li], (int64_t)(nc + k), x};
          ent[ne++] = t;
        }
      }
    for (size_t k = 0; k < nk; k++)
      for (size_t c = 0; c < nbm; c++) vd_push(&bc, k < rr ? r[k * nbm + c] : 0.0);
    nc += nk;
    cptr[g + 1] = (int64_t)nc;
    free(q);
    free(r);
    free(kept);
    vi_free(&dofs);
  }
  for (size_t g = 0; g < nagg; g++) vi_free(&members[g]);
  free(members);
  qsort(ent, ne, sizeof(trip), trip_cmp);
  csr p;
  p.nrows = n;
  p.ncols = nc;
  p.ptr = (int64_t *)ora_xcalloc(n + 1, sizeof(int64_t));
  for (size_t e = 0; e < ne; e++) p.ptr[ent[e].r + 1]++;
  for (size_t i = 0; i < n; i++) p.ptr[i + 1] += p.ptr[i];
  p.idx = (int64_t *)ora_xmalloc((ne + 1) * sizeof(int64_t));
  p.val = (double *)ora_xmalloc((ne + 1) * sizeof(double));
  for (size_t e = 0; e < ne; e++) {
    p.idx[e] = ent[e].c;
    p.val[e] = ent[e].v;
  }
  free(ent);
  *bc_out = bc.a;
  *cptr_out = cptr;
  *ncb_out = nagg;
  return p;
}

/* src/solver/iterative/amg.rs:260-276 */
static double spectral_radius(const csr *a, const double *diag) {
  size_t n = a->nrows;
  double *x = (double *)ora_xmalloc((n + 1) * sizeof(double));
  double *y = (double *)ora_xcalloc(n + 1, sizeof(double));
  for (size_t i = 0; i < n; i++) x[i] = 1.0 + (double)((i * 7919) % 13) / 13.0;
  double rho = 1.0;
  for (int it = 0; it < 15; it++) {
    double nx = 0.0;
    for (size_t i = 0; i < n; i++) nx += x[i] * x[i];
    nx = sqrt(nx);
    for (size_t i = 0; i < n; i++) x[i] /= nx;
    csr_matvec(a, x, y);
    for (size_t i = 0; i < n; i++) y[i] /= diag[i];
    double s = 0.0;
    for (size_t i = 0; i < n; i++) s += y[i] * y[i];
    rho = sqrt(s);
    double *t = x;
    x = y;
    y = t;
  }
  free(x);
  free(y);
  return rho * 1.05;
}

/* src/solver/iterative/amg.rs:413-445 */
static csr sub_scaled(const csr *pt, const csr *ap, const double *diag, double omega) {
  size_t n = pt->nrows;
  csr m;
  m.nrows = n;
  m.ncols = pt->ncols;
  m.ptr = (int64_t *)ora_xmalloc((n + 1) * sizeof(int64_t));
  vi idx = {0};
  vd val = {0};
  m.ptr[0] = 0;
  for (size_t i = 0; i < n; i++) {
    double s = omega / diag[i];
    int64_t p = pt->ptr[i], pe = pt->ptr[i + 1];
    int64_t q = ap->ptr[i], qe = ap->ptr[i + 1];
    while (p < pe || q < qe) {
      int64_t cp = p < pe ? pt->idx[p] : INT64_MAX;
      int64_t cq = q < qe ? ap->idx[q] : INT64_MAX;
      int64_t c;
      double v;
      if (cp == cq) {
        v = pt->val[p] - s * ap->val[q];
        p++;
        q++;
        c = cp;
      } else if (cp < cq) {
        p++;
        c = cp;
        v = pt->val[p - 1];
      } else {
        q++;
        c = cq;
        v = -s * ap->val[q - 1];
      }
      if (v != 0.0) {
        vi_push(&idx, c);
        vd_push(&val, v);
      }
    }
    m.ptr[i + 1] = (int64_t)idx.n;
  }
  m.idx = idx.a ? idx.a : (int64_t *)ora_xmalloc(8);
  m.val = val.a ? val.a : (double *)ora_xmalloc(8);
  return m;
}

/* src/solver/iterative/amg.rs:282-341; returns 0 on success, -1 when the coarse matrix breaks down */
static int amg_new(precond *pc, csr a, int64_t *bptr, size_t nblk, double *b, size_t nbm, ora_amg *rec) {
  size_t cap = 16;
  pc->A = (csr *)ora_xcalloc(cap, sizeof(csr));
  pc->P = (csr *)ora_xcalloc(cap, sizeof(csr));
  pc->R = (csr *)ora_xcalloc(cap, sizeof(csr));
  pc->diag = (double **)ora_xcalloc(cap, sizeof(double *));
  pc->nlev = 0;
  size_t nsz = 0;
  pc->sizes[nsz++] = (int64_t)a.nrows;
  double fine_nnz = (double)csr_nnz(&a), total = fine_nnz;
  double theta = 0.08;
  int own_b = 0;
  while (a.nrows > 1500 && pc->nlev < 12) {
    double *diag = csr_diag(&a);
    for (size_t i = 0; i < a.nrows; i++)
      if (!(diag[i] > 0.0)) {
        free(diag);
        csr_free(&a);
        if (own_b) { free(b); free(bptr); }
        return -1;
      }
    size_t nagg = 0;
    int64_t *agg = aggregate(&a, bptr, nblk, theta, &nagg);
    double *bc = NULL;
    int64_t *cptr = NULL;
    size_t ncb = 0;
    csr pt = tentative(bptr, nblk, agg, nagg, b, nbm, &bc, &cptr, &ncb);
    if (rec && (double)pt.ncols <= 0.85 * (double)a.nrows) {
      size_t l = pc->nlev;
      rec->A[l] = csr_copy(&a);
      rec->Pt[l] = csr_copy(&pt);
      rec->agg[l] = (int64_t *)ora_xmalloc((nblk + 1) * sizeof(int64_t));
      memcpy(rec->agg[l], agg, nblk * sizeof(int64_t));
      rec->nagg[l] = nagg;
      rec->nblk[l] = nblk;
      rec->bptr[l] = (int64_t *)ora_xmalloc((nblk + 1) * sizeof(int64_t));
      memcpy(rec->bptr[l], bptr, (nblk + 1) * sizeof(int64_t));
      rec->B[l] = (double *)ora_xmalloc((a.nrows * nbm + 1) * sizeof(double));
      memcpy(rec->B[l], b, a.nrows * nbm * sizeof(double));
      rec->nbm = nbm;
    }
    free(agg);
    if ((double)pt.ncols > 0.85 * (double)a.nrows) {
      csr_free(&pt);
      free(bc);
      free(cptr);
      free(diag);
      break;
    }
    double omega = 4.0 / 3.0 / spectral_radius(&a, diag);
    csr ap = csr_mul(&a, &pt);
    csr p = sub_scaled(&pt, &ap, diag, omega);
    csr_free(&ap);
    csr_free(&pt);
    csr r = csr_transpose(&p);
    csr apm = csr_mul(&a, &p);
    csr ac = csr_mul(&r, &apm);
    csr_free(&apm);
    total += (double)csr_nnz(&ac);
    if (nsz < 16) pc->sizes[nsz++] = (int64_t)ac.nrows;
    if (rec) rec->P[pc->nlev] = csr_copy(&p);
    pc->A[pc->nlev] = a;
    pc->P[pc->nlev] = p;
    pc->R[pc->nlev] = r;
    pc->diag[pc->nlev] = diag;
    pc->nlev++;
    a = ac;
    if (own_b) { free(b); free(bptr); }
    bptr = cptr;
    nblk = ncb;
    b = bc;
    own_b = 1;
    theta *= 0.5;
  }
  for (size_t i = nsz; i < 16; i++) pc->sizes[i] = 0;
  if (rec) {
    size_t l = pc->nlev;
    rec->nlev = l;
    rec->A[l] = csr_copy(&a);
    rec->nblk[l] = nblk;
    rec->bptr[l] = (int64_t *)ora_xmalloc((nblk + 1) * sizeof(int64_t));
    memcpy(rec->bptr[l], bptr, (nblk + 1) * sizeof(int64_t));
    rec->B[l] = (double *)ora_xmalloc((a.nrows * nbm + 1) * sizeof(double));
    memcpy(rec->B[l], b, a.nrows * nbm * sizeof(double));
  }
  if (own_b) { free(b); free(bptr); }
  size_t n = a.nrows;
  double *l = (double *)ora_xcalloc(n * n + 1, sizeof(double));
  for (size_t i = 0; i < n; i++)
    for (int64_t p = a.ptr[i]; p < a.ptr[i + 1]; p++) l[i * n + a.idx[p]] = a.val[p];
  double *d = (double *)ora_xcalloc(n + 1, sizeof(double));
  double dmax = 0.0;
  for (size_t i = 0; i < n; i++) dmax = fmax(dmax, fabs(l[i * n + i]));
  for (size_t j = 0; j < n; j++) {
    double dj = l[j * n + j];
    for (size_t k = 0; k < j; k++) dj -= l[j * n + k] * l[j * n + k] * d[k];
    if (!(dj > 1e-13 * dmax)) {
      free(l);
      free(d);
      csr_free(&a);
      return -1;
    }
    d[j] = dj;
    for (size_t i = j + 1; i < n; i++) {
      double s = l[i * n + j];
      for (size_t k = 0; k < j; k++) s -= l[i * n + k] * l[j * n + k] * d[k];
      l[i * n + j] = s / dj;
    }
  }
  csr_free(&a);
  pc->coarse_n = n;
  pc->cl = l;
  pc->cd = d;
  pc->complexity = total / fine_nnz;
  return 0;
}

/* src/solver/iterative/amg.rs:343-364 */
static void coarse_solve(const precond *pc, const double *b, double *x) {
  size_t n = pc->coarse_n;
  const double *l = pc->cl;
  memcpy(x, b, n * sizeof(double));
  for (size_t i = 0; i < n; i++) {
    double s = x[i];
    for (size_t k = 0; k < i; k++) s -= l[i * n + k] * x[k];
    x[i] = s;
  }
  for (size_t i = 0; i < n; i++) x[i] /= pc->cd[i];
  for (size_t i = n; i-- > 0;) {
    double s = x[i];
    for (size_t k = i + 1; k < n; k++) s -= l[k * n + i] * x[k];
    x[i] = s;
  }
}

/* src/solver/iterative/amg.rs:366-409 */
static void vcycle(const precond *pc, size_t lev, const double *b, double *x) {
  if (lev == pc->nlev) {
    coarse_solve(pc, b, x);
    return;
  }
  const csr *a = &pc->A[lev];
  const double *dg = pc->diag[lev];
  size_t n = a->nrows;
  for (size_t i = 0; i < n; i++) x[i] = 0.0;
  for (size_t i = 0; i < n; i++) {
    double s = b[i];
    for (int64_t p = a->ptr[i]; p < a->ptr[i + 1]; p++) {
      int64_t j = a->idx[p];
      if ((size_t)j != i) s -= a->val[p] * x[j];
    }
    x[i] = s / dg[i];
  }
  double *ax = (double *)ora_xmalloc((n + 1) * sizeof(double));
  csr_matvec(a, x, ax);
  double *res = (double *)ora_xmalloc((n + 1) * sizeof(double));
  for (size_t i = 0; i < n; i++) res[i] = b[i] - ax[i];
  size_t nc = pc->R[lev].nrows;
  double *rc = (double *)ora_xcalloc(nc + 1, sizeof(double));
  csr_matvec(&pc->R[lev], res, rc);
  double *xc = (double *)ora_xcalloc(nc + 1, sizeof(double));
  vcycle(pc, lev + 1, rc, xc);
  double *corr = ax; /* reuse */
  csr_matvec(&pc->P[lev], xc, corr);
  for (size_t i = 0; i < n; i++) x[i] += corr[i];
  for (size_t i = n; i-- > 0;) {
    double s = b[i];
    for (int64_t p = a->ptr[i]; p < a->ptr[i + 1]; p++) {
      int64_t j = a->idx[p];
      if ((size_t)j != i) s -= a->val[p] * x[j];
    }
    x[i] = s / dg[i];
  }
  free(ax);
  free(res);
  free(rc);
  free(xc);
}

static void precond_apply(const precond *pc, const double *r, double *z, size_t n) {
  if (pc->kind == 2) {
    for (size_t i = 0; i < n; i++) z[i] = r[i] * pc->inv[i];
  } else if (pc->kind == 1) {
    ic0_apply(pc, r, z);
  } else {
    vcycle(pc, 0, r, z);
  }
}

static void precond_free(precond *pc) {
  free(pc->inv);
  free(pc->ptr);
  free(pc->idx);
  free(pc->lval);
  free(pc->d);
  for (size_t i = 0; i < pc->nlev; i++) {
    csr_free(&pc->A[i]);
    csr_free(&pc->P[i]);
    csr_free(&pc->R[i]);
    free(pc->diag[i]);
  }
  free(pc->A);
  free(pc->P);
  free(pc->R);
  free(pc->diag);
  free(pc->cl);
  free(pc->cd);
}

static double dotv(const double *a, const double *b, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* src/solver/iterative/csr.rs:341-380 */
static int64_t pcg(const csr *a, const double *b, double *x, const precond *m, double tol, int64_t maxit,
                   double *rel_out, int32_t *conv) {
  size_t n = a->nrows;
  for (size_t i = 0; i < n; i++) x[i] = 0.0;
  double bn = sqrt(dotv(b, b, n));
  if (bn == 0.0) {
    *rel_out = 0.0;
    *conv = 1;
    return 0;
  }
  double *r = (double *)ora_xmalloc((n + 1) * sizeof(double));
  double *z = (double *)ora_xcalloc(n + 1, sizeof(double));
  double *p = (double *)ora_xmalloc((n + 1) * sizeof(double));
  double *q = (double *)ora_xcalloc(n + 1, sizeof(double));
  memcpy(r, b, n * sizeof(double));
  precond_apply(m, r, z, n);
  memcpy(p, z, n * sizeof(double));
  double rz = dotv(r, z, n);
  double rel = 1.0;
  int64_t it;
  for (it = 1; it <= maxit; it++) {
    csr_matvec(a, p, q);
    double pq = dotv(p, q, n);
    if (!(pq > 0.0)) {
      *rel_out = rel;
      *conv = 0;
      goto done;
    }
    double alpha = rz / pq;
    for (size_t i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    rel = sqrt(dotv(r, r, n)) / bn;
    if (rel <= tol) {
      *rel_out = rel;
      *conv = 1;
      goto done;
    }
    precond_apply(m, r, z, n);
    double rz_new = dotv(r, z, n);
    double beta = rz_new / rz;
    rz = rz_new;
    for (size_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
  }
  it = maxit;
  *rel_out = rel;
  *conv = 0;
done:
  free(r);
  free(z);
  free(p);
  free(q);
  return it;
}

/* src/solver/iterative.rs:275-292 */
int64_t ora_solve_lower(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val,
                        const double *b, int kind, size_t nn, const double *coords, const int64_t *ndof,
                        double tol, int64_t maxit, double *x, double *rel, int32_t *converged,
                        int64_t *amg_sizes, double *complexity) {
  csr a = csr_from_lower(n, colptr, rowidx, val);
  precond pc;
  memset(&pc, 0, sizeof(pc));
  pc.kind = kind;
  if (kind == 2) {
    double *d = csr_diag(&a);
    pc.inv = (double *)ora_xmalloc((n + 1) * sizeof(double));
    for (size_t i = 0; i < n; i++) pc.inv[i] = d[i] != 0.0 ? 1.0 / d[i] : 1.0;
    free(d);
  } else if (kind == 1) {
    ic0_new(&pc, &a);
  } else {
    int64_t *bptr = (int64_t *)ora_xmalloc((nn + 1) * sizeof(int64_t));
    bptr[0] = 0;
    for (size_t v = 0; v < nn; v++) bptr[v + 1] = bptr[v] + ndof[v];
    double *bm = (double *)ora_xmalloc((n * 6 + 1) * sizeof(double));
    size_t nbm = ora_rigid_body_modes(nn, coords, ndof, bm);
    csr acopy;
    acopy.nrows = a.nrows;
    acopy.ncols = a.ncols;
    size_t nnz = csr_nnz(&a);
    acopy.ptr = (int64_t *)ora_xmalloc((n + 1) * sizeof(int64_t));
    acopy.idx = (int64_t *)ora_xmalloc((nnz + 1) * sizeof(int64_t));
    acopy.val = (double *)ora_xmalloc((nnz + 1) * sizeof(double));
    memcpy(acopy.ptr, a.ptr, (n + 1) * sizeof(int64_t));
    memcpy(acopy.idx, a.idx, nnz * sizeof(int64_t));
    memcpy(acopy.val, a.val, nnz * sizeof(double));
    if (amg_new(&pc, acopy, bptr, nn, bm, nbm, NULL) != 0) {
      free(bptr);
      free(bm);
      precond_free(&pc);
      csr_free(&a);
      *converged = 0;
      *rel = NAN;
      return -1;
    }
    free(bptr);
    free(bm);
    if (amg_sizes) memcpy(amg_sizes, pc.sizes, sizeof(pc.sizes));
    if (complexity) *complexity = pc.complexity;
  }
  if (maxit <= 0) maxit = (int64_t)(n > 20000 ? n : 20000);
  int64_t it = pcg(&a, b, x, &pc, tol, maxit, rel, converged);
  precond_free(&pc);
  csr_free(&a);
  return it;
}

/* IC(0) factor of the lower CSC matrix (src/solver/iterative/csr.rs:229-257):
 * strictly lower L row-wise (ptr n+1, idx/val nnz(L)) and D; returns nnz(L).
 * With NULL outputs only the count is returned. */
int64_t ora_ic0_factor(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val, int64_t *ptr,
                       int64_t *idx, double *lval, double *d, double *shift) {
  csr a = csr_from_lower(n, colptr, rowidx, val);
  precond pc;
  memset(&pc, 0, sizeof(pc));
  pc.kind = 1;
  ic0_new(&pc, &a);
  int64_t nnz = pc.ptr[n];
  if (ptr) memcpy(ptr, pc.ptr, (n + 1) * sizeof(int64_t));
  if (idx) memcpy(idx, pc.idx, (size_t)nnz * sizeof(int64_t));
  if (lval) memcpy(lval, pc.lval, (size_t)nnz * sizeof(double));
  if (d) memcpy(d, pc.d, n * sizeof(double));
  if (shift) *shift = pc.shift;
  precond_free(&pc);
  csr_free(&a);
  return nnz;
}

/* z = M r for the IC(0) (kind 1) or Jacobi (kind 2) preconditioner of the
 * lower CSC matrix (Precond::apply, src/solver/iterative/csr.rs:193-217,
 * 296-316). */
void ora_precond_apply(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val, int kind,
                       const double *r, double *z) {
  csr a = csr_from_lower(n, colptr, rowidx, val);
  precond pc;
  memset(&pc, 0, sizeof(pc));
  pc.kind = kind;
  if (kind == 2) {
    double *dg = csr_diag(&a);
    pc.inv = (double *)ora_xmalloc((n + 1) * sizeof(double));
    for (size_t i = 0; i < n; i++) pc.inv[i] = dg[i] != 0.0 ? 1.0 / dg[i] : 1.0;
    free(dg);
  } else {
    pc.kind = 1;
    ic0_new(&pc, &a);
  }
  precond_apply(&pc, r, z, n);
  precond_free(&pc);
  csr_free(&a);
}

/* src/solver/iterative.rs:74-97 */
size_t ora_natural_pattern(size_t nn, const int64_t *adj_ptr, const int64_t *adj_idx, const int64_t *ndof,
                           int64_t *first, int64_t *colptr, int64_t *rowidx) {
  int64_t *fst = (int64_t *)ora_xmalloc((nn + 1) * sizeof(int64_t));
  int64_t n = 0;
  for (size_t v = 0; v < nn; v++) {
    fst[v] = n;
    n += ndof[v];
  }
  if (first) memcpy(first, fst, nn * sizeof(int64_t));
  size_t nnz = 0;
  if (colptr) colptr[0] = 0;
  vi nb = {0};
  for (size_t v = 0; v < nn; v++) {
    nb.n = 0;
    for (int64_t p = adj_ptr[v]; p < adj_ptr[v + 1]; p++)
      if ((size_t)adj_idx[p] > v) vi_push(&nb, adj_idx[p]);
    ora_sort_i64(nb.a, nb.n);
    for (int64_t c = fst[v]; c < fst[v] + ndof[v]; c++) {
      for (int64_t r = c; r < fst[v] + ndof[v]; r++) {
        if (rowidx) rowidx[nnz] = r;
        nnz++;
      }
      for (size_t q = 0; q < nb.n; q++) {
        int64_t w = nb.a[q];
        for (int64_t r = fst[w]; r < fst[w] + ndof[w]; r++) {
          if (rowidx) rowidx[nnz] = r;
          nnz++;
        }
      }
      if (colptr) colptr[c + 1] = (int64_t)nnz;
    }
  }
  vi_free(&nb);
  free(fst);
  return nnz;
}

/* ---- AMG hierarchy and aggregation hooks (for the tests) ---- */

/* src/solver/iterative/amg.rs:108-181 alone, on a full CSR matrix */
int64_t ora_amg_aggregate(size_t n, const int64_t *ptr, const int64_t *idx, const double *val, size_t nblk,
                          const int64_t *bptr, double theta, int64_t *agg_out) {
  csr a;
  a.nrows = a.ncols = n;
  a.ptr = (int64_t *)ptr;
  a.idx = (int64_t *)idx;
  a.val = (double *)val;
  size_t nagg = 0;
  int64_t *agg = aggregate(&a, bptr, nblk, theta, &nagg);
  memcpy(agg_out, agg, nblk * sizeof(int64_t));
  free(agg);
  return (int64_t)nagg;
}

/* Amg::new (src/solver/iterative/amg.rs:282-341) on the lower triangle with
 * rigid body modes, recording every level; NULL when the setup fails */
ora_amg *ora_amg_build(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val, size_t nn,
                       const double *coords, const int64_t *ndof) {
  csr a = csr_from_lower(n, colptr, rowidx, val);
  int64_t *bptr = (int64_t *)ora_xmalloc((nn + 1) * sizeof(int64_t));
  bptr[0] = 0;
  for (size_t v = 0; v < nn; v++) bptr[v + 1] = bptr[v] + ndof[v];
  double *bm = (double *)ora_xmalloc((n * 6 + 1) * sizeof(double));
  size_t nbm = ora_rigid_body_modes(nn, coords, ndof, bm);
  precond pc;
  memset(&pc, 0, sizeof(pc));
  pc.kind = 0;
  ora_amg *rec = (ora_amg *)ora_xcalloc(1, sizeof(ora_amg));
  int rc = amg_new(&pc, a, bptr, nn, bm, nbm, rec);
  free(bptr);
  free(bm);
  precond_free(&pc);
  if (rc != 0) {
    ora_amg_free(rec);
    return NULL;
  }
  return rec;
}

size_t ora_amg_nlev(const ora_amg *h) { return h->nlev; }

/* level l (0..nlev): what -1 info[0..7] = rows, nnz(A), blocks, aggregates,
 * null-space columns, coarse columns, nnz(P~), nnz(P); 0 A, 1 P~, 2 P (CSR
 * into iout/idx/dout); 3 aggregates, 4 block bounds (iout); 5 B (dout) */
int32_t ora_amg_level(const ora_amg *h, size_t l, int32_t what, int64_t *iout, int64_t *idx, double *dout) {
  if (l > h->nlev) return -1;
  const int coarsest = l == h->nlev;
  const csr *A = &h->A[l];
  const csr *m = NULL;
  switch (what) {
    case -1: {
      int64_t info[8] = {(int64_t)A->nrows, (int64_t)csr_nnz(A), (int64_t)h->nblk[l], 0, (int64_t)h->nbm, 0, 0, 0};
      if (!coarsest) {
        info[3] = (int64_t)h->nagg[l];
        info[5] = (int64_t)h->P[l].ncols;
        info[6] = (int64_t)csr_nnz(&h->Pt[l]);
        info[7] = (int64_t)csr_nnz(&h->P[l]);
      }
      memcpy(iout, info, sizeof info);
      return 0;
    }
    case 0: m = A; break;
    case 1: if (coarsest) return -1; m = &h->Pt[l]; break;
    case 2: if (coarsest) return -1; m = &h->P[l]; break;
    case 3:
      if (coarsest) return -1;
      memcpy(iout, h->agg[l], h->nblk[l] * sizeof(int64_t));
      return 0;
    case 4:
      memcpy(iout, h->bptr[l], (h->nblk[l] + 1) * sizeof(int64_t));
      return 0;
    case 5:
      memcpy(dout, h->B[l], A->nrows * h->nbm * sizeof(double));
      return 0;
    default: return -1;
  }
  size_t nnz = csr_nnz(m);
  if (iout) memcpy(iout, m->ptr, (m->nrows + 1) * sizeof(int64_t));
  if (idx) memcpy(idx, m->idx, nnz * sizeof(int64_t));
  if (dout) memcpy(dout, m->val, nnz * sizeof(double));
  return 0;
}

void ora_amg_free(ora_amg *h) {
  if (!h) return;
  for (size_t l = 0; l < ORA_MAXLEV; l++) {
    if (h->A[l].ptr) csr_free(&h->A[l]);
    if (h->Pt[l].ptr) csr_free(&h->Pt[l]);
    if (h->P[l].ptr) csr_free(&h->P[l]);
    free(h->agg[l]);
    free(h->bptr[l]);
    free(h->B[l]);
  }
  free(h);
}
