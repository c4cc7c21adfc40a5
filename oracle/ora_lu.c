/*
 * Oracle: multifrontal LU for unsymmetric values on a symmetric pattern,
 * diagonal threshold pivoting with delays and partial pivoting at the root.
 * Restates src/sparse/lu.rs.  Test infrastructure only (see ora.h).
 */
#include "ora_internal.h"
#include <math.h>

#define LU_ZERO_TOL 1e-12
#define LU_KB 64

typedef struct {
  vi piv;
  vi rperm;
  double *lu11;
  vi rows;
  double *l21, *u12;
  size_t k, nr;
} lufront;

struct ora_lufactor {
  lufront *fr;
  size_t nf, cap;
  int64_t delayed;
};

typedef struct {
  const ora_sym *S;
  const int64_t *colptr, *rowidx;
  const double *lo, *ut;
  double *scale;
  double thr;
} lctx;

typedef struct {
  double *upd;
  vi idx;
  lufront *fronts;
  size_t nfronts, cap;
  int64_t delayed, err;
} lout;

static void lout_push(lout *o, lufront f) {
  if (o->nfronts == o->cap) {
    o->cap = o->cap ? 2 * o->cap : 8;
    o->fronts = (lufront *)ora_xrealloc(o->fronts, o->cap * sizeof(lufront));
  }
  o->fronts[o->nfronts++] = f;
}

static void lufront_free(lufront *f) {
  vi_free(&f->piv);
  vi_free(&f->rperm);
  free(f->lu11);
  vi_free(&f->rows);
  free(f->l21);
  free(f->u12);
}

static void lout_release(lout *o) {
  free(o->upd);
  vi_free(&o->idx);
  for (size_t i = 0; i < o->nfronts; i++) lufront_free(&o->fronts[i]);
  free(o->fronts);
  memset(o, 0, sizeof(*o));
}

/* src/sparse/lu.rs:71-87 */
static double *full_scale(const int64_t *colptr, const int64_t *rowidx, const double *lo, const double *ut,
                          size_t n) {
  double *s = (double *)ora_xcalloc(n ? n : 1, sizeof(double));
  for (int t = 0; t < 2; t++) {
    const double *val = t ? ut : lo;
    for (size_t c = 0; c < n; c++)
      for (int64_t p = colptr[c]; p < colptr[c + 1]; p++) {
        int64_t r = rowidx[p];
        if (t && (size_t)r == c) continue;
        double v = fabs(val[p]);
        s[c] = fmax(s[c], v);
        s[r] = fmax(s[r], v);
      }
  }
  return s;
}

/* src/sparse/lu.rs:246-277; returns -1 or the singular variable */
static int64_t dense_lu(double *s, size_t r, const int64_t *vars, const double *scale, int64_t *perm) {
  for (size_t i = 0; i < r; i++) perm[i] = (int64_t)i;
  for (size_t p = 0; p < r; p++) {
    size_t bi = p;
    double big = 0.0;
    for (size_t i = p; i < r; i++)
      if (fabs(s[i * r + p]) > big) {
        big = fabs(s[i * r + p]);
        bi = i;
      }
    if (!(big > LU_ZERO_TOL * scale[vars[p]])) return vars[p];
    if (bi != p) {
      for (size_t c = 0; c < r; c++) {
        double t = s[p * r + c];
        s[p * r + c] = s[bi * r + c];
        s[bi * r + c] = t;
      }
      int64_t t = perm[p];
      perm[p] = perm[bi];
      perm[bi] = t;
    }
    double dp = s[p * r + p];
    for (size_t i = p + 1; i < r; i++) {
      double l = s[i * r + p] / dp;
      s[i * r + p] = l;
      if (l != 0.0)
        for (size_t c = p + 1; c < r; c++) s[i * r + c] -= l * s[p * r + c];
    }
  }
  return -1;
}

/* src/sparse/lu.rs:279-290 */
static void lswap(double *f, size_t m, int64_t *fs, size_t p, size_t q) {
  if (p == q) return;
  for (size_t c = 0; c < m; c++) {
    double t = f[p * m + c];
    f[p * m + c] = f[q * m + c];
    f[q * m + c] = t;
  }
  for (size_t i = 0; i < m; i++) {
    double t = f[i * m + p];
    f[i * m + p] = f[i * m + q];
    f[i * m + q] = t;
  }
  int64_t t = fs[p];
  fs[p] = fs[q];
  fs[q] = t;
}

/* src/sparse/lu.rs:292-365 */
static size_t leliminate(double *f, size_t m, size_t kf, int64_t *fs, const lctx *c, int threads) {
  size_t p = 0, width = LU_KB;
  while (p < kf) {
    size_t p0 = p;
    size_t b1 = p + width < kf ? p + width : kf;
    while (p < b1) {
      size_t pick = (size_t)-1;
      for (size_t j = p; j < b1; j++) {
        double ajj = fabs(f[j * m + j]);
        if (!(ajj > LU_ZERO_TOL * c->scale[fs[j]])) continue;
        double g = 0.0;
        for (size_t i = p; i < m; i++)
          if (i != j) g = fmax(g, fabs(f[i * m + j]));
        if (ajj >= c->thr * g) {
          pick = j;
          break;
        }
      }
      if (pick == (size_t)-1) break;
      lswap(f, m, fs, p, pick);
      double dp = f[p * m + p];
      size_t nu = m - p - 1;
      double *urow = (double *)ora_xmalloc((nu ? nu : 1) * sizeof(double));
      memcpy(urow, f + p * m + p + 1, nu * sizeof(double));
      for (size_t i = p + 1; i < m; i++) {
        double li = f[i * m + p] / dp;
        f[i * m + p] = li;
        if (li == 0.0) continue;
        size_t end = i < b1 ? m : b1;
        double *row = f + i * m + p + 1;
        size_t cnt = end - p - 1;
        for (size_t q = 0; q < cnt; q++) row[q] -= li * urow[q];
      }
      free(urow);
      p += 1;
    }
    size_t npb = p - p0;
    if (npb > 0 && b1 < m) {
      size_t rows = m - b1;
      double *a = (double *)ora_xmalloc((rows * npb + 1) * sizeof(double));
      for (size_t r = 0; r < rows; r++) memcpy(a + r * npb, f + (b1 + r) * m + p0, npb * sizeof(double));
      double *b = (double *)ora_xmalloc((npb * rows + 1) * sizeof(double));
      for (size_t q = 0; q < npb; q++) memcpy(b + q * rows, f + (p0 + q) * m + b1, rows * sizeof(double));
      ora_par_gemm_rows(f + b1 * m, m, rows, b1, npb, a, b, threads);
      free(a);
      free(b);
    }
    if (npb == 0) {
      if (b1 == kf) break;
      width *= 2;
    } else {
      width = LU_KB;
    }
  }
  return p;
}

typedef struct {
  const lctx *c;
  int64_t t;
  int threads, root;
  lout out;
} ljob;

static void lnode(ljob *job);
static void ljob_run(void *p) { lnode((ljob *)p); }

/* src/sparse/lu.rs:139-242 */
static void lnode(ljob *job) {
  const lctx *c = job->c;
  const ora_sym *S = c->S;
  const ora_tnode *node = &S->tree[job->t];
  size_t nch = node->children.n;
  ljob cj[2];
  memset(cj, 0, sizeof(cj));
  for (size_t i = 0; i < nch; i++) {
    cj[i].c = c;
    cj[i].t = node->children.a[i];
  }
  if (nch == 2 && job->threads > 1 && fmin(S->tree[cj[0].t].work, S->tree[cj[1].t].work) > 2e7) {
    int half = job->threads / 2;
    cj[0].threads = half;
    cj[1].threads = job->threads - half;
    void *args[2] = {&cj[1], &cj[0]};
    ora_run_parallel(ljob_run, args, 2);
  } else {
    for (size_t i = 0; i < nch; i++) {
      cj[i].threads = job->threads;
      lnode(&cj[i]);
    }
  }
  memset(&job->out, 0, sizeof(job->out));
  job->out.err = -1;
  for (size_t i = 0; i < nch; i++)
    if (cj[i].out.err >= 0) {
      job->out.err = cj[i].out.err;
      for (size_t j = 0; j < nch; j++) lout_release(&cj[j].out);
      return;
    }
  size_t start = (size_t)node->start, k = (size_t)node->k, u = node->rows.n;
  vi fs = {0};
  for (size_t i = 0; i < k; i++) vi_push(&fs, (int64_t)(start + i));
  int64_t delayed = 0;
  for (size_t i = 0; i < nch; i++) {
    lout *o = &cj[i].out;
    for (size_t q = 0; q < o->nfronts; q++) lout_push(&job->out, o->fronts[q]);
    free(o->fronts);
    o->fronts = NULL;
    o->nfronts = o->cap = 0;
    delayed += o->delayed;
    for (size_t q = 0; q < o->idx.n; q++)
      if ((size_t)o->idx.a[q] < start) vi_push(&fs, o->idx.a[q]);
  }
  size_t kf = fs.n;
  delayed += (int64_t)(kf - k);
  size_t m = kf + u;
  size_t nd = kf - k;
  int64_t *dkeys = (int64_t *)ora_xmalloc((nd ? nd : 1) * sizeof(int64_t));
  int64_t *dord = (int64_t *)ora_xmalloc((nd ? nd : 1) * sizeof(int64_t));
  for (size_t i = 0; i < nd; i++) dord[i] = (int64_t)i;
  ora_sort_by_ikey(dord, nd, fs.a + k);
  for (size_t i = 0; i < nd; i++) dkeys[i] = fs.a[k + dord[i]];
  const int64_t *rows = node->rows.a;
#define LLOCAL(g)                                                                 \
  (((size_t)(g) >= start && (size_t)(g) < start + k) ? (size_t)(g) - start        \
   : ((size_t)(g) < start) ? k + (size_t)dord[ora_bsearch(dkeys, nd, (g))]        \
                           : kf + (size_t)ora_bsearch(rows, u, (g)))
  double *f = (double *)ora_xcalloc(m * m, sizeof(double));
  for (size_t cc = 0; cc < k; cc++) {
    size_t col = start + cc;
    for (int64_t p = c->colptr[col]; p < c->colptr[col + 1]; p++) {
      size_t lr = LLOCAL(c->rowidx[p]);
      f[lr * m + cc] += c->lo[p];
      if (lr != cc) f[cc * m + lr] += c->ut[p];
    }
  }
  for (size_t i = 0; i < nch; i++) {
    lout *o = &cj[i].out;
    size_t s = o->idx.n;
    size_t *map = (size_t *)ora_xmalloc((s ? s : 1) * sizeof(size_t));
    for (size_t x = 0; x < s; x++) map[x] = LLOCAL(o->idx.a[x]);
    for (size_t x = 0; x < s; x++) {
      size_t fx = map[x] * m;
      for (size_t y = 0; y < s; y++) f[fx + map[y]] += o->upd[x * s + y];
    }
    free(map);
    lout_release(o);
  }
#undef LLOCAL
  free(dkeys);
  free(dord);
  size_t npiv = leliminate(f, m, kf, fs.a, c, job->threads);
  vi rowsv = {0};
  for (size_t i = npiv; i < kf; i++) vi_push(&rowsv, fs.a[i]);
  for (size_t i = 0; i < u; i++) vi_push(&rowsv, rows[i]);
  size_t nr = rowsv.n;
  {
    lufront F;
    memset(&F, 0, sizeof(F));
    size_t kk = npiv;
    F.k = kk;
    F.nr = nr;
    for (size_t i = 0; i < kk; i++) vi_push(&F.piv, fs.a[i]);
    F.lu11 = (double *)ora_xmalloc((kk * kk + 1) * sizeof(double));
    for (size_t i = 0; i < kk; i++) memcpy(F.lu11 + i * kk, f + i * m, kk * sizeof(double));
    F.l21 = (double *)ora_xmalloc((nr * kk + 1) * sizeof(double));
    F.u12 = (double *)ora_xmalloc((kk * nr + 1) * sizeof(double));
    for (size_t r = 0; r < nr; r++) memcpy(F.l21 + r * kk, f + (kk + r) * m, kk * sizeof(double));
    for (size_t i = 0; i < kk; i++) memcpy(F.u12 + i * nr, f + i * m + kk, nr * sizeof(double));
    for (size_t i = 0; i < nr; i++) vi_push(&F.rows, rowsv.a[i]);
    lout_push(&job->out, F);
  }
  if (job->root && npiv < kf) {
    size_t r = kf - npiv;
    double *s = (double *)ora_xmalloc((r * r + 1) * sizeof(double));
    for (size_t i = 0; i < r; i++) memcpy(s + i * r, f + (npiv + i) * m + npiv, r * sizeof(double));
    int64_t *perm = (int64_t *)ora_xmalloc((r + 1) * sizeof(int64_t));
    int64_t e = dense_lu(s, r, fs.a + npiv, c->scale, perm);
    if (e >= 0) {
      job->out.err = e;
      free(s);
      free(perm);
      free(f);
      vi_free(&fs);
      vi_free(&rowsv);
      return;
    }
    lufront F;
    memset(&F, 0, sizeof(F));
    F.k = r;
    F.nr = 0;
    for (size_t i = 0; i < r; i++) {
      vi_push(&F.piv, fs.a[npiv + i]);
      vi_push(&F.rperm, perm[i]);
    }
    F.lu11 = s;
    F.l21 = (double *)ora_xmalloc(sizeof(double));
    F.u12 = (double *)ora_xmalloc(sizeof(double));
    lout_push(&job->out, F);
    free(perm);
    free(f);
    vi_free(&fs);
    vi_free(&rowsv);
    job->out.delayed = delayed;
    return;
  }
  double *upd = (double *)ora_xmalloc((nr * nr + 1) * sizeof(double));
  for (size_t x = 0; x < nr; x++) memcpy(upd + x * nr, f + (npiv + x) * m + npiv, nr * sizeof(double));
  free(f);
  vi_free(&fs);
  job->out.upd = upd;
  job->out.idx = rowsv;
  job->out.delayed = delayed;
}

ora_lufactor *ora_factor_lu(const ora_sym *S, const int64_t *colptr, const int64_t *rowidx, const double *lower,
                            const double *upper_t, double threshold, int64_t *err) {
  *err = -1;
  double thr = threshold < 1e-8 ? 1e-8 : (threshold > 1.0 ? 1.0 : threshold);
  lctx c = {S, colptr, rowidx, lower, upper_t, full_scale(colptr, rowidx, lower, upper_t, S->n), thr};
  ora_lufactor *F = (ora_lufactor *)ora_xcalloc(1, sizeof(ora_lufactor));
  for (size_t i = 0; i < S->roots.n; i++) {
    ljob job;
    memset(&job, 0, sizeof(job));
    job.c = &c;
    job.t = S->roots.a[i];
    job.threads = ora_threads();
    job.root = 1;
    lnode(&job);
    if (job.out.err >= 0) {
      *err = job.out.err;
      lout_release(&job.out);
      free(c.scale);
      ora_lufactor_free(F);
      return NULL;
    }
    for (size_t q = 0; q < job.out.nfronts; q++) {
      if (F->nf == F->cap) {
        F->cap = F->cap ? 2 * F->cap : 16;
        F->fr = (lufront *)ora_xrealloc(F->fr, F->cap * sizeof(lufront));
      }
      F->fr[F->nf++] = job.out.fronts[q];
    }
    F->delayed += job.out.delayed;
    free(job.out.fronts);
    free(job.out.upd);
    vi_free(&job.out.idx);
  }
  free(c.scale);
  return F;
}

void ora_lufactor_free(ora_lufactor *F) {
  if (!F) return;
  for (size_t i = 0; i < F->nf; i++) lufront_free(&F->fr[i]);
  free(F->fr);
  free(F);
}

void ora_lufactor_stats(const ora_lufactor *F, int64_t *out2) {
  size_t tot = 0;
  for (size_t i = 0; i < F->nf; i++) tot += F->fr[i].k * F->fr[i].k + 2 * F->fr[i].nr * F->fr[i].k;
  out2[0] = F->delayed;
  out2[1] = (int64_t)tot;
}

size_t ora_lufactor_pivots(const ora_lufactor *F, int64_t *eq, double *d, double *e, int32_t *kind) {
  size_t c = 0;
  for (size_t fi = 0; fi < F->nf; fi++) {
    const lufront *fr = &F->fr[fi];
    for (size_t i = 0; i < fr->k; i++, c++) {
      eq[c] = fr->piv.a[i];
      d[c] = fr->lu11[i * fr->k + i];
      e[c] = fr->rperm.n ? (double)fr->rperm.a[i] : 0.0;
      kind[c] = fr->rperm.n ? 3 : 1;
    }
  }
  return c;
}

/* src/sparse/lu.rs:369-404 */
void ora_lufactor_solve(const ora_lufactor *F, double *x) {
  double *xp = NULL, *xr = NULL;
  size_t cap = 0, rcap = 0;
  for (size_t fi = 0; fi < F->nf; fi++) {
    const lufront *fr = &F->fr[fi];
    size_t k = fr->k;
    if (k > cap) { cap = k; xp = (double *)ora_xrealloc(xp, cap * sizeof(double)); }
    if (fr->rperm.n == 0) {
      for (size_t i = 0; i < k; i++) xp[i] = x[fr->piv.a[i]];
    } else {
      for (size_t i = 0; i < k; i++) xp[i] = x[fr->piv.a[fr->rperm.a[i]]];
    }
    for (size_t i = 0; i < k; i++) {
      double s = ora_dot(fr->lu11 + i * k, xp, i);
      xp[i] -= s;
    }
    for (size_t r = 0; r < fr->nr; r++) x[fr->rows.a[r]] -= ora_dot(fr->l21 + r * k, xp, k);
    for (size_t i = 0; i < k; i++) x[fr->piv.a[i]] = xp[i];
  }
  for (size_t fi = F->nf; fi-- > 0;) {
    const lufront *fr = &F->fr[fi];
    size_t k = fr->k, nr = fr->nr;
    if (k > cap) { cap = k; xp = (double *)ora_xrealloc(xp, cap * sizeof(double)); }
    if (nr > rcap) { rcap = nr; xr = (double *)ora_xrealloc(xr, rcap * sizeof(double)); }
    for (size_t r = 0; r < nr; r++) xr[r] = x[fr->rows.a[r]];
    for (size_t i = 0; i < k; i++) xp[i] = x[fr->piv.a[i]];
    for (size_t i = 0; i < k; i++) xp[i] -= ora_dot(fr->u12 + i * nr, xr, nr);
    for (size_t i = k; i-- > 0;) {
      double s = ora_dot(fr->lu11 + i * k + i + 1, xp + i + 1, k - i - 1);
      xp[i] = (xp[i] - s) / fr->lu11[i * k + i];
    }
    for (size_t i = 0; i < k; i++) x[fr->piv.a[i]] = xp[i];
  }
  free(xp);
  free(xr);
}
