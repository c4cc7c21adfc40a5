/*
 * Oracle: multifrontal LDL^T for symmetric indefinite matrices with 1x1/2x2
 * threshold pivoting, delayed pivots and Bunch-Kaufman at the root.
 * Restates src/sparse/indef.rs.  Test infrastructure only (see ora.h).
 */
#include "ora_internal.h"
#include <math.h>

#define IND_ZERO_TOL 1e-12
#define IND_ALPHA 0.6403882032022079
#define IND_KB 64

typedef struct {
  vi piv;
  double *l11, *d, *e;
  uint8_t *kind;
  vi rows;
  double *l21;
  size_t l11n, l21n;
} ifront;

struct ora_ifactor {
  ifront *fr;
  size_t nf, cap;
  int64_t neg, zero, pos, delayed, two;
};

typedef struct {
  const ora_sym *S;
  const int64_t *colptr, *rowidx;
  const double *val;
  double *scale;
  double thr;
} ictx;

typedef struct {
  double *upd;
  vi idx;
  vi frs; /* indices into a per-job list of fronts (appended in postorder) */
  ifront *fronts;
  size_t nfronts, cap;
  int64_t delayed;
  int64_t err;
} iout;

static void iout_push(iout *o, ifront f) {
  if (o->nfronts == o->cap) {
    o->cap = o->cap ? 2 * o->cap : 8;
    o->fronts = (ifront *)ora_xrealloc(o->fronts, o->cap * sizeof(ifront));
  }
  o->fronts[o->nfronts++] = f;
}

static void ifront_free(ifront *f) {
  vi_free(&f->piv);
  free(f->l11);
  free(f->d);
  free(f->e);
  free(f->kind);
  vi_free(&f->rows);
  free(f->l21);
}

/* src/sparse/indef.rs:98-108 */
static double *column_scale(const int64_t *colptr, const int64_t *rowidx, const double *val, size_t n) {
  double *s = (double *)ora_xcalloc(n ? n : 1, sizeof(double));
  for (size_t c = 0; c < n; c++)
    for (int64_t p = colptr[c]; p < colptr[c + 1]; p++) {
      int64_t r = rowidx[p];
      double v = fabs(val[p]);
      s[c] = fmax(s[c], v);
      s[r] = fmax(s[r], v);
    }
  return s;
}

/* src/sparse/indef.rs:262-273 */
static void iswap(double *f, size_t m, size_t kf, int64_t *fs, size_t p, size_t q) {
  if (p == q) return;
  for (size_t c = 0; c < kf; c++) {
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

/* src/sparse/indef.rs:277-285 */
static double colmax(const double *f, size_t m, size_t p, size_t j, size_t x) {
  double g = 0.0;
  for (size_t i = p; i < m; i++)
    if (i != j && i != x) g = fmax(g, fabs(f[i * m + j]));
  return g;
}

/* 0 none, 1 one (j), 2 two (j, r) */
static int choose(const double *f, size_t m, size_t p, size_t end, const int64_t *fs, const ictx *c, size_t *pj,
                  size_t *pr) {
  double thr = c->thr;
  for (size_t j = p; j < end; j++) {
    double ajj = f[j * m + j];
    double sj = c->scale[fs[j]];
    double gj = colmax(f, m, p, j, (size_t)-1);
    if (fabs(ajj) > IND_ZERO_TOL * sj && fabs(ajj) >= thr * gj) {
      *pj = j;
      return 1;
    }
    size_t r = (size_t)-1;
    double best = 0.0;
    for (size_t i = p; i < end; i++)
      if (i != j && fabs(f[i * m + j]) > best) {
        best = fabs(f[i * m + j]);
        r = i;
      }
    if (r == (size_t)-1) continue;
    double a = ajj, b = f[r * m + j], cc = f[r * m + r];
    double det = a * cc - b * b;
    double sr = c->scale[fs[r]];
    if (!(fabs(det) > IND_ZERO_TOL * sj * sr)) continue;
    double gj2 = colmax(f, m, p, j, r);
    double gr2 = colmax(f, m, p, r, j);
    double ad = fabs(det);
    if ((fabs(cc) * gj2 + fabs(b) * gr2) / ad <= 1.0 / thr && (fabs(b) * gj2 + fabs(a) * gr2) / ad <= 1.0 / thr) {
      *pj = j;
      *pr = r;
      return 2;
    }
  }
  return 0;
}

/* src/sparse/indef.rs:326-372; returns -1 on success with kind in *kind, else the singular equation */
static int64_t choose_forced(const double *f, size_t m, size_t p, size_t kf, const int64_t *fs, const ictx *c,
                             int *kind, size_t *pj, size_t *pr) {
  size_t bi = p, bj = p;
  double big = -1.0;
  for (size_t j = p; j < kf; j++)
    for (size_t i = j; i < kf; i++) {
      double v = fabs(f[i * m + j]);
      if (v > big) {
        big = v;
        bi = i;
        bj = j;
      }
    }
#define ONE(jj)                                                      \
  do {                                                               \
    if (fabs(f[(jj) * m + (jj)]) > IND_ZERO_TOL * c->scale[fs[(jj)]]) { \
      *kind = 1;                                                     \
      *pj = (jj);                                                    \
      return -1;                                                     \
    }                                                                \
    return fs[(jj)];                                                 \
  } while (0)
  if (!(big > IND_ZERO_TOL * fmax(c->scale[fs[bj]], c->scale[fs[bi]]))) return fs[bj];
  if (bi == bj) ONE(bj);
  size_t j = bj, r = bi;
  double lam = big;
  double ajj = fabs(f[j * m + j]);
  if (ajj >= IND_ALPHA * lam) ONE(j);
  double sig = 0.0;
  for (size_t i = p; i < kf; i++) {
    if (i == r) continue;
    double v = fabs(i >= r ? f[i * m + r] : f[r * m + i]);
    sig = fmax(sig, v);
  }
  if (ajj * sig >= IND_ALPHA * lam * lam) ONE(j);
  if (fabs(f[r * m + r]) >= IND_ALPHA * sig) ONE(r);
  double a = f[j * m + j], b = f[r * m + j], cc = f[r * m + r];
  double det = a * cc - b * b;
  if (!(fabs(det) > IND_ZERO_TOL * c->scale[fs[j]] * c->scale[fs[r]])) return fs[j];
  *kind = 2;
  *pj = j;
  *pr = r;
  return -1;
#undef ONE
}

/* src/sparse/indef.rs:375-389 */
static double elim1(double *f, size_t m, size_t p, size_t bend) {
  double dp = f[p * m + p];
  size_t nw = bend > p + 1 ? bend - p - 1 : 0;
  double *w = (double *)ora_xmalloc((nw ? nw : 1) * sizeof(double));
  for (size_t q = 0; q < nw; q++) w[q] = f[(p + 1 + q) * m + p];
  for (size_t i = p + 1; i < m; i++) {
    double li = f[i * m + p] / dp;
    if (li != 0.0) {
      double *row = f + i * m + p + 1;
      for (size_t q = 0; q < nw; q++) row[q] -= li * w[q];
    }
    f[i * m + p] = li;
  }
  free(w);
  return dp;
}

/* src/sparse/indef.rs:392-411 */
static void elim2(double *f, size_t m, size_t p, size_t bend, double *oa, double *ob, double *oc) {
  double a = f[p * m + p], b = f[(p + 1) * m + p], c = f[(p + 1) * m + p + 1];
  double det = a * c - b * b;
  size_t nw = bend > p + 2 ? bend - p - 2 : 0;
  double *w0 = (double *)ora_xmalloc((nw ? nw : 1) * sizeof(double));
  double *w1 = (double *)ora_xmalloc((nw ? nw : 1) * sizeof(double));
  for (size_t q = 0; q < nw; q++) {
    w0[q] = f[(p + 2 + q) * m + p];
    w1[q] = f[(p + 2 + q) * m + p + 1];
  }
  for (size_t i = p + 2; i < m; i++) {
    double x0 = f[i * m + p], x1 = f[i * m + p + 1];
    double l0 = (x0 * c - x1 * b) / det;
    double l1 = (x1 * a - x0 * b) / det;
    if (l0 != 0.0 || l1 != 0.0) {
      double *row = f + i * m + p + 2;
      for (size_t q = 0; q < nw; q++) row[q] -= l0 * w0[q] + l1 * w1[q];
    }
    f[i * m + p] = l0;
    f[i * m + p + 1] = l1;
  }
  free(w0);
  free(w1);
  *oa = a;
  *ob = b;
  *oc = c;
}

/* src/sparse/indef.rs:504-532 */
static void trailing_update(double *f, size_t m, size_t p0, size_t npb, size_t b1, const double *d, const double *e,
                            const uint8_t *kind, int threads) {
  size_t rows = m - b1;
  double *w = (double *)ora_xcalloc(rows * npb + 1, sizeof(double));
  double *lt = (double *)ora_xcalloc(npb * rows + 1, sizeof(double));
  for (size_t r = 0; r < rows; r++) {
    const double *fr = f + (b1 + r) * m + p0;
    double *wr = w + r * npb;
    size_t q = 0;
    while (q < npb) {
      if (kind[q] == 2) {
        double l0 = fr[q], l1 = fr[q + 1];
        wr[q] = l0 * d[q] + l1 * e[q];
        wr[q + 1] = l0 * e[q] + l1 * d[q + 1];
        q += 2;
      } else {
        wr[q] = fr[q] * d[q];
        q += 1;
      }
    }
    for (q = 0; q < npb; q++) lt[q * rows + r] = fr[q];
  }
  ora_par_syrk(f + b1 * m, m, rows, b1, npb, w, lt, threads);
  free(w);
  free(lt);
}

typedef struct {
  vd d, e;
  vi kind;
} dek;

/* candidates left to the Bunch-Kaufman loop of a root in the last factorization (a test probe) */
static int64_t g_forced = 0;
int64_t ora_indef_last_forced(void) { return g_forced; }

/* src/sparse/indef.rs:416-500; returns npiv or -(1+equation) on a singular root */
static int64_t eliminate(double *f, size_t m, size_t kf, int64_t *fs, const ictx *c, int root, int threads, dek *o) {
  size_t p = 0, width = IND_KB;
  while (p < kf) {
    size_t p0 = p;
    size_t b1 = p + width < kf ? p + width : kf;
    while (p < b1) {
      size_t j = 0, r = 0;
      int kd = choose(f, m, p, b1, fs, c, &j, &r);
      if (kd == 1) {
        iswap(f, m, kf, fs, p, j);
        vd_push(&o->d, elim1(f, m, p, b1));
        vd_push(&o->e, 0.0);
        vi_push(&o->kind, 1);
        p += 1;
      } else if (kd == 2) {
        size_t rr = (r == p) ? j : r;
        iswap(f, m, kf, fs, p, j);
        iswap(f, m, kf, fs, p + 1, rr);
        double a, b, cc;
        elim2(f, m, p, b1, &a, &b, &cc);
        vd_push(&o->d, a);
        vd_push(&o->d, cc);
        vd_push(&o->e, b);
        vd_push(&o->e, 0.0);
        vi_push(&o->kind, 2);
        vi_push(&o->kind, 0);
        p += 2;
      } else {
        break;
      }
    }
    size_t npb = p - p0;
    if (npb > 0 && b1 < m) {
      uint8_t *kk = (uint8_t *)ora_xmalloc(npb);
      for (size_t q = 0; q < npb; q++) kk[q] = (uint8_t)o->kind.a[p0 + q];
      trailing_update(f, m, p0, npb, b1, o->d.a + p0, o->e.a + p0, kk, threads);
      free(kk);
      for (size_t i = b1; i < kf; i++)
        for (size_t jj = p; jj < i; jj++) f[jj * m + i] = f[i * m + jj];
    }
    if (npb == 0) {
      if (b1 == kf) break;
      width *= 2;
    } else {
      width = IND_KB;
    }
  }
  if (p < kf && root) {
    g_forced += (int64_t)(kf - p);
    while (p < kf) {
      int kd = 0;
      size_t j = 0, r = 0;
      int64_t e = choose_forced(f, m, p, kf, fs, c, &kd, &j, &r);
      if (e >= 0) return -(1 + e);
      if (kd == 1) {
        iswap(f, m, kf, fs, p, j);
        vd_push(&o->d, elim1(f, m, p, kf));
        vd_push(&o->e, 0.0);
        vi_push(&o->kind, 1);
        p += 1;
      } else {
        size_t rr = (r == p) ? j : r;
        iswap(f, m, kf, fs, p, j);
        iswap(f, m, kf, fs, p + 1, rr);
        double a, b, cc;
        elim2(f, m, p, kf, &a, &b, &cc);
        vd_push(&o->d, a);
        vd_push(&o->d, cc);
        vd_push(&o->e, b);
        vd_push(&o->e, 0.0);
        vi_push(&o->kind, 2);
        vi_push(&o->kind, 0);
        p += 2;
      }
    }
  }
  return (int64_t)p;
}

typedef struct {
  const ictx *c;
  int64_t t;
  int threads;
  int root;
  iout out;
} ijob;

static void inode(ijob *job);
static void ijob_run(void *p) { inode((ijob *)p); }

static void iout_release(iout *o) {
  free(o->upd);
  vi_free(&o->idx);
  for (size_t i = 0; i < o->nfronts; i++) ifront_free(&o->fronts[i]);
  free(o->fronts);
  memset(o, 0, sizeof(*o));
}

static int64_t bsearch_pair(const int64_t *keys, size_t n, int64_t g) { return ora_bsearch(keys, n, g); }

/* src/sparse/indef.rs:158-254 */
static void inode(ijob *job) {
  const ictx *c = job->c;
  const ora_sym *S = c->S;
  const ora_tnode *node = &S->tree[job->t];
  size_t nch = node->children.n;
  ijob cj[2];
  memset(cj, 0, sizeof(cj));
  for (size_t i = 0; i < nch; i++) {
    cj[i].c = c;
    cj[i].t = node->children.a[i];
    cj[i].root = 0;
  }
  if (nch == 2 && job->threads > 1 && fmin(S->tree[cj[0].t].work, S->tree[cj[1].t].work) > 2e7) {
    int half = job->threads / 2;
    cj[0].threads = half;
    cj[1].threads = job->threads - half;
    void *args[2] = {&cj[1], &cj[0]};
    ora_run_parallel(ijob_run, args, 2);
  } else {
    for (size_t i = 0; i < nch; i++) {
      cj[i].threads = job->threads;
      inode(&cj[i]);
    }
  }
  memset(&job->out, 0, sizeof(job->out));
  job->out.err = -1;
  for (size_t i = 0; i < nch; i++)
    if (cj[i].out.err >= 0) {
      job->out.err = cj[i].out.err;
      for (size_t j = 0; j < nch; j++) iout_release(&cj[j].out);
      return;
    }
  size_t start = (size_t)node->start, k = (size_t)node->k, u = node->rows.n;
  vi fs = {0};
  for (size_t i = 0; i < k; i++) vi_push(&fs, (int64_t)(start + i));
  int64_t delayed = 0;
  for (size_t i = 0; i < nch; i++) {
    iout *o = &cj[i].out;
    for (size_t q = 0; q < o->nfronts; q++) iout_push(&job->out, o->fronts[q]);
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
  /* dmap sorted by equation */
  size_t nd = kf - k;
  int64_t *dkeys = (int64_t *)ora_xmalloc((nd ? nd : 1) * sizeof(int64_t));
  int64_t *dord = (int64_t *)ora_xmalloc((nd ? nd : 1) * sizeof(int64_t));
  for (size_t i = 0; i < nd; i++) dord[i] = (int64_t)i;
  {
    int64_t *keys = fs.a + k;
    ora_sort_by_ikey(dord, nd, keys);
    for (size_t i = 0; i < nd; i++) dkeys[i] = keys[dord[i]];
  }
  const int64_t *rows = node->rows.a;
#define ILOCAL(g)                                                                              \
  (((size_t)(g) >= start && (size_t)(g) < start + k) ? (size_t)(g) - start                     \
   : ((size_t)(g) < start) ? k + (size_t)dord[bsearch_pair(dkeys, nd, (g))]                    \
                           : kf + (size_t)ora_bsearch(rows, u, (g)))
  double *f = (double *)ora_xcalloc(m * m, sizeof(double));
  for (size_t cc = 0; cc < k; cc++) {
    size_t col = start + cc;
    for (int64_t p = c->colptr[col]; p < c->colptr[col + 1]; p++) f[ILOCAL(c->rowidx[p]) * m + cc] += c->val[p];
  }
  for (size_t i = 0; i < nch; i++) {
    iout *o = &cj[i].out;
    size_t s = o->idx.n;
    size_t *map = (size_t *)ora_xmalloc((s ? s : 1) * sizeof(size_t));
    for (size_t x = 0; x < s; x++) map[x] = ILOCAL(o->idx.a[x]);
    for (size_t x = 0; x < s; x++)
      for (size_t y = 0; y <= x; y++) {
        size_t a = map[x], b = map[y];
        size_t ii = a >= b ? a : b, jj = a >= b ? b : a;
        f[ii * m + jj] += o->upd[x * s + y];
      }
    free(map);
    iout_release(o);
  }
#undef ILOCAL
  free(dkeys);
  free(dord);
  for (size_t i = 0; i < kf; i++)
    for (size_t j = 0; j < i; j++) f[j * m + i] = f[i * m + j];
  dek o = {{0}, {0}, {0}};
  int64_t np = eliminate(f, m, kf, fs.a, c, job->root, job->threads, &o);
  if (np < 0) {
    job->out.err = -np - 1;
    free(f);
    vi_free(&fs);
    vd_free(&o.d);
    vd_free(&o.e);
    vi_free(&o.kind);
    return;
  }
  size_t npiv = (size_t)np;
  ifront F;
  memset(&F, 0, sizeof(F));
  F.l11n = npiv * (npiv ? npiv - 1 : 0) / 2;
  F.l11 = (double *)ora_xmalloc((F.l11n + 1) * sizeof(double));
  size_t off = 0;
  for (size_t i = 0; i < npiv; i++) {
    memcpy(F.l11 + off, f + i * m, i * sizeof(double));
    if (o.kind.a[i] == 0) F.l11[off + i - 1] = 0.0;
    off += i;
  }
  for (size_t i = npiv; i < kf; i++) vi_push(&F.rows, fs.a[i]);
  for (size_t i = 0; i < u; i++) vi_push(&F.rows, rows[i]);
  size_t nr = F.rows.n;
  F.l21n = nr * npiv;
  F.l21 = (double *)ora_xmalloc((F.l21n + 1) * sizeof(double));
  double *upd = (double *)ora_xcalloc(nr * nr + 1, sizeof(double));
  for (size_t r = 0; r < nr; r++) {
    const double *fr = f + (npiv + r) * m;
    memcpy(F.l21 + r * npiv, fr, npiv * sizeof(double));
    memcpy(upd + r * nr, fr + npiv, (r + 1) * sizeof(double));
  }
  for (size_t i = 0; i < npiv; i++) vi_push(&F.piv, fs.a[i]);
  F.d = (double *)ora_xmalloc((npiv + 1) * sizeof(double));
  F.e = (double *)ora_xmalloc((npiv + 1) * sizeof(double));
  F.kind = (uint8_t *)ora_xmalloc(npiv + 1);
  for (size_t i = 0; i < npiv; i++) {
    F.d[i] = o.d.a[i];
    F.e[i] = o.e.a[i];
    F.kind[i] = (uint8_t)o.kind.a[i];
  }
  vd_free(&o.d);
  vd_free(&o.e);
  vi_free(&o.kind);
  free(f);
  vi_free(&fs);
  job->out.idx.n = 0;
  for (size_t i = 0; i < nr; i++) vi_push(&job->out.idx, F.rows.a[i]);
  iout_push(&job->out, F);
  job->out.upd = upd;
  job->out.delayed = delayed;
}

ora_ifactor *ora_factor_indef(const ora_sym *S, const int64_t *colptr, const int64_t *rowidx, const double *val,
                              double threshold, int64_t *err) {
  *err = -1;
  g_forced = 0;
  double thr = threshold < 1e-8 ? 1e-8 : (threshold > 0.5 ? 0.5 : threshold);
  ictx c = {S, colptr, rowidx, val, column_scale(colptr, rowidx, val, S->n), thr};
  ora_ifactor *F = (ora_ifactor *)ora_xcalloc(1, sizeof(ora_ifactor));
  for (size_t i = 0; i < S->roots.n; i++) {
    ijob job;
    memset(&job, 0, sizeof(job));
    job.c = &c;
    job.t = S->roots.a[i];
    job.threads = ora_threads();
    job.root = 1;
    inode(&job);
    if (job.out.err >= 0) {
      *err = job.out.err;
      iout_release(&job.out);
      free(c.scale);
      ora_ifactor_free(F);
      return NULL;
    }
    for (size_t q = 0; q < job.out.nfronts; q++) {
      if (F->nf == F->cap) {
        F->cap = F->cap ? 2 * F->cap : 16;
        F->fr = (ifront *)ora_xrealloc(F->fr, F->cap * sizeof(ifront));
      }
      F->fr[F->nf++] = job.out.fronts[q];
    }
    F->delayed += job.out.delayed;
    free(job.out.fronts);
    free(job.out.upd);
    vi_free(&job.out.idx);
  }
  free(c.scale);
  /* inertia (src/sparse/indef.rs:125-154) */
  for (size_t fi = 0; fi < F->nf; fi++) {
    ifront *fr = &F->fr[fi];
    size_t i = 0;
    while (i < fr->piv.n) {
      if (fr->kind[i] == 2) {
        F->two++;
        double a = fr->d[i], b = fr->e[i], cc = fr->d[i + 1];
        double det = a * cc - b * b;
        if (det < 0.0) {
          F->neg++;
          F->pos++;
        } else if (a + cc < 0.0) {
          F->neg += 2;
        } else {
          F->pos += 2;
        }
        i += 2;
      } else {
        if (fr->d[i] < 0.0) F->neg++;
        else if (fr->d[i] > 0.0) F->pos++;
        else F->zero++;
        i += 1;
      }
    }
  }
  return F;
}

void ora_ifactor_free(ora_ifactor *F) {
  if (!F) return;
  for (size_t i = 0; i < F->nf; i++) ifront_free(&F->fr[i]);
  free(F->fr);
  free(F);
}

void ora_ifactor_stats(const ora_ifactor *F, int64_t *out6) {
  size_t fsz = 0;
  for (size_t i = 0; i < F->nf; i++) fsz += F->fr[i].l11n + F->fr[i].l21n;
  out6[0] = F->neg;
  out6[1] = F->zero;
  out6[2] = F->pos;
  out6[3] = F->delayed;
  out6[4] = F->two;
  out6[5] = (int64_t)fsz;
}

size_t ora_ifactor_pivots(const ora_ifactor *F, int64_t *eq, double *d, double *e, int32_t *kind) {
  size_t c = 0;
  for (size_t fi = 0; fi < F->nf; fi++) {
    const ifront *fr = &F->fr[fi];
    for (size_t i = 0; i < fr->piv.n; i++, c++) {
      eq[c] = fr->piv.a[i];
      d[c] = fr->d[i];
      e[c] = fr->e[i];
      kind[c] = fr->kind[i];
    }
  }
  return c;
}

/* src/sparse/indef.rs:536-591 */
void ora_ifactor_solve(const ora_ifactor *F, double *x) {
  double *xp = NULL;
  size_t cap = 0;
  for (size_t fi = 0; fi < F->nf; fi++) {
    const ifront *fr = &F->fr[fi];
    size_t k = fr->piv.n;
    if (k > cap) { cap = k; xp = (double *)ora_xrealloc(xp, cap * sizeof(double)); }
    for (size_t i = 0; i < k; i++) xp[i] = x[fr->piv.a[i]];
    for (size_t i = 0; i < k; i++) {
      size_t r = i * (i ? i - 1 : 0) / 2;
      double s = ora_dot(fr->l11 + r, xp, i);
      xp[i] -= s;
    }
    for (size_t r = 0; r < fr->rows.n; r++) x[fr->rows.a[r]] -= ora_dot(fr->l21 + r * k, xp, k);
    for (size_t i = 0; i < k; i++) x[fr->piv.a[i]] = xp[i];
  }
  for (size_t fi = 0; fi < F->nf; fi++) {
    const ifront *fr = &F->fr[fi];
    size_t i = 0;
    while (i < fr->piv.n) {
      if (fr->kind[i] == 2) {
        double a = fr->d[i], b = fr->e[i], c = fr->d[i + 1];
        double det = a * c - b * b;
        int64_t g0 = fr->piv.a[i], g1 = fr->piv.a[i + 1];
        double y0 = x[g0], y1 = x[g1];
        x[g0] = (c * y0 - b * y1) / det;
        x[g1] = (a * y1 - b * y0) / det;
        i += 2;
      } else {
        x[fr->piv.a[i]] /= fr->d[i];
        i += 1;
      }
    }
  }
  for (size_t fi = F->nf; fi-- > 0;) {
    const ifront *fr = &F->fr[fi];
    size_t k = fr->piv.n;
    if (k > cap) { cap = k; xp = (double *)ora_xrealloc(xp, cap * sizeof(double)); }
    for (size_t i = 0; i < k; i++) xp[i] = x[fr->piv.a[i]];
    for (size_t r = 0; r < fr->rows.n; r++) {
      double xr = x[fr->rows.a[r]];
      if (xr != 0.0) {
        const double *l = fr->l21 + r * k;
        for (size_t p = 0; p < k; p++) xp[p] -= l[p] * xr;
      }
    }
    for (size_t i = k; i-- > 0;) {
      double xi = xp[i];
      size_t r = i * (i ? i - 1 : 0) / 2;
      for (size_t q = 0; q < i; q++) xp[q] -= fr->l11[r + q] * xi;
    }
    for (size_t i = 0; i < k; i++) x[fr->piv.a[i]] = xp[i];
  }
  free(xp);
}

/* y = A x for the symmetric matrix in the lower triangle (src/sparse/indef.rs:600-611) */
void ora_sym_matvec(size_t n, const int64_t *colptr, const int64_t *rowidx, const double *val, const double *x,
                    double *y) {
  for (size_t i = 0; i < n; i++) y[i] = 0.0;
  for (size_t c = 0; c < n; c++)
    for (int64_t p = colptr[c]; p < colptr[c + 1]; p++) {
      int64_t r = rowidx[p];
      double v = val[p];
      y[r] += v * x[c];
      if ((size_t)r != c) y[c] += v * x[r];
    }
}
