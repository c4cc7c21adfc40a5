/*
 * Oracle: the C interface of ffi/src/lib.rs (input validation, uncoupled
 * DOF split, node compaction, nested dissection with leaf 16, entry filter,
 * solve with permutation), restated with an ora_ prefix.
 * Test infrastructure only (see ora.h).
 */
#include "ora_internal.h"
#include <math.h>

static size_t g_ffi_leaf = 16; /* ffi/src/lib.rs:17 */
size_t ora_ffi_leaf(void) { return g_ffi_leaf; }
void ora_set_ffi_leaf(size_t leaf) { g_ffi_leaf = leaf ? leaf : 16; }

enum { MODE_PD = 0, MODE_INDEF = 1, MODE_LU = 2 };

struct ora_handle {
  int mode;
  ora_sym *sym;
  ora_factor *pd;
  ora_ifactor *ind;
  ora_lufactor *lu;
  size_t n;
  int64_t *eq; /* -1 = uncoupled */
  int64_t *iso_i;
  double *iso_d;
  size_t niso;
  size_t n_sym;
  double work;
  size_t factor_size;
  int has_inertia;
  int64_t inertia[3];
};

static void handle_free(ora_handle *h) {
  if (!h) return;
  ora_factor_free(h->pd);
  ora_ifactor_free(h->ind);
  ora_lufactor_free(h->lu);
  ora_sym_free(h->sym);
  free(h->eq);
  free(h->iso_i);
  free(h->iso_d);
  free(h);
}

/* ffi/src/lib.rs:100-229; returns handle or NULL with *err set; d_out filled for PD */
static ora_handle *build(size_t n, size_t nn, const int64_t *loc, const double *coords, const int64_t *indptr,
                         const int64_t *indices, const double *data, int mode, double *d_out, int64_t *err) {
  *err = -1;
  size_t nnz = (size_t)indptr[n];
  for (size_t i = 0; i < n; i++)
    if (loc[i] < 0 || (size_t)loc[i] >= nn) return NULL;
  for (size_t p = 0; p < nnz; p++)
    if (indices[p] < 0 || (size_t)indices[p] >= n) return NULL;
  uint8_t *coupled = (uint8_t *)ora_xcalloc(n ? n : 1, 1);
  double *diag = (double *)ora_xcalloc(n ? n : 1, sizeof(double));
  for (size_t c = 0; c < n; c++)
    for (int64_t p = indptr[c]; p < indptr[c + 1]; p++) {
      size_t r = (size_t)indices[p];
      if (r == c) diag[c] += data[p];
      else if (data[p] != 0.0) {
        coupled[c] = 1;
        coupled[r] = 1;
      }
    }
  ora_handle *h = (ora_handle *)ora_xcalloc(1, sizeof(ora_handle));
  h->mode = mode;
  h->n = n;
  h->iso_i = (int64_t *)ora_xmalloc((n ? n : 1) * sizeof(int64_t));
  h->iso_d = (double *)ora_xmalloc((n ? n : 1) * sizeof(double));
  for (size_t i = 0; i < n; i++) {
    if (coupled[i]) continue;
    int ok = mode == MODE_PD ? (diag[i] > 0.0) : (diag[i] != 0.0 && isfinite(diag[i]));
    if (!ok) {
      *err = (int64_t)i;
      free(coupled);
      free(diag);
      handle_free(h);
      return NULL;
    }
    h->iso_i[h->niso] = (int64_t)i;
    h->iso_d[h->niso] = diag[i];
    h->niso++;
  }
  int64_t *nid = (int64_t *)ora_xmalloc((nn ? nn : 1) * sizeof(int64_t));
  for (size_t v = 0; v < nn; v++) nid[v] = -1;
  vd xyz = {0};
  size_t nc = 0;
  for (size_t i = 0; i < n; i++) {
    if (!coupled[i]) continue;
    int64_t v = loc[i];
    if (nid[v] < 0) {
      nid[v] = (int64_t)nc++;
      vd_push(&xyz, coords[3 * v]);
      vd_push(&xyz, coords[3 * v + 1]);
      vd_push(&xyz, coords[3 * v + 2]);
    }
  }
  int64_t *ndof = (int64_t *)ora_xcalloc(nc ? nc : 1, sizeof(int64_t));
  int64_t *local = (int64_t *)ora_xcalloc(n ? n : 1, sizeof(int64_t));
  for (size_t i = 0; i < n; i++) {
    if (!coupled[i]) continue;
    int64_t v = nid[loc[i]];
    local[i] = ndof[v];
    ndof[v]++;
  }
  /* adjacency: collect pairs, sort and dedup per node */
  vi *adj = (vi *)ora_xcalloc(nc ? nc : 1, sizeof(vi));
  for (size_t c = 0; c < n; c++) {
    if (!coupled[c]) continue;
    for (int64_t p = indptr[c]; p < indptr[c + 1]; p++) {
      size_t r = (size_t)indices[p];
      if (coupled[r] && nid[loc[r]] != nid[loc[c]]) {
        vi_push(&adj[nid[loc[c]]], nid[loc[r]]);
        vi_push(&adj[nid[loc[r]]], nid[loc[c]]);
      }
    }
  }
  int64_t *aptr = (int64_t *)ora_xcalloc(nc + 1, sizeof(int64_t));
  for (size_t v = 0; v < nc; v++) {
    ora_sort_i64(adj[v].a, adj[v].n);
    adj[v].n = ora_dedup_i64(adj[v].a, adj[v].n);
    aptr[v + 1] = aptr[v] + (int64_t)adj[v].n;
  }
  int64_t *aidx = (int64_t *)ora_xmalloc(((size_t)aptr[nc] + 1) * sizeof(int64_t));
  for (size_t v = 0; v < nc; v++) {
    memcpy(aidx + aptr[v], adj[v].a, adj[v].n * sizeof(int64_t));
    vi_free(&adj[v]);
  }
  free(adj);
  ora_sym *sym = ora_analyse(nc, xyz.a, aptr, aidx, ndof, g_ffi_leaf);
  vd_free(&xyz);
  h->sym = sym;
  h->eq = (int64_t *)ora_xmalloc((n ? n : 1) * sizeof(int64_t));
  for (size_t i = 0; i < n; i++) h->eq[i] = coupled[i] ? sym->first_dof[nid[loc[i]]] + local[i] : -1;
  size_t ns = sym->n;
  int64_t *owner = (int64_t *)ora_xcalloc(ns ? ns : 1, sizeof(int64_t));
  for (size_t i = 0; i < n; i++)
    if (h->eq[i] >= 0) owner[h->eq[i]] = (int64_t)i;
  size_t pnnz = ora_sym_pattern(sym, aptr, aidx, ndof, NULL, NULL);
  int64_t *colptr = (int64_t *)ora_xmalloc((ns + 1) * sizeof(int64_t));
  int64_t *rowidx = (int64_t *)ora_xmalloc((pnnz + 1) * sizeof(int64_t));
  ora_sym_pattern(sym, aptr, aidx, ndof, colptr, rowidx);
  colptr[0] = 0;
  free(aptr);
  free(aidx);
  free(ndof);
  free(local);
  free(nid);
  double *val = (double *)ora_xcalloc(pnnz + 1, sizeof(double));
  double *uval = mode == MODE_LU ? (double *)ora_xcalloc(pnnz + 1, sizeof(double)) : NULL;
  int bad = 0;
#define LADD(arr, r, c, v)                                                               \
  do {                                                                                   \
    int64_t q_ = ora_bsearch(rowidx + colptr[(c)], (size_t)(colptr[(c) + 1] - colptr[(c)]), (r)); \
    if (q_ < 0) bad = 1; else (arr)[colptr[(c)] + q_] += (v);                           \
  } while (0)
  for (size_t c = 0; c < n && !bad; c++) {
    if (!coupled[c]) continue;
    int64_t qc = h->eq[c];
    for (int64_t p = indptr[c]; p < indptr[c + 1]; p++) {
      size_t r = (size_t)indices[p];
      int64_t qr = h->eq[r];
      if (qr < 0) continue;
      if (mode == MODE_LU) {
        if (qr >= qc) LADD(val, qr, qc, data[p]);
        else LADD(uval, qc, qr, data[p]);
      } else if (mode == MODE_INDEF) {
        if (r >= c) LADD(val, qr > qc ? qr : qc, qr > qc ? qc : qr, data[p]);
      } else {
        if (qr >= qc) LADD(val, qr, qc, data[p]);
      }
    }
  }
#undef LADD
  h->n_sym = ns;
  h->work = ora_sym_work(sym);
  int64_t e = -1;
  if (!bad) {
    if (mode == MODE_LU) {
      h->lu = ora_factor_lu(sym, colptr, rowidx, val, uval, 0.01, &e);
      if (h->lu) {
        int64_t st[2];
        ora_lufactor_stats(h->lu, st);
        h->factor_size = (size_t)st[1];
      }
    } else if (mode == MODE_INDEF) {
      h->ind = ora_factor_indef(sym, colptr, rowidx, val, 0.01, &e);
      if (h->ind) {
        int64_t st[6];
        ora_ifactor_stats(h->ind, st);
        h->factor_size = (size_t)st[5];
        h->has_inertia = 1;
        h->inertia[0] = st[0];
        h->inertia[1] = st[1];
        h->inertia[2] = st[2];
        for (size_t i = 0; i < h->niso; i++) {
          if (h->iso_d[i] < 0.0) h->inertia[0]++; else h->inertia[2]++;
        }
      }
    } else {
      h->pd = ora_factor_pd(sym, colptr, rowidx, val, &e);
      if (h->pd) {
        h->factor_size = ora_sym_factor_size(sym);
        h->has_inertia = 1;
        h->inertia[0] = 0;
        h->inertia[1] = 0;
        h->inertia[2] = (int64_t)n;
        if (d_out) {
          double *dq = (double *)ora_xmalloc((ns + 1) * sizeof(double));
          ora_factor_pivots(h->pd, dq);
          for (size_t i = 0; i < n; i++) d_out[i] = coupled[i] ? dq[h->eq[i]] : diag[i];
          free(dq);
        }
      }
    }
  }
  free(colptr);
  free(rowidx);
  free(val);
  free(uval);
  free(coupled);
  free(diag);
  if (bad || (!h->pd && !h->ind && !h->lu)) {
    *err = bad ? -1 : owner[e];
    free(owner);
    handle_free(h);
    return NULL;
  }
  free(owner);
  return h;
}

ora_handle *ora_ffi_factor(size_t n, size_t nn, const int64_t *loc, const double *coords, const int64_t *indptr,
                           const int64_t *indices, const double *data, double *d_out, int64_t *err_dof) {
  return build(n, nn, loc, coords, indptr, indices, data, MODE_PD, d_out, err_dof);
}

ora_handle *ora_ffi_factor_indef(size_t n, size_t nn, const int64_t *loc, const double *coords,
                                 const int64_t *indptr, const int64_t *indices, const double *data,
                                 int64_t *inertia_out, int64_t *err_dof) {
  ora_handle *h = build(n, nn, loc, coords, indptr, indices, data, MODE_INDEF, NULL, err_dof);
  if (h && inertia_out) memcpy(inertia_out, h->inertia, 3 * sizeof(int64_t));
  return h;
}

ora_handle *ora_ffi_factor_unsym(size_t n, size_t nn, const int64_t *loc, const double *coords,
                                 const int64_t *indptr, const int64_t *indices, const double *data,
                                 int64_t *err_dof) {
  return build(n, nn, loc, coords, indptr, indices, data, MODE_LU, NULL, err_dof);
}

int32_t ora_ffi_inertia(const ora_handle *h, int64_t *out) {
  if (!h->has_inertia) return -1;
  memcpy(out, h->inertia, 3 * sizeof(int64_t));
  return 0;
}

/* ffi/src/lib.rs:336-371 (sequential over right-hand sides; same arithmetic per column) */
int32_t ora_ffi_solve(const ora_handle *h, size_t nrhs, double *b) {
  size_t n = h->n;
  if (n == 0 || nrhs == 0) return 0;
  double *x = (double *)ora_xcalloc(h->n_sym + 1, sizeof(double));
  for (size_t j = 0; j < nrhs; j++) {
    double *col = b + j * n;
    memset(x, 0, h->n_sym * sizeof(double));
    for (size_t i = 0; i < n; i++)
      if (h->eq[i] >= 0) x[h->eq[i]] = col[i];
    if (h->pd) ora_factor_solve(h->pd, x);
    else if (h->ind) ora_ifactor_solve(h->ind, x);
    else ora_lufactor_solve(h->lu, x);
    for (size_t i = 0; i < n; i++)
      if (h->eq[i] >= 0) col[i] = x[h->eq[i]];
    for (size_t q = 0; q < h->niso; q++) col[h->iso_i[q]] /= h->iso_d[q];
  }
  free(x);
  return 0;
}

void ora_ffi_info(const ora_handle *h, double *out) {
  out[0] = h->work;
  out[1] = (double)h->factor_size;
}

void ora_ffi_order(const ora_handle *h, int64_t *out) {
  int64_t next = (int64_t)h->n_sym;
  for (size_t i = 0; i < h->n; i++) out[i] = h->eq[i] < 0 ? next++ : h->eq[i];
}

void ora_ffi_free(ora_handle *h) { handle_free(h); }

size_t ora_ffi_pivots(const ora_handle *h, int64_t *eq, double *d, double *e, int32_t *kind) {
  if (h->ind) return ora_ifactor_pivots(h->ind, eq, d, e, kind);
  if (h->lu) return ora_lufactor_pivots(h->lu, eq, d, e, kind);
  ora_factor_pivots(h->pd, d);
  for (size_t i = 0; i < h->n_sym; i++) {
    eq[i] = (int64_t)i;
    e[i] = 0.0;
    kind[i] = 1;
  }
  return h->n_sym;
}

void ora_ffi_stats(const ora_handle *h, int64_t *out6) {
  memset(out6, 0, 6 * sizeof(int64_t));
  if (h->ind) {
    ora_ifactor_stats(h->ind, out6);
  } else if (h->lu) {
    int64_t st[2];
    ora_lufactor_stats(h->lu, st);
    out6[3] = st[0];
    out6[5] = st[1];
  } else {
    out6[2] = (int64_t)h->n_sym;
    out6[5] = (int64_t)h->factor_size;
  }
}

const ora_sym *ora_ffi_sym(const ora_handle *h) { return h ? h->sym : NULL; }
