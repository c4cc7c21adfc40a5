/*
 * Oracle: shift-invert block subspace iteration with Rayleigh-Ritz and the
 * cyclic Jacobi method for the projected problems.  Restates
 * src/solver/eigen.rs:159-214 (jacobi), :243-252 (the xorshift start
 * vectors), :258-375 (subspace, lincomb) and :378-387 (normalize), with Op^-1
 * applied by the oracle's PD handle (ora_ffi_solve) one vector at a time,
 * sequential dot products and sums in the reference's loop order.  Test
 * infrastructure only (see ora.h).
 */
#include "ora_internal.h"
#include <math.h>

/* src/solver/eigen.rs:159-214 */
static void jacobi(const double *a0, size_t n, double *d, double *v) {
  double *a = (double *)malloc(n * n * sizeof(double));
  memcpy(a, a0, n * n * sizeof(double));
  for (size_t i = 0; i < n * n; ++i) v[i] = 0.0;
  for (size_t i = 0; i < n; ++i) v[i * n + i] = 1.0;
  double fro = 0.0;
  for (size_t i = 0; i < n * n; ++i) fro += a[i] * a[i];
  double scale = sqrt(fro);
  if (scale < 1e-300) scale = 1e-300;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j)
        if (j != i) off += a[i * n + j] * a[i * n + j];
    if (sqrt(off) <= 1e-15 * scale) break;
    for (size_t p = 0; p < n; ++p)
      for (size_t q = p + 1; q < n; ++q) {
        const double apq = a[p * n + q];
        if (fabs(apq) <= 1e-300) continue;
        const double app = a[p * n + p], aqq = a[q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double sg = theta > 0.0 ? 1.0 : (theta < 0.0 ? -1.0 : 0.0);
        double t = sg / (fabs(theta) + sqrt(theta * theta + 1.0));
        if (theta == 0.0) t = 1.0;
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (size_t k = 0; k < n; ++k) {
          const double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = c * akp - s * akq;
          a[k * n + q] = s * akp + c * akq;
        }
        for (size_t k = 0; k < n; ++k) {
          const double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = c * apk - s * aqk;
          a[q * n + k] = s * apk + c * aqk;
        }
        for (size_t k = 0; k < n; ++k) {
          const double vkp = v[k * n + p], vkq = v[k * n + q];
          v[k * n + p] = c * vkp - s * vkq;
          v[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  for (size_t i = 0; i < n; ++i) d[i] = a[i * n + i];
  free(a);
}

void ora_jacobi_eigen(size_t n, const double *a, double *d, double *v) { jacobi(a, n, d, v); }

static uint64_t rng_state;
static double rng_next(void) {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return (double)(rng_state >> 11) / (double)(1ull << 53) - 0.5;
}

/* y = B x for B in CSC with both triangles (= its CSR) */
static void bmul(size_t n, const int64_t *bp, const int64_t *bi, const double *bv, const double *x, double *y) {
  for (size_t r = 0; r < n; ++r) {
    double s = 0.0;
    for (int64_t p = bp[r]; p < bp[r + 1]; ++p) s += bv[p] * x[bi[p]];
    y[r] = s;
  }
}

static double dot(size_t n, const double *a, const double *b) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* out = sum_i vs[i] w[i r + jj] (src/solver/eigen.rs:366-376) */
static void lincomb(size_t n, double *const *vs, size_t m, const double *w, size_t r, size_t jj, double *out) {
  for (size_t k = 0; k < n; ++k) out[k] = 0.0;
  for (size_t i = 0; i < m; ++i) {
    const double c = w[i * r + jj];
    if (c != 0.0)
      for (size_t k = 0; k < n; ++k) out[k] += c * vs[i][k];
  }
}

static int cmp_abs_desc_idx(const void *a, const void *b, void *ctx) {
  const double *th = (const double *)ctx;
  const size_t i = *(const size_t *)a, j = *(const size_t *)b;
  const double x = fabs(th[i]), y = fabs(th[j]);
  if (x > y) return -1;
  if (x < y) return 1;
  return (i > j) - (i < j);
}

/* insertion sort by |theta| descending, stable (the reference's sort_by) */
static void order_desc(const double *th, size_t r, size_t *order) {
  for (size_t i = 0; i < r; ++i) order[i] = i;
  for (size_t i = 1; i < r; ++i) {
    size_t x = order[i], j = i;
    while (j > 0 && cmp_abs_desc_idx(&order[j - 1], &x, (void *)th) > 0) {
      order[j] = order[j - 1];
      --j;
    }
    order[j] = x;
  }
}

static double **alloc_vecs(size_t count, size_t n) {
  double **v = (double **)calloc(count ? count : 1, sizeof(double *));
  for (size_t i = 0; i < count; ++i) v[i] = (double *)calloc(n ? n : 1, sizeof(double));
  return v;
}
static void free_vecs(double **v, size_t count) {
  if (!v) return;
  for (size_t i = 0; i < count; ++i) free(v[i]);
  free(v);
}

/* src/solver/eigen.rs:258-364 (subspace) and :378-387 (normalize).  Returns
 * the number of pairs found, -1 (no convergence), -2 (B zero on the block). */
int64_t ora_eigen(const ora_handle *h, size_t n, const int64_t *bp, const int64_t *bi, const double *bv, size_t nev,
                  size_t vectors, size_t max_iter, double tol, double *theta_out, double *x_out,
                  int64_t *iterations) {
  if (n == 0 || nev == 0) return 0;
  if (nev > n) nev = n;
  size_t q = vectors > 0 ? (vectors > nev ? vectors : nev) : (2 * nev < nev + 8 ? 2 * nev : nev + 8);
  if (q > n) q = n;
  rng_state = 0x9E3779B97F4A7C15ull;
  double **y = alloc_vecs(q, n), **xb = alloc_vecs(q, n), **z = alloc_vecs(q, n);
  double **x = alloc_vecs(q, n), **opx = alloc_vecs(q, n), **bx = alloc_vecs(q, n);
  double *tmp = (double *)malloc(n * sizeof(double));
  double *theta = (double *)calloc(q, sizeof(double));
  size_t my = q, nx = 0;
  int64_t found = -1;
  for (size_t j = 0; j < q; ++j) {
    for (size_t i = 0; i < n; ++i) tmp[i] = rng_next();
    bmul(n, bp, bi, bv, tmp, y[j]);
  }
  for (size_t it = 1; it <= max_iter; ++it) {
    for (size_t j = 0; j < my; ++j) {
      memcpy(xb[j], y[j], n * sizeof(double));
      ora_ffi_solve(h, 1, xb[j]);
    }
    if (nx > 0) {
      const size_t k = nev < nx ? nev : nx;
      int ok = 1;
      for (size_t i = 0; i < k; ++i) {
        double rr = 0.0;
        for (size_t j = 0; j < n; ++j) rr += (xb[i][j] - theta[i] * x[i][j]) * (y[i][j] - theta[i] * opx[i][j]);
        if (!(sqrt(fabs(rr)) / fabs(theta[i]) <= tol)) ok = 0;
      }
      if (ok && (nx >= nev || it > 5)) {
        found = (int64_t)k;
        if (iterations) *iterations = (int64_t)(it - 1);
        break;
      }
    }
    const size_t m = my;
    for (size_t j = 0; j < m; ++j) bmul(n, bp, bi, bv, xb[j], z[j]);
    double *g = (double *)calloc(m * m, sizeof(double)), *a = (double *)calloc(m * m, sizeof(double));
    for (size_t i = 0; i < m; ++i)
      for (size_t j = 0; j <= i; ++j) {
        const double gij = 0.5 * (dot(n, xb[i], y[j]) + dot(n, xb[j], y[i]));
        const double aij = 0.5 * (dot(n, xb[i], z[j]) + dot(n, xb[j], z[i]));
        g[i * m + j] = g[j * m + i] = gij;
        a[i * m + j] = a[j * m + i] = aij;
      }
    double gmax = 0.0;
    for (size_t i = 0; i < m; ++i)
      if (g[i * m + i] > gmax) gmax = g[i * m + i];
    double *sc = (double *)calloc(m, sizeof(double));
    for (size_t i = 0; i < m; ++i)
      sc[i] = g[i * m + i] > 1e-300 * (gmax > 1e-300 ? gmax : 1e-300) ? 1.0 / sqrt(g[i * m + i]) : 0.0;
    for (size_t i = 0; i < m; ++i)
      for (size_t j = 0; j < m; ++j) {
        g[i * m + j] *= sc[i] * sc[j];
        a[i * m + j] *= sc[i] * sc[j];
      }
    double *d = (double *)calloc(m, sizeof(double)), *u = (double *)calloc(m * m, sizeof(double));
    jacobi(g, m, d, u);
    double dmax = 0.0;
    for (size_t j = 0; j < m; ++j)
      if (d[j] > dmax) dmax = d[j];
    if (dmax <= 0.0 || gmax <= 0.0) {
      free(g); free(a); free(sc); free(d); free(u);
      found = -2;
      break;
    }
    size_t *keep = (size_t *)calloc(m, sizeof(size_t)), r = 0;
    for (size_t j = 0; j < m; ++j)
      if (d[j] > 1e-13 * dmax) keep[r++] = j;
    double *t = (double *)calloc(m * r + 1, sizeof(double)), *at = (double *)calloc(m * r + 1, sizeof(double));
    double *c = (double *)calloc(r * r + 1, sizeof(double));
    for (size_t jj = 0; jj < r; ++jj) {
      const double s = 1.0 / sqrt(d[keep[jj]]);
      for (size_t i = 0; i < m; ++i) t[i * r + jj] = u[i * m + keep[jj]] * s;
    }
    for (size_t i = 0; i < m; ++i)
      for (size_t jj = 0; jj < r; ++jj) {
        double s = 0.0;
        for (size_t k = 0; k < m; ++k) s += a[i * m + k] * t[k * r + jj];
        at[i * r + jj] = s;
      }
    for (size_t ii = 0; ii < r; ++ii)
      for (size_t jj = 0; jj < r; ++jj) {
        double s = 0.0;
        for (size_t k = 0; k < m; ++k) s += t[k * r + ii] * at[k * r + jj];
        c[ii * r + jj] = s;
      }
    for (size_t ii = 0; ii < r; ++ii)
      for (size_t jj = 0; jj < ii; ++jj) {
        const double s = 0.5 * (c[ii * r + jj] + c[jj * r + ii]);
        c[ii * r + jj] = s;
        c[jj * r + ii] = s;
      }
    double *th = (double *)calloc(r + 1, sizeof(double)), *v = (double *)calloc(r * r + 1, sizeof(double));
    jacobi(c, r, th, v);
    size_t *order = (size_t *)calloc(r + 1, sizeof(size_t));
    order_desc(th, r, order);
    double *w = (double *)calloc(m * r + 1, sizeof(double));
    for (size_t i = 0; i < m; ++i)
      for (size_t jj = 0; jj < r; ++jj) {
        double s = 0.0;
        for (size_t k = 0; k < r; ++k) s += t[i * r + k] * v[k * r + order[jj]];
        w[i * r + jj] = sc[i] * s;
      }
    for (size_t jj = 0; jj < r; ++jj) {
      lincomb(n, xb, m, w, r, jj, x[jj]);
      lincomb(n, y, m, w, r, jj, opx[jj]);
      lincomb(n, z, m, w, r, jj, bx[jj]);
      theta[jj] = th[order[jj]];
    }
    nx = r;
    for (size_t jj = 0; jj < r; ++jj) memcpy(y[jj], bx[jj], n * sizeof(double));
    my = r;
    while (my < q) {
      for (size_t i = 0; i < n; ++i) tmp[i] = rng_next();
      bmul(n, bp, bi, bv, tmp, y[my++]);
    }
    if (iterations) *iterations = (int64_t)it;
    free(g); free(a); free(sc); free(d); free(u); free(keep); free(t); free(at); free(c); free(th); free(v);
    free(order); free(w);
  }
  if (found > 0) {
    for (size_t i = 0; i < (size_t)found; ++i) {
      bmul(n, bp, bi, bv, x[i], tmp);
      const double s = dot(n, x[i], tmp) / theta[i];
      double f = 1.0;
      if (s > 0.0) f = 1.0 / sqrt(s);
      theta_out[i] = theta[i];
      if (x_out)
        for (size_t k = 0; k < n; ++k) x_out[i * n + k] = x[i][k] * f;
    }
  }
  free_vecs(y, q); free_vecs(xb, q); free_vecs(z, q); free_vecs(x, q); free_vecs(opx, q); free_vecs(bx, q);
  free(tmp);
  free(theta);
  return found;
}
