/*
 * Oracle (test infrastructure): linear elastic element stiffness and the
 * sequential element loop with its scatter, restated from the reference.
 *
 *   C3D8  src/elements/continuum3d.rs:39-84 (hex8_ip, hex8_ips), :117-121
 *   C3D4  src/elements/solid_higher.rs:109, :137-189 (ip_at), :453-463;
 *         src/elements/solid_higher_shapes.rs:232 (Tet4 shape), :445-447 (Tet(1))
 *   CPS4, CPE4  src/elements/continuum2d.rs:24-27, :35-80 (ips_2d), :119-123
 *   B, B-bar, D B, B^T D B  src/elements/common.rs:48-70, :142-202, :248-297
 *   D  src/materials/elastic.rs:55-75
 *   scatter  src/fem/assembly.rs:378-398
 *
 * The order of every operation follows the reference (the library is built
 * with -ffp-contract=off).  The sums that the reference writes as iterator
 * sums start here from +0.0; they can differ from the reference's only in the
 * sign of an exact zero.
 *
 * Parity: pinned by closed forms in tests/test_oracle_assembly.py (the
 * Hessian of the tensor-form strain energy, the rank-one volumetric part of
 * the B-bar elements, rigid-body modes, constant-strain energies) and against
 * the independent structured assemblies of rustyfem_amd/meshes.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ora.h"

#define G2 0.5773502691896258

typedef struct {
  int nn, dim, ns, ng, bbar; /* bbar: 0 none, 2 planar, 3 volumetric */
  int plane_stress;
} ora_kind;

static int kind_of(int32_t etype, ora_kind *k) {
  switch (etype) {
    case 1: *k = (ora_kind){8, 3, 6, 8, 3, 0}; return 0;
    case 2: *k = (ora_kind){4, 3, 6, 1, 0, 0}; return 0;
    case 3: *k = (ora_kind){4, 2, 3, 4, 0, 1}; return 0;
    case 4: *k = (ora_kind){4, 2, 4, 4, 2, 0}; return 0;
    default: return -1;
  }
}

/* src/elements/common.rs:48-70 */
static double inv3(const double m[3][3], double ji[3][3]) {
  const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                     m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  const double id = 1.0 / det;
  ji[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
  ji[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
  ji[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
  ji[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
  ji[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
  ji[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
  ji[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
  ji[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
  ji[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
  return det;
}

static const double HEX8_NAT[8][3] = {{-1.0, -1.0, -1.0}, {1.0, -1.0, -1.0}, {1.0, 1.0, -1.0}, {-1.0, 1.0, -1.0},
                                      {-1.0, -1.0, 1.0},  {1.0, -1.0, 1.0},  {1.0, 1.0, 1.0},  {-1.0, 1.0, 1.0}};

/* hex8_ip (src/elements/continuum3d.rs:39-69); returns 0 or -1 (distorted) */
static int hex8_ip(const double x[][3], const double p[3], double wg, double dndx[][3], double *w) {
  double dn[3][8];
  for (int i = 0; i < 8; ++i) {
    const double *nat = HEX8_NAT[i];
    const double f[3] = {1.0 + p[0] * nat[0], 1.0 + p[1] * nat[1], 1.0 + p[2] * nat[2]};
    dn[0][i] = 0.125 * nat[0] * f[1] * f[2];
    dn[1][i] = 0.125 * nat[1] * f[0] * f[2];
    dn[2][i] = 0.125 * nat[2] * f[0] * f[1];
  }
  double j[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int i = 0; i < 8; ++i) s += dn[r][i] * x[i][c];
      j[r][c] = s;
    }
  double ji[3][3];
  const double det = inv3(j, ji);
  if (det <= 0.0) return -1;
  for (int i = 0; i < 8; ++i)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int r = 0; r < 3; ++r) s += ji[c][r] * dn[r][i];
      dndx[i][c] = s;
    }
  *w = det * wg;
  return 0;
}

/* ip_at for Tet4, ThreeD (src/elements/solid_higher.rs:137-181) */
static int tet4_ip(const double x[][3], double gw, double dndx[][3], double *w) {
  /* dN/dxi of (1 - xi - eta - zeta, xi, eta, zeta) (solid_higher_shapes.rs:232) */
  static const double dn[4][3] = {{-1.0, -1.0, -1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  double j[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int i = 0; i < 4; ++i) s += dn[i][r] * x[i][c];
      j[r][c] = s;
    }
  double ji[3][3];
  const double det = inv3(j, ji);
  for (int i = 0; i < 4; ++i)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int r = 0; r < 3; ++r) s += ji[c][r] * dn[i][r];
      dndx[i][c] = s;
    }
  if (!(det > 0.0)) return -1;
  *w = det * gw * 1.0;
  return 0;
}

/* ips_2d, four nodes (src/elements/continuum2d.rs:49-78) */
static int quad4_ips(const double x[][3], double thickness, double dndx[4][4][3], double w[4]) {
  static const double nat[4][2] = {{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0}};
  static const double pts[4][2] = {{-G2, -G2}, {G2, -G2}, {-G2, G2}, {G2, G2}};
  for (int g = 0; g < 4; ++g) {
    const double xi = pts[g][0], eta = pts[g][1];
    double dxi[4], deta[4];
    for (int i = 0; i < 4; ++i) {
      dxi[i] = 0.25 * nat[i][0] * (1.0 + eta * nat[i][1]);
      deta[i] = 0.25 * nat[i][1] * (1.0 + xi * nat[i][0]);
    }
    double j[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int i = 0; i < 4; ++i) {
      j[0][0] += dxi[i] * x[i][0];
      j[0][1] += dxi[i] * x[i][1];
      j[1][0] += deta[i] * x[i][0];
      j[1][1] += deta[i] * x[i][1];
    }
    const double det = j[0][0] * j[1][1] - j[0][1] * j[1][0];
    if (det <= 0.0) return -1;
    for (int i = 0; i < 4; ++i) {
      dndx[g][i][0] = (j[1][1] * dxi[i] - j[0][1] * deta[i]) / det;
      dndx[g][i][1] = (-j[1][0] * dxi[i] + j[0][0] * deta[i]) / det;
      dndx[g][i][2] = 0.0;
    }
    w[g] = det * thickness;
  }
  return 0;
}

/* bmat (src/elements/common.rs:142-174): row-wise ns x nd */
static void bmat(const ora_kind *k, const double dndx[][3], double *b) {
  const int nd = k->dim * k->nn;
  memset(b, 0, sizeof(double) * (size_t)(k->ns * nd));
  for (int i = 0; i < k->nn; ++i) {
    const double *d = dndx[i];
    if (k->dim == 3) {
      const int c = 3 * i;
      b[c] = d[0];
      b[nd + c + 1] = d[1];
      b[2 * nd + c + 2] = d[2];
      b[3 * nd + c] = d[1];
      b[3 * nd + c + 1] = d[0];
      b[4 * nd + c] = d[2];
      b[4 * nd + c + 2] = d[0];
      b[5 * nd + c + 1] = d[2];
      b[5 * nd + c + 2] = d[1];
    } else {
      const int shear = k->ns - 1;
      b[2 * i] = d[0];
      b[nd + 2 * i + 1] = d[1];
      b[shear * nd + 2 * i] = d[1];
      b[shear * nd + 2 * i + 1] = d[0];
    }
  }
}

/* dmat (src/materials/elastic.rs:55-75) */
static void dmat(const ora_kind *k, double e, double nu, double *d) {
  const int n = k->ns;
  memset(d, 0, sizeof(double) * (size_t)(n * n));
  if (k->plane_stress) {
    const double c = e / (1.0 - nu * nu);
    d[0] = c;
    d[1] = c * nu;
    d[3] = c * nu;
    d[4] = c;
    d[8] = c * (1.0 - nu) / 2.0;
    return;
  }
  const double lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = e / (2.0 * (1.0 + nu));
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) d[i * n + j] = lam;
    d[i * n + i] = lam + 2.0 * mu;
  }
  for (int i = 3; i < n; ++i) d[i * n + i] = mu;
}

/* add_btdb (src/elements/common.rs:278-297) */
static void add_btdb(double *ke, const double *b, const double *d, int ns, int nd, double w) {
  double db[6 * 24];
  memset(db, 0, sizeof(double) * (size_t)(ns * nd));
  for (int r = 0; r < ns; ++r)
    for (int m = 0; m < ns; ++m) {
      const double drm = d[r * ns + m];
      if (drm != 0.0)
        for (int c = 0; c < nd; ++c) db[r * nd + c] += drm * b[m * nd + c];
    }
  for (int r = 0; r < ns; ++r)
    for (int i = 0; i < nd; ++i) {
      const double bri = b[r * nd + i] * w;
      if (bri != 0.0)
        for (int j = 0; j < nd; ++j) ke[i * nd + j] += bri * db[r * nd + j];
    }
}

int32_t ora_element_stiffness(int32_t etype, const double *xin, double young, double nu, double thickness,
                              double *ke) {
  ora_kind k;
  if (kind_of(etype, &k)) return -2;
  const int nd = k.dim * k.nn, ns = k.ns;
  double x[8][3];
  for (int i = 0; i < k.nn; ++i)
    for (int c = 0; c < 3; ++c) x[i][c] = xin[3 * i + c];
  double dndx[8][8][3], w[8];
  if (etype == 1) {
    /* hex8_ips: zeta outer, eta, xi inner (continuum3d.rs:72-84) */
    const double gp[2] = {-G2, G2};
    int g = 0;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 2; ++c, ++g) {
          const double p[3] = {gp[c], gp[b], gp[a]};
          if (hex8_ip((const double(*)[3])x, p, 1.0, dndx[g], &w[g])) return -1;
        }
  } else if (etype == 2) {
    if (tet4_ip((const double(*)[3])x, 1.0 / 6.0, dndx[0], &w[0])) return -1;
  } else {
    double d4[4][4][3];
    if (quad4_ips((const double(*)[3])x, thickness, d4, w)) return -1;
    for (int g = 0; g < 4; ++g)
      for (int i = 0; i < 4; ++i)
        for (int c = 0; c < 3; ++c) dndx[g][i][c] = d4[g][i][c];
  }
  /* bmats with B-bar (src/elements/common.rs:177-202) */
  double bs[8][6 * 24];
  for (int g = 0; g < k.ng; ++g) bmat(&k, (const double(*)[3])dndx[g], bs[g]);
  if (k.bbar) {
    double wsum = 0.0;
    for (int g = 0; g < k.ng; ++g) wsum += w[g];
    double avg[24];
    for (int c = 0; c < nd; ++c) {
      double s = 0.0;
      for (int g = 0; g < k.ng; ++g) s += w[g] * (bs[g][c] + bs[g][nd + c] + bs[g][2 * nd + c]);
      avg[c] = s / wsum;
    }
    for (int g = 0; g < k.ng; ++g)
      for (int c = 0; c < nd; ++c) {
        double *b = bs[g];
        const double vol = b[c] + b[nd + c] + b[2 * nd + c];
        if (k.bbar == 2) {
          const double corr = (avg[c] - vol) / 2.0;
          b[c] += corr;
          b[nd + c] += corr;
        } else {
          const double corr = (avg[c] - vol) / 3.0;
          for (int r = 0; r < 3; ++r) b[r * nd + c] += corr;
        }
      }
  }
  /* stiffness (src/elements/common.rs:261-271) */
  double d[36];
  dmat(&k, young, nu, d);
  memset(ke, 0, sizeof(double) * (size_t)(nd * nd));
  for (int g = 0; g < k.ng; ++g) add_btdb(ke, bs[g], d, ns, nd, w[g]);
  return 0;
}

static int64_t find_entry(const int64_t *indptr, const int64_t *indices, int64_t row, int64_t col) {
  int64_t lo = indptr[col], hi = indptr[col + 1];
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (indices[mid] < row)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < indptr[col + 1] && indices[lo] == row) ? lo : -1;
}

int64_t ora_assemble(int32_t etype, int64_t nelem, const int64_t *conn, int64_t nn, const double *coords,
                     const int64_t *eq, int64_t n, const double *young, const double *nu, const double *thickness,
                     const int64_t *indptr, const int64_t *indices, double *data) {
  ora_kind k;
  if (kind_of(etype, &k)) return -3;
  const int nd = k.dim * k.nn;
  (void)nn;
  memset(data, 0, sizeof(double) * (size_t)indptr[n]);
  double ke[24 * 24], x[8 * 3];
  int64_t map[24];
  for (int64_t e = 0; e < nelem; ++e) {
    for (int a = 0; a < k.nn; ++a) {
      const int64_t v = conn[e * k.nn + a];
      for (int c = 0; c < 3; ++c) x[3 * a + c] = coords[3 * v + c];
      for (int s = 0; s < k.dim; ++s) map[k.dim * a + s] = eq[v * k.dim + s];
    }
    if (ora_element_stiffness(etype, x, young[e], nu[e], thickness ? thickness[e] : 1.0, ke)) return e;
    /* src/fem/assembly.rs:378-398: the symmetric part for ia >= ib, nonzero
       values only; the caller's both-triangle pattern gets the mirror too */
    for (int a = 0; a < nd; ++a) {
      const int64_t ia = map[a];
      if (ia < 0) continue;
      for (int b = 0; b < nd; ++b) {
        const int64_t ib = map[b];
        if (ib < 0 || ia < ib) continue;
        const double v = 0.5 * (ke[a * nd + b] + ke[b * nd + a]);
        if (v == 0.0) continue;
        const int64_t p = find_entry(indptr, indices, ia, ib);
        if (p < 0) return -2;
        data[p] += v;
        if (ia != ib) {
          const int64_t q = find_entry(indptr, indices, ib, ia);
          if (q < 0) return -2;
          data[q] += v;
        }
      }
    }
  }
  return -1;
}
