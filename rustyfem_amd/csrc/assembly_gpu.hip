// Element assembly on one MI355X (interface and scope: assembly.hpp).
//
// Compiled without floating-point contraction and written so that every
// element-matrix entry is formed by the operations of the reference in their
// order (integration point by point, the sums from zero, the divisions and
// the B-bar correction as written there); each assembled value then sums the
// symmetric parts 0.5 (K_e[a][b] + K_e[b][a]) of its elements from zero in
// the order of the reference's element loop (src/fem/assembly.rs:378-398).
// The result equals a sequential CPU evaluation of the reference's formulas
// bit for bit (up to the sign of exact zeros, which the reference's skipping
// of zero terms leaves open).
//
// Layout: an element takes NN x NN lanes of a wave (a hexahedron a whole
// wave, four tetrahedra or quadrilaterals share one); lane (a, b) forms the
// DIM x DIM block K_e[DIM a.., DIM b..] from B w and D B held in LDS, trades
// it with lane (b, a) for the transposed block, and the lanes a <= b store
// the symmetric part of their block.  A stored element matrix is its
// NN (NN + 1) / 2 upper node-pair blocks of DIM x DIM doubles, in row-major
// pair order.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "assembly.hpp"

namespace rf {

namespace {

#define AHIPCHK(expr)                                                                                        \
  do {                                                                                                       \
    hipError_t e_ = (expr);                                                                                  \
    if (e_ != hipSuccess) {                                                                                  \
      char b_[512];                                                                                          \
      snprintf(b_, sizeof b_, "HIP error %s at %s:%d: %s", hipGetErrorString(e_), __FILE__, __LINE__, #expr); \
      throw std::runtime_error(b_);                                                                          \
    }                                                                                                        \
  } while (0)

// 1 / sqrt(3), the two-point Gauss abscissa (src/elements/common.rs:13)
constexpr double G2 = 0.5773502691896258;

// Element families.  KIND: 0 C3D8, 1 C3D4, 2 CPS4, 3 CPE4.
template <int KIND>
struct Elem;
template <>
struct Elem<0> {
  static constexpr int NN = 8, DIM = 3, NS = 6, NG = 8, BBAR = 3;
};
template <>
struct Elem<1> {
  static constexpr int NN = 4, DIM = 3, NS = 6, NG = 1, BBAR = 0;
};
template <>
struct Elem<2> {
  static constexpr int NN = 4, DIM = 2, NS = 3, NG = 4, BBAR = 0;
};
template <>
struct Elem<3> {
  static constexpr int NN = 4, DIM = 2, NS = 4, NG = 4, BBAR = 2;
};

// Is row r of B (after B-bar) structurally nonzero in the column of
// component k?  (src/elements/common.rs:142-202)
template <int KIND>
__device__ constexpr bool nz_b(int r, int k) {
  if (Elem<KIND>::DIM == 3) {
    if (r < 3) return Elem<KIND>::BBAR ? true : r == k;
    if (r == 3) return k != 2;
    if (r == 4) return k != 1;
    return k != 0;
  }
  if (KIND == 2) return r < 2 ? r == k : true;  // plane stress: e11, e22, g12
  return r == 2 ? false : true;                 // plane strain, planar B-bar: e11, e22, e33 = 0, g12
}

// Is row r of D B structurally nonzero in the column of component k?  The
// normal rows mix all normal strains; a shear row is the shear row of B.
template <int KIND>
__device__ constexpr bool nz_db(int r, int k) {
  if (Elem<KIND>::DIM == 3) return r < 3 ? true : nz_b<KIND>(r, k);
  return true;
}

__device__ void inv3(const double m[3][3], double ji[3][3], double& det) {
  // src/elements/common.rs:48-70
  det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
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
}

// dN/dx of integration point g and its weight; false when the Jacobian
// determinant fails the reference's test.
template <int KIND>
__device__ bool ip_point(int g, const double (*x)[3], double thick, double (*dndx)[3], double& w);

template <>
__device__ bool ip_point<0>(int g, const double (*x)[3], double, double (*dndx)[3], double& w) {
  // hex8_ip at the 2x2x2 points, xi fastest (src/elements/continuum3d.rs:39-84)
  const double nat[8][3] = {{-1.0, -1.0, -1.0}, {1.0, -1.0, -1.0}, {1.0, 1.0, -1.0}, {-1.0, 1.0, -1.0},
                            {-1.0, -1.0, 1.0},  {1.0, -1.0, 1.0},  {1.0, 1.0, 1.0},  {-1.0, 1.0, 1.0}};
  const double p[3] = {(g & 1) ? G2 : -G2, (g & 2) ? G2 : -G2, (g & 4) ? G2 : -G2};
  double dn[3][8];
  for (int i = 0; i < 8; ++i) {
    const double f0 = 1.0 + p[0] * nat[i][0], f1 = 1.0 + p[1] * nat[i][1], f2 = 1.0 + p[2] * nat[i][2];
    dn[0][i] = 0.125 * nat[i][0] * f1 * f2;
    dn[1][i] = 0.125 * nat[i][1] * f0 * f2;
    dn[2][i] = 0.125 * nat[i][2] * f0 * f1;
  }
  double j[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int i = 0; i < 8; ++i) s = s + dn[r][i] * x[i][c];
      j[r][c] = s;
    }
  double ji[3][3], det;
  inv3(j, ji, det);
  for (int i = 0; i < 8; ++i)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int r = 0; r < 3; ++r) s = s + ji[c][r] * dn[r][i];
      dndx[i][c] = s;
    }
  w = det * 1.0;
  return !(det <= 0.0);
}

template <>
__device__ bool ip_point<1>(int, const double (*x)[3], double, double (*dndx)[3], double& w) {
  // ip_at for Tet4 at (1/4, 1/4, 1/4), weight 1/6 (src/elements/solid_higher.rs:137-181,
  // src/elements/solid_higher_shapes.rs:232, :445-447): dN/dxi of
  // (1 - xi - eta - zeta, xi, eta, zeta)
  const double dn[4][3] = {{-1.0, -1.0, -1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  double j[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int i = 0; i < 4; ++i) s = s + dn[i][r] * x[i][c];
      j[r][c] = s;
    }
  double ji[3][3], det;
  inv3(j, ji, det);
  for (int i = 0; i < 4; ++i)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int r = 0; r < 3; ++r) s = s + ji[c][r] * dn[i][r];
      dndx[i][c] = s;
    }
  w = det * (1.0 / 6.0) * 1.0;
  return det > 0.0;
}

__device__ bool quad_point(int g, const double (*x)[3], double thick, double (*dndx)[3], double& w) {
  // ips_2d for four nodes, xi fastest (src/elements/continuum2d.rs:49-78)
  const double nat[4][2] = {{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0}};
  const double xi = (g & 1) ? G2 : -G2, eta = (g & 2) ? G2 : -G2;
  double dxi[4], deta[4];
  for (int i = 0; i < 4; ++i) {
    dxi[i] = 0.25 * nat[i][0] * (1.0 + eta * nat[i][1]);
    deta[i] = 0.25 * nat[i][1] * (1.0 + xi * nat[i][0]);
  }
  double j00 = 0.0, j01 = 0.0, j10 = 0.0, j11 = 0.0;
  for (int i = 0; i < 4; ++i) {
    j00 = j00 + dxi[i] * x[i][0];
    j01 = j01 + dxi[i] * x[i][1];
    j10 = j10 + deta[i] * x[i][0];
    j11 = j11 + deta[i] * x[i][1];
  }
  const double det = j00 * j11 - j01 * j10;
  for (int i = 0; i < 4; ++i) {
    dndx[i][0] = (j11 * dxi[i] - j01 * deta[i]) / det;
    dndx[i][1] = (-j10 * dxi[i] + j00 * deta[i]) / det;
    dndx[i][2] = 0.0;
  }
  w = det * thick;
  return !(det <= 0.0);
}

template <>
__device__ bool ip_point<2>(int g, const double (*x)[3], double thick, double (*dndx)[3], double& w) {
  return quad_point(g, x, thick, dndx, w);
}
template <>
__device__ bool ip_point<3>(int g, const double (*x)[3], double thick, double (*dndx)[3], double& w) {
  return quad_point(g, x, thick, dndx, w);
}

// Row r of B (before B-bar) in the column of component k of a node with
// derivatives d (src/elements/common.rs:142-174).
template <int KIND>
__device__ double bentry(int r, int k, const double* d) {
  if (Elem<KIND>::DIM == 3) {
    switch (r) {
      case 0: return k == 0 ? d[0] : 0.0;
      case 1: return k == 1 ? d[1] : 0.0;
      case 2: return k == 2 ? d[2] : 0.0;
      case 3: return k == 0 ? d[1] : (k == 1 ? d[0] : 0.0);
      case 4: return k == 0 ? d[2] : (k == 2 ? d[0] : 0.0);
      default: return k == 1 ? d[2] : (k == 2 ? d[1] : 0.0);
    }
  }
  constexpr int shear = Elem<KIND>::NS - 1;
  if (r == 0) return k == 0 ? d[0] : 0.0;
  if (r == 1) return k == 1 ? d[1] : 0.0;
  if (r == shear) return k == 0 ? d[1] : d[0];
  return 0.0;
}

// D[r][m] of the isotropic material (src/materials/elastic.rs:55-75):
// plane stress for CPS4, the 3-D form (ncomp 6 or 4) otherwise
template <int KIND>
struct Dmat {
  double v[Elem<KIND>::NS][Elem<KIND>::NS];
  __device__ Dmat(double e, double nu) {
    constexpr int ns = Elem<KIND>::NS;
    for (int r = 0; r < ns; ++r)
      for (int m = 0; m < ns; ++m) v[r][m] = 0.0;
    if (KIND == 2) {
      const double c = e / (1.0 - nu * nu);
      v[0][0] = c;
      v[0][1] = c * nu;
      v[1][0] = c * nu;
      v[1][1] = c;
      v[2][2] = c * (1.0 - nu) / 2.0;
    } else {
      const double lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
      const double mu = e / (2.0 * (1.0 + nu));
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) v[i][j] = lam;
        v[i][i] = lam + 2.0 * mu;
      }
      for (int i = 3; i < ns; ++i) v[i][i] = mu;
    }
  }
};

template <int KIND>
struct Lds {
  static constexpr int NN = Elem<KIND>::NN, DIM = Elem<KIND>::DIM, NS = Elem<KIND>::NS, NG = Elem<KIND>::NG;
  static constexpr int ND = NN * DIM;
  double x[NN][3];
  double dndx[NG][NN][3];
  double w[NG];
  double avg[ND];
  double bw[NG][NS][ND];  // B w
  double db[NG][NS][ND];  // D B
};

// Element stiffness blocks: 64-lane workgroups, 64 / NN^2 elements each.
template <int KIND>
__global__ __launch_bounds__(64) void k_elem(const int32_t* __restrict__ conn, const double* __restrict__ coords,
                                             const double* __restrict__ young, const double* __restrict__ pois,
                                             const double* __restrict__ thick, int64_t nelem,
                                             double* __restrict__ ke, unsigned long long* __restrict__ bad) {
  using E = Elem<KIND>;
  constexpr int NN = E::NN, DIM = E::DIM, NS = E::NS, NG = E::NG, ND = NN * DIM;
  constexpr int LPE = NN * NN, EPW = 64 / LPE, NPAIR = NN * (NN + 1) / 2;
  __shared__ Lds<KIND> S[EPW];
  const int t = threadIdx.x, slot = t / LPE, l = t % LPE;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * EPW + slot;
  const bool live = e < nelem;
  const int64_t ec = live ? e : nelem - 1;  // lanes past the end mirror the last element and store nothing
  Lds<KIND>& s = S[slot];
  if (l < NN * 3) s.x[l / 3][l % 3] = coords[static_cast<int64_t>(conn[ec * NN + l / 3]) * 3 + l % 3];
  __syncthreads();
  if (l < NG) {
    double w;
    const bool ok = ip_point<KIND>(l, s.x, thick[ec], s.dndx[l], w);
    s.w[l] = w;
    if (!ok && live) atomicMin(bad, static_cast<unsigned long long>(e));
  }
  __syncthreads();
  if (E::BBAR && l < ND) {
    // B-bar (src/elements/common.rs:177-202): the w-weighted mean of the
    // volumetric row sum over the points; the row sum of a column is the
    // one derivative it holds
    double wsum = 0.0;
    for (int g = 0; g < NG; ++g) wsum = wsum + s.w[g];
    double acc = 0.0;
    for (int g = 0; g < NG; ++g) acc = acc + s.w[g] * s.dndx[g][l / DIM][l % DIM];
    s.avg[l] = acc / wsum;
  }
  __syncthreads();
  {
    const Dmat<KIND> D(young[ec], pois[ec]);
    for (int q = l; q < NG * NN; q += LPE) {
      const int g = q / NN, i = q % NN;
      const double* d = s.dndx[g][i];
      const double wg = s.w[g];
      for (int k = 0; k < DIM; ++k) {
        const int c = DIM * i + k;
        double b[NS];
        for (int r = 0; r < NS; ++r) b[r] = bentry<KIND>(r, k, d);
        if (E::BBAR) {
          const double corr = (s.avg[c] - d[k]) / static_cast<double>(E::BBAR);
          for (int r = 0; r < E::BBAR; ++r) b[r] = b[r] + corr;
        }
        for (int r = 0; r < NS; ++r) {
          s.bw[g][r][c] = b[r] * wg;
          double acc = 0.0;
          for (int m = 0; m < NS; ++m)
            if (D.v[r][m] != 0.0) acc = acc + D.v[r][m] * b[m];
          s.db[g][r][c] = acc;
        }
      }
    }
  }
  __syncthreads();
  // lane (a, b): K_e[DIM a + ki][DIM b + kj] = sum over points, rows of
  // (B[r][.] w) (D B)[r][.] (src/elements/common.rs:289-296)
  const int a = l / NN, bn = l % NN;
  double k[DIM][DIM];
  for (int ki = 0; ki < DIM; ++ki)
    for (int kj = 0; kj < DIM; ++kj) k[ki][kj] = 0.0;
  for (int g = 0; g < NG; ++g) {
#pragma unroll
    for (int r = 0; r < NS; ++r) {
      double bv[DIM], dv[DIM];
#pragma unroll
      for (int q = 0; q < DIM; ++q) {
        bv[q] = nz_b<KIND>(r, q) ? s.bw[g][r][DIM * a + q] : 0.0;
        dv[q] = nz_db<KIND>(r, q) ? s.db[g][r][DIM * bn + q] : 0.0;
      }
#pragma unroll
      for (int ki = 0; ki < DIM; ++ki)
#pragma unroll
        for (int kj = 0; kj < DIM; ++kj)
          if (nz_b<KIND>(r, ki) && nz_db<KIND>(r, kj)) k[ki][kj] = k[ki][kj] + bv[ki] * dv[kj];
    }
  }
  // the transposed block from lane (b, a)
  const int partner = slot * LPE + bn * NN + a;
  double kt[DIM][DIM];
#pragma unroll
  for (int ki = 0; ki < DIM; ++ki)
#pragma unroll
    for (int kj = 0; kj < DIM; ++kj) kt[kj][ki] = __shfl(k[ki][kj], partner, 64);
  if (live && a <= bn) {
    const int pair = a * NN - a * (a - 1) / 2 + (bn - a);
    double* o = ke + (ec * NPAIR + pair) * (DIM * DIM);
#pragma unroll
    for (int ki = 0; ki < DIM; ++ki)
#pragma unroll
      for (int kj = 0; kj < DIM; ++kj) o[ki * DIM + kj] = 0.5 * (k[ki][kj] + kt[ki][kj]);
  }
}

// out[p] = sum of the element-matrix entries of entry p, in list order, from zero
__global__ __launch_bounds__(256) void k_gather_sum(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                                                    const double* __restrict__ ke, int64_t nout,
                                                    double* __restrict__ out) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (p >= nout) return;
  double s = 0.0;
  const uint32_t k1 = off[p + 1];
  for (uint32_t k = off[p]; k < k1; ++k) s = s + ke[src[k]];
  out[p] = s;
}

template <typename X>
X* dup(const std::vector<X>& v) {
  X* p = nullptr;
  AHIPCHK(hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(X)));
  if (!v.empty()) AHIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice));
  return p;
}

struct Contribution {
  int32_t pos;   // row position within the column
  uint32_t e;    // element
  int32_t key;   // order of the reference's (a, b) loops within the element
  uint32_t src;  // index into the stored element matrices
};

class Assembly final : public GpuAssembly {
 public:
  Assembly(int64_t nelem, ElemType etype, const int64_t* conn, int64_t nn, const double* coords, const int64_t* eq,
           int64_t n)
      : nelem_(nelem), kind_(static_cast<int>(etype) - 1), n_(n) {
    int nen = 0, dim = 0;
    elem_shape(static_cast<int32_t>(etype), &nen, &dim);
    if (nen == 0) throw std::runtime_error("unknown element type");
    if (nelem <= 0 || nn <= 0 || n <= 0) throw std::runtime_error("empty mesh");
    nen_ = nen;
    const auto t0 = std::chrono::steady_clock::now();
    const int nde = dim * nen, npair = nen * (nen + 1) / 2;
    per_elem_ = npair * dim * dim;
    if (static_cast<double>(nelem) * per_elem_ > 4294967295.0 ||
        static_cast<double>(nelem) * nde * nde > 4294967295.0)
      throw std::runtime_error("mesh too large for 32-bit element-matrix indices");
    for (int64_t i = 0; i < nelem * nen; ++i)
      if (conn[i] < 0 || conn[i] >= nn) throw std::runtime_error("connectivity refers to a node out of range");
    for (int64_t i = 0; i < nn * dim; ++i)
      if (eq[i] < -1 || eq[i] >= n) throw std::runtime_error("equation number out of range");
    // elements of each node, increasing, each once
    auto first_at = [&](int64_t e, int a) {
      for (int b = 0; b < a; ++b)
        if (conn[e * nen + b] == conn[e * nen + a]) return false;
      return true;
    };
    std::vector<int64_t> eptr(nn + 1, 0);
    for (int64_t e = 0; e < nelem; ++e)
      for (int a = 0; a < nen; ++a)
        if (first_at(e, a)) ++eptr[conn[e * nen + a] + 1];
    for (int64_t v = 0; v < nn; ++v) eptr[v + 1] += eptr[v];
    std::vector<uint32_t> elist(eptr[nn]);
    {
      std::vector<int64_t> pos(eptr.begin(), eptr.end() - 1);
      for (int64_t e = 0; e < nelem; ++e)
        for (int a = 0; a < nen; ++a)
          if (first_at(e, a)) elist[pos[conn[e * nen + a]]++] = static_cast<uint32_t>(e);
    }
    // the (node, component) of each equation
    std::vector<int64_t> enode(n, -1);
    std::vector<int8_t> ecomp(n, 0);
    for (int64_t v = 0; v < nn; ++v)
      for (int k = 0; k < dim; ++k)
        if (eq[v * dim + k] >= 0) {
          if (enode[eq[v * dim + k]] >= 0) throw std::runtime_error("two DOFs share an equation");
          enode[eq[v * dim + k]] = v;
          ecomp[eq[v * dim + k]] = static_cast<int8_t>(k);
        }
    for (int64_t q = 0; q < n; ++q)
      if (enode[q] < 0) throw std::runtime_error("an equation has no DOF");
    // the stored index of local entry (i, j) of an element matrix
    std::vector<int32_t> sidx(nde * nde);
    for (int i = 0; i < nde; ++i)
      for (int j = 0; j < nde; ++j) {
        int a = i / dim, ki = i % dim, b = j / dim, kj = j % dim;
        if (a > b) {
          std::swap(a, b);
          std::swap(ki, kj);
        }
        const int pair = a * nen - a * (a - 1) / 2 + (b - a);
        sidx[i * nde + j] = pair * dim * dim + ki * dim + kj;
      }
    // Per column c, its rows and the contributions of each entry.  Entry
    // (row, c) takes, from every element e holding both DOFs in increasing e,
    // the symmetric parts of the local pairs (i, j) mapping to it, ordered as
    // the reference's loops over the row-local index of the larger equation
    // then the other (src/fem/assembly.rs:378-398), so the lower and the upper
    // entry of a pair sum the same terms in the same order.
    auto column = [&](int64_t c, std::vector<int64_t>& cand, std::vector<Contribution>& con) {
      const int64_t p = enode[c];
      const int k = ecomp[c];
      cand.clear();
      con.clear();
      for (int64_t q = eptr[p]; q < eptr[p + 1]; ++q) {
        const int64_t e = elist[q];
        for (int a = 0; a < nen; ++a)
          for (int r = 0; r < dim; ++r) {
            const int64_t row = eq[conn[e * nen + a] * dim + r];
            if (row >= 0) cand.push_back(row);
          }
      }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      for (int64_t q = eptr[p]; q < eptr[p + 1]; ++q) {
        const int64_t e = elist[q];
        for (int a = 0; a < nen; ++a)
          for (int r = 0; r < dim; ++r) {
            const int64_t row = eq[conn[e * nen + a] * dim + r];
            if (row < 0) continue;
            const int32_t pos =
                static_cast<int32_t>(std::lower_bound(cand.begin(), cand.end(), row) - cand.begin());
            const int i = dim * a + r;
            for (int b = 0; b < nen; ++b) {
              if (conn[e * nen + b] != p) continue;
              const int j = dim * b + k;
              const int key = row >= c ? i * nde + j : j * nde + i;
              con.push_back({pos, static_cast<uint32_t>(e), key,
                             static_cast<uint32_t>(e * per_elem_ + sidx[i * nde + j])});
            }
          }
      }
      std::sort(con.begin(), con.end(), [](const Contribution& x, const Contribution& y) {
        if (x.pos != y.pos) return x.pos < y.pos;
        if (x.e != y.e) return x.e < y.e;
        return x.key < y.key;
      });
    };
    const int nt = static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    auto run = [&](auto&& body) {
      std::vector<std::thread> th;
      for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
          std::vector<int64_t> cand;
          std::vector<Contribution> con;
          for (int64_t c = t; c < n; c += nt) {
            column(c, cand, con);
            body(c, cand, con);
          }
        });
      for (auto& x : th) x.join();
    };
    std::vector<int64_t> nrow(n, 0), ncon(n, 0);
    run([&](int64_t c, const std::vector<int64_t>& cand, const std::vector<Contribution>& con) {
      nrow[c] = static_cast<int64_t>(cand.size());
      ncon[c] = static_cast<int64_t>(con.size());
    });
    indptr_.assign(n + 1, 0);
    std::vector<int64_t> cptr(n + 1, 0);
    for (int64_t c = 0; c < n; ++c) {
      indptr_[c + 1] = indptr_[c] + nrow[c];
      cptr[c + 1] = cptr[c] + ncon[c];
    }
    ncon_ = cptr[n];
    indices_.resize(indptr_[n]);
    off_h_.assign(indptr_[n] + 1, 0);
    src_h_.resize(ncon_);
    run([&](int64_t c, const std::vector<int64_t>& cand, const std::vector<Contribution>& con) {
      std::copy(cand.begin(), cand.end(), indices_.begin() + indptr_[c]);
      int64_t at = cptr[c];
      size_t i = 0;
      for (size_t pos = 0; pos < cand.size(); ++pos) {
        while (i < con.size() && con[i].pos == static_cast<int32_t>(pos)) src_h_[at++] = con[i++].src;
        off_h_[indptr_[c] + pos + 1] = static_cast<uint32_t>(at);
      }
    });
    std::vector<int32_t> conn32(conn, conn + nelem * nen);
    std::vector<double> xyz(coords, coords + nn * 3);
    d_conn_ = dup(conn32);
    d_coords_ = dup(xyz);
    d_off_ = dup(off_h_);
    d_src_ = dup(src_h_);
    AHIPCHK(hipMalloc(&d_ke_, static_cast<size_t>(nelem) * per_elem_ * sizeof(double)));
    AHIPCHK(hipMalloc(&d_mat_, static_cast<size_t>(nelem) * 3 * sizeof(double)));
    AHIPCHK(hipMalloc(&d_data_, std::max<int64_t>(indptr_[n], 1) * sizeof(double)));
    AHIPCHK(hipMalloc(&d_bad_, sizeof(unsigned long long)));
    for (auto& ev : ev_) AHIPCHK(hipEventCreate(&ev));
    build_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  ~Assembly() override {
    for (void* p : {static_cast<void*>(d_conn_), static_cast<void*>(d_coords_), static_cast<void*>(d_off_),
                    static_cast<void*>(d_src_), static_cast<void*>(d_ke_), static_cast<void*>(d_mat_),
                    static_cast<void*>(d_data_), static_cast<void*>(d_bad_), static_cast<void*>(d_boff_),
                    static_cast<void*>(d_bsrc_), static_cast<void*>(d_bout_)})
      if (p) (void)hipFree(p);
    for (hipEvent_t e : ev_)
      if (e) (void)hipEventDestroy(e);
  }

  int64_t n() const override { return n_; }
  int64_t nnz() const override { return indptr_.empty() ? 0 : indptr_.back(); }
  const std::vector<int64_t>& indptr() const override { return indptr_; }
  const std::vector<int64_t>& indices() const override { return indices_; }

  int64_t values(const double* young, const double* nu, const double* thickness, double* data,
                 bool on_device) override {
    const int64_t bad = element_matrices(young, nu, thickness);
    if (bad >= 0) return bad;
    double* out = on_device ? data : d_data_;
    gather_ms_ = gather(d_off_, d_src_, nnz(), out);
    if (!on_device) AHIPCHK(hipMemcpy(data, d_data_, nnz() * sizeof(double), hipMemcpyDeviceToHost));
    return -1;
  }

  void bind(const std::vector<int64_t>& vmap, int64_t nout, uint64_t owner) override {
    if (static_cast<int64_t>(vmap.size()) != nnz()) throw std::runtime_error("bind: one position per stored entry");
    if (nout < 0) throw std::runtime_error("bind: negative size");
    std::vector<int64_t> from(nout, -1);
    for (size_t p = 0; p < vmap.size(); ++p) {
      if (vmap[p] < 0) continue;
      if (vmap[p] >= nout) throw std::runtime_error("bind: position out of range");
      if (from[vmap[p]] >= 0) throw std::runtime_error("bind: two entries go to one position");
      from[vmap[p]] = static_cast<int64_t>(p);
    }
    std::vector<uint32_t> boff(nout + 1, 0);
    std::vector<uint32_t> bsrc;
    bsrc.reserve(ncon_ / 2 + nout);
    for (int64_t q = 0; q < nout; ++q) {
      if (from[q] >= 0)
        bsrc.insert(bsrc.end(), src_h_.begin() + off_h_[from[q]], src_h_.begin() + off_h_[from[q] + 1]);
      boff[q + 1] = static_cast<uint32_t>(bsrc.size());
    }
    for (void* p : {static_cast<void*>(d_boff_), static_cast<void*>(d_bsrc_), static_cast<void*>(d_bout_)})
      if (p) (void)hipFree(p);
    d_boff_ = nullptr;
    d_bsrc_ = nullptr;
    d_bout_ = nullptr;
    owner_ = 0;
    d_boff_ = dup(boff);
    d_bsrc_ = dup(bsrc);
    AHIPCHK(hipMalloc(&d_bout_, std::max<int64_t>(nout, 1) * sizeof(double)));
    nbound_ = nout;
    owner_ = owner;
  }

  uint64_t bound_owner() const override { return owner_; }
  const double* bound_values() const override { return d_bout_; }

  int64_t values_bound(const double* young, const double* nu, const double* thickness) override {
    if (!d_boff_) throw std::runtime_error("no bound layout");
    const int64_t bad = element_matrices(young, nu, thickness);
    if (bad >= 0) return bad;
    bound_ms_ = gather(d_boff_, d_bsrc_, nbound_, d_bout_);
    return -1;
  }

  void stats(double* out, int len) const override {
    const double v[6] = {ke_ms_, gather_ms_, build_s_, static_cast<double>(ncon_), static_cast<double>(per_elem_),
                         bound_ms_};
    for (int i = 0; i < len && i < 6; ++i) out[i] = v[i];
  }

 private:
  // element matrices into d_ke_; returns -1 or the lowest distorted element
  int64_t element_matrices(const double* young, const double* nu, const double* thickness) {
    const size_t ne = static_cast<size_t>(nelem_);
    std::vector<double> mat(3 * ne);
    for (size_t e = 0; e < ne; ++e) {
      mat[e] = young[e];
      mat[ne + e] = nu[e];
      mat[2 * ne + e] = thickness ? thickness[e] : 1.0;
    }
    AHIPCHK(hipMemcpy(d_mat_, mat.data(), mat.size() * sizeof(double), hipMemcpyHostToDevice));
    const unsigned long long none = ~0ull;
    AHIPCHK(hipMemcpy(d_bad_, &none, sizeof none, hipMemcpyHostToDevice));
    const int epw = 64 / (nen_ * nen_);
    const unsigned blocks = static_cast<unsigned>((nelem_ + epw - 1) / epw);
    const double *dy = d_mat_, *dn = d_mat_ + ne, *dt = d_mat_ + 2 * ne;
    AHIPCHK(hipEventRecord(ev_[0], nullptr));
    switch (kind_) {
      case 0:
        hipLaunchKernelGGL(k_elem<0>, dim3(blocks), dim3(64), 0, nullptr, d_conn_, d_coords_, dy, dn, dt, nelem_,
                           d_ke_, d_bad_);
        break;
      case 1:
        hipLaunchKernelGGL(k_elem<1>, dim3(blocks), dim3(64), 0, nullptr, d_conn_, d_coords_, dy, dn, dt, nelem_,
                           d_ke_, d_bad_);
        break;
      case 2:
        hipLaunchKernelGGL(k_elem<2>, dim3(blocks), dim3(64), 0, nullptr, d_conn_, d_coords_, dy, dn, dt, nelem_,
                           d_ke_, d_bad_);
        break;
      default:
        hipLaunchKernelGGL(k_elem<3>, dim3(blocks), dim3(64), 0, nullptr, d_conn_, d_coords_, dy, dn, dt, nelem_,
                           d_ke_, d_bad_);
        break;
    }
    AHIPCHK(hipGetLastError());
    AHIPCHK(hipEventRecord(ev_[1], nullptr));
    AHIPCHK(hipEventSynchronize(ev_[1]));
    float ms = 0;
    AHIPCHK(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
    ke_ms_ = ms;
    unsigned long long bad = 0;
    AHIPCHK(hipMemcpy(&bad, d_bad_, sizeof bad, hipMemcpyDeviceToHost));
    return bad == ~0ull ? -1 : static_cast<int64_t>(bad);
  }

  // one thread per output entry; returns the kernel's ms
  double gather(const uint32_t* off, const uint32_t* src, int64_t nout, double* out) {
    if (nout == 0) return 0.0;
    AHIPCHK(hipEventRecord(ev_[2], nullptr));
    hipLaunchKernelGGL(k_gather_sum, dim3(static_cast<unsigned>((nout + 255) / 256)), dim3(256), 0, nullptr, off, src,
                       d_ke_, nout, out);
    AHIPCHK(hipGetLastError());
    AHIPCHK(hipEventRecord(ev_[3], nullptr));
    AHIPCHK(hipEventSynchronize(ev_[3]));
    float ms = 0;
    AHIPCHK(hipEventElapsedTime(&ms, ev_[2], ev_[3]));
    return ms;
  }

  int64_t nelem_;
  int kind_;
  int64_t n_;
  int nen_ = 0, per_elem_ = 0;
  int64_t ncon_ = 0, nbound_ = 0;
  std::vector<int64_t> indptr_, indices_;
  std::vector<uint32_t> off_h_, src_h_;
  int32_t* d_conn_ = nullptr;
  double* d_coords_ = nullptr;
  uint32_t* d_off_ = nullptr;
  uint32_t* d_src_ = nullptr;
  uint32_t* d_boff_ = nullptr;
  uint32_t* d_bsrc_ = nullptr;
  double* d_bout_ = nullptr;
  uint64_t owner_ = 0;
  double* d_ke_ = nullptr;
  double* d_mat_ = nullptr;
  double* d_data_ = nullptr;
  unsigned long long* d_bad_ = nullptr;
  hipEvent_t ev_[4] = {nullptr, nullptr, nullptr, nullptr};
  double ke_ms_ = 0, gather_ms_ = 0, build_s_ = 0, bound_ms_ = 0;
};

}  // namespace

void elem_shape(int32_t etype, int* nodes, int* dofs) {
  switch (etype) {
    case 1: *nodes = 8; *dofs = 3; return;
    case 2: *nodes = 4; *dofs = 3; return;
    case 3:
    case 4: *nodes = 4; *dofs = 2; return;
    default: *nodes = 0; *dofs = 0;
  }
}

std::unique_ptr<GpuAssembly> make_gpu_assembly(int64_t nelem, ElemType etype, const int64_t* conn, int64_t nn,
                                               const double* coords, const int64_t* eq, int64_t n) {
  return std::make_unique<Assembly>(nelem, etype, conn, nn, coords, eq, n);
}

}  // namespace rf
