// Smoothed-aggregation AMG setup (host, C++ threads) and the GPU
// preconditioned Krylov solver built on it.
//
// The setup restates src/solver/iterative/amg.rs:37-341 and the sparse
// kernels of src/solver/iterative/csr.rs:7-190: rigid body modes, Vanek
// three-pass aggregation with strength theta (halved per level), QR of the
// near-null space per aggregate (tentative prolongator), one damped Jacobi
// smoothing step with omega = 4 / (3 rho(D^-1 A)), Galerkin products, and a
// dense LDL^T at the coarsest level.  The spectral radius is computed on the
// GPU by the same 15-step power method.
#pragma once
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace rf {

// The hierarchy cannot be built: a level diagonal is not positive or the
// coarse LDL^T breaks down (the Err(()) of src/solver/iterative/amg.rs:282-341).
struct AmgBreakdown : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// scalar CSR with sorted column indices
struct Csr {
  int64_t nrows = 0, ncols = 0;
  std::vector<int64_t> ptr{0};
  std::vector<int32_t> idx;
  std::vector<double> val;
  int64_t nnz() const { return ptr.empty() ? 0 : ptr[nrows]; }
};

// full symmetric CSR from the lower triangle in CSC (src/solver/iterative/csr.rs:21-60)
Csr csr_from_lower(int64_t n, const int64_t* colptr, const int64_t* rowidx, const double* val);
Csr csr_transpose(const Csr& a);                 // src/solver/iterative/csr.rs:111-134
Csr csr_mul(const Csr& a, const Csr& b);         // Gustavson, src/solver/iterative/csr.rs:136-189
std::vector<double> csr_diag(const Csr& a);      // src/solver/iterative/csr.rs:96-108

// rigid body modes, row-major n x nb (src/solver/iterative/amg.rs:39-95)
std::vector<double> rigid_body_modes(int64_t nn, const double* coords, const int64_t* ndof, int& nb);

struct AmgHostLevel {
  Csr A, P, R;
  std::vector<double> diag;
  double rho = 0.0;     // 1.05 x the power-method estimate of rho(D^-1 A)
  double omega = 0.0;
  int64_t nblk = 0;     // node blocks of this level
  std::vector<int64_t> bptr;
  // kept only when amg_setup is asked to keep the hierarchy (parity checks):
  std::vector<int64_t> agg;  // aggregate of every block
  Csr Pt;                    // tentative prolongator
  std::vector<double> B;     // near-null space of this level (rows x nbm, row-major)
  int nbm = 0;
};

struct AmgHost {
  std::vector<AmgHostLevel> levels;   // smoothed levels (finest first)
  Csr coarse;                         // the coarsest matrix
  std::vector<int64_t> coarse_bptr;   // its DOF blocks (one per aggregate of the last level)
  std::vector<double> coarse_l, coarse_d;  // its dense LDL^T (unit L row-major, D)
  std::vector<int64_t> sizes;         // rows per level, finest first, coarsest last
  double complexity = 0.0;            // total nnz / fine nnz
  std::vector<double> coarse_B;       // near-null space of the coarsest level (kept hierarchy only)
};

// Builds the hierarchy (src/solver/iterative/amg.rs:282-341).  spectral_radius
// returns 1.05 x the 15-step power-method estimate of rho(D^-1 A) for a level.
// Throws when the coarse LDL^T breaks down (d <= 1e-13 dmax) or a diagonal is
// not positive.
// keep = true also records the aggregates, the tentative prolongators and the
// near-null spaces of every level (AmgHostLevel::agg, Pt, B; coarse_B).
AmgHost amg_setup(Csr a, std::vector<int64_t> bptr, std::vector<double> b, int nb,
                  const std::function<double(const Csr&, const std::vector<double>&)>& spectral_radius,
                  bool keep = false);

// A_c^-1 = L^-T D^-1 L^-1 as a dense row-major matrix, from the coarse LDL^T
std::vector<double> coarse_inverse(const std::vector<double>& l, const std::vector<double>& d, int64_t n);

// aggregation of one level (src/solver/iterative/amg.rs:108-181); exposed for tests
std::vector<int64_t> amg_aggregate(const Csr& a, const std::vector<int64_t>& bptr, double theta, int64_t& nagg);

// the three passes of the aggregation (src/solver/iterative/amg.rs:140-181)
// over the strong neighbours of every block (CSR sptr/sidx, each list in the
// order of first appearance in the block's rows)
std::vector<int64_t> aggregate_passes(int64_t nb, const std::vector<int64_t>& sptr, const std::vector<int32_t>& sidx,
                                      int64_t& nagg);

// dense LDL^T of the coarsest matrix (src/solver/iterative/amg.rs:319-340):
// unit L row-major and D; throws AmgBreakdown when a pivot d <= 1e-13 max|a_ii|
void coarse_ldlt(const Csr& a, std::vector<double>& l, std::vector<double>& d);

}  // namespace rf
