// Shift-invert block subspace iteration for K x = lambda B x on one MI355X
// (src/solver/eigen.rs:229-375): the block vectors, the products with B and
// the Gram and combination products live in device memory; the small
// projected problems are solved on the host.
#pragma once
#include <cstdint>
#include <functional>
#include <vector>

namespace rf {

struct EigenOptions {
  int nev = 1;             // requested pairs
  int vectors = 0;         // iteration vectors (0 = min(2 nev, nev + 8))
  int max_iter = 2000;
  double tol = 1e-10;      // relative residual of the Ritz vectors (Op norm)
  bool allow_unconverged = false;
};

struct EigenResult {
  std::vector<double> theta;    // largest |theta| first: Op x = (1/theta) B x
  std::vector<double> vectors;  // n x (theta.size()), column-major, x^T Op x = 1
  int iterations = 0;
  bool converged = false;
};

// Op^-1 applied in place to q device vectors (n x q, column-major).
using BlockSolve = std::function<void(double* d_y, int q)>;

// B is symmetric, given by its rows (CSR, both triangles; for a symmetric
// matrix the caller's CSC arrays are the same).  Throws on failure.
EigenResult subspace_gpu(int64_t n, const int64_t* bptr, const int64_t* bidx, const double* bval,
                         const BlockSolve& solve, const EigenOptions& opt);

// Cyclic Jacobi for a dense symmetric m x m matrix (row-major): eigenvalues
// (diagonal after the sweeps) and eigenvectors (v[i * m + j] = component i of
// vector j).  The projected problems of subspace_gpu use it.
void jacobi_eigen(std::vector<double> a, int m, std::vector<double>& d, std::vector<double>& v);

}  // namespace rf
