// Indefinite LDL^T (1x1/2x2 threshold pivots, delayed pivots, Bunch-Kaufman
// at the roots) and unsymmetric-value LU (diagonal threshold pivots, delayed
// pivots, partial pivoting at the roots) on the GPU, on the same
// nested-dissection tree as the positive definite solver.
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

#include "symbolic.hpp"

namespace rf {

class IndefGpu {
 public:
  virtual ~IndefGpu() = default;
  virtual void set_caller_map(const std::vector<int64_t>& eq, const std::vector<double>& iso_diag) = 0;
  // lower values of the pattern (LU: lower values followed by the transposed
  // strictly-upper values on the same pattern).  Returns -1 or the singular equation.
  virtual int64_t factor(const double* vals) = 0;
  virtual void solve(double* b, int nrhs) = 0;  // caller numbering, host pointer
  virtual void inertia(int64_t out[3]) const = 0;
  virtual int64_t factor_size() const = 0;
  // neg, zero, pos, delayed, two_by_two, factor_size (LU: 0, 0, 0, delayed, 0, size)
  virtual void stats(int64_t out[6]) const = 0;
  // pivot sequence in front postorder: internal equation, d, e, kind (LDL^T:
  // 1 = 1x1, 2 = first of a 2x2 with e its off-diagonal, 0 = second; LU: d =
  // U diagonal, kind 1, or 3 with e the row permutation entry for the
  // partial-pivoting block of a root).  Returns the count.
  virtual int64_t pivots(int64_t* eq, double* d, double* e, int32_t* kind) const = 0;
};

std::unique_ptr<IndefGpu> make_indef_gpu(Symbolic sym, const std::vector<int64_t>& colptr,
                                         const std::vector<int64_t>& rowidx, bool lu);

}  // namespace rf
