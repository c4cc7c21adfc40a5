// Incomplete Cholesky without fill, A ~ L D L^T on the pattern of A, on the
// GPU (src/solver/iterative/csr.rs:219-324): factorization and application
// are level-scheduled over the dependency graph of L.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "amg.hpp"

namespace rf {

class Ic0Gpu {
 public:
  virtual ~Ic0Gpu() = default;
  // z = (L D L^T)^-1 r for device vectors r, z (n each, z != r); every kernel
  // returns at once when done != nullptr and *done != 0
  virtual void apply(const double* r, double* z, const int32_t* done) = 0;
  // the diagonal shift of the attempt that succeeded (0, 1e-3, 4e-3, ...)
  virtual double shift() const = 0;
  virtual int64_t forward_levels() const = 0;
  virtual int64_t backward_levels() const = 0;
  virtual double factor_ms() const = 0;
  virtual int attempts() const = 0;
};

// Factors the full symmetric CSR matrix `a` (rows sorted) on stream s; every
// row of L may hold at most IC0_MAX_ROW entries.  Throws after 40 failed
// shifts.
constexpr int IC0_MAX_ROW = 1024;
std::unique_ptr<Ic0Gpu> make_ic0_gpu(const Csr& a, hipStream_t s);

}  // namespace rf
