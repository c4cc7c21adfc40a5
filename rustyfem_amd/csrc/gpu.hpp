// Device side of the multifrontal solver: owns the HBM buffers of one
// factorization and runs the level-batched HIP kernels on its own stream.
#pragma once
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "mf.hpp"

namespace rf {

enum class Precision : int32_t { FP64 = 0, MIXED = 1 };

struct Timings {
  double factor_ms = 0, solve_ms = 0, refine_ms = 0;
  int32_t refine_steps = 0;
  double rel_residual = -1;
  // per-launch HIP-event timing of the Schur-update GEMM kernel (when enabled)
  double gemm_ms = 0, gemm_launches = 0, gemm_flops = 0;
  // GEMM autotuning at the first factorization: its time, the GEMM launches
  // of the plan and how many of them run the AGPR-accumulator variant
  double tune_ms = 0, tuned_launches = 0, agpr_launches = 0, ilv_launches = 0, q_launches = 0;
  // worst relative residual after 0, 1, 2, 3 refinement corrections (-1: not reached)
  double refine_history[4] = {-1, -1, -1, -1};
  // 1 when the last refinement took back a correction that raised the worst residual
  int32_t refine_undone = 0;
  // 1 when the last solve missed the refinement target and was repeated by the
  // fallback (an FP64 factorization); the relative residual it reached
  int32_t fell_back = 0;
  double fallback_rel = -1;
  // tree levels whose solves run through the fused small-front kernels
  double small_levels = 0;
};

// Solves in internal numbering on device vectors (b, x: n x nrhs), returns the
// relative residual reached, or a negative value when it cannot solve.
using SolveFallback = std::function<double(const double* d_b, double* d_x, int nrhs)>;

class GpuFactor {
 public:
  virtual ~GpuFactor() = default;
  // Numeric factorization from Lower values (internal numbering, host or device
  // pointer).  Returns -1, or the first equation whose pivot fails the test.
  virtual int64_t factor(const double* lower_vals, bool on_device) = 0;
  // Caller numbering: eq[i] is the internal equation of caller DOF i (-1 for an
  // uncoupled DOF, solved as b / diag with its entry in iso_diag).
  virtual void set_caller_map(const std::vector<int64_t>& eq, const std::vector<double>& iso_diag) = 0;
  // In-place solve of nrhs columns in caller numbering (column-major
  // n_caller x nrhs, host or device pointer).  With max_steps > 0 the solve
  // is followed by FP64 iterative refinement against the FP64 matrix until
  // ||b - A x|| <= tol ||b|| for every column, max_steps corrections, or a
  // correction that does not halve the worst residual (FP64 floor reached).
  virtual void solve(double* b, int nrhs, bool on_device, double tol, int max_steps) = 0;
  // solve() on a device pointer without allocating device memory: the work
  // vectors come from a fixed per-device scratch array (throws when they do
  // not fit); calls that overlap on one device race on that array.
  // Returns false when the refinement missed its target (no fallback runs here).
  virtual bool solve_noalloc(double* d_b, int nrhs, double tol, int max_steps) = 0;
  // Solve in internal numbering: d_b, d_x device vectors (n x nrhs); d_b is
  // not changed.  Returns the worst relative residual after refinement (-1
  // without refinement).
  virtual double solve_internal(const double* d_b, double* d_x, int nrhs, double tol, int max_steps) = 0;
  // Called by solve() when the refinement misses tol: it solves the same
  // internal right-hand sides again (for the mixed factor: with an FP64 one).
  virtual void set_fallback(SolveFallback f) = 0;
  // the Lower values of the last factorization, in device memory (internal numbering)
  virtual const double* device_lower_values() const = 0;
  virtual void pivots(double* d_host) = 0;
  virtual const Timings& timings() const = 0;
  virtual double flops() const = 0;
  virtual size_t device_bytes() const = 0;
  virtual void sync() = 0;
  // bracket every Schur-update GEMM launch with HIP events during factor()
  virtual void set_kernel_timing(bool on) = 0;
  // Independent problems sharing the factor (a batch): internal equations
  // [bounds[i], bounds[i+1]) belong to problem i, and refinement tests the
  // residual of every problem and column separately.
  virtual void set_segments(const std::vector<int64_t>& bounds) = 0;
};

// lower: Lower pattern (internal numbering), used to build the refinement matrix
std::unique_ptr<GpuFactor> make_gpu_factor(Plan plan, const std::vector<int64_t>& colptr,
                                           const std::vector<int64_t>& rowidx, Precision prec);

// last HIP error message of this thread (empty if none)
std::string& last_error();

// synchronous copies between host and device memory (throw on a HIP error)
void copy_to_host(void* dst, const void* src, size_t bytes);
void copy_to_device(void* dst, const void* src, size_t bytes);

}  // namespace rf
