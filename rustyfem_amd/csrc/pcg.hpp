// GPU preconditioned Krylov solver for the lower-triangle matrices of the
// iterative path (src/solver/iterative.rs:275-292, csr.rs:326-380): AMG
// V-cycle (Chebyshev or damped-Jacobi smoothing, FP64 or FP32), IC(0),
// Jacobi; standard PCG or flexible CG.
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

namespace rf {

struct AmgHost;

struct PcgOptions {
  int krylov = 0;        // 0: PCG (csr.rs:341-380); 1: flexible CG (FCG(1), Notay)
  int vcycle_fp32 = 0;   // V-cycle arithmetic: 0 FP64, 1 FP32 (matrices and vectors)
  int smoother = 0;      // 0: Chebyshev of degree `degree`; 1: damped Jacobi (omega = 4 / (3 rho))
  int degree = 2;        // Chebyshev degree, pre and post
  double cheb_ratio = 20.0;  // Chebyshev interval [rho / ratio, rho]
  int batch = 8;         // iterations between convergence checks on the host
  bool keep_hierarchy = false;       // keep the host AMG hierarchy (PcgGpu::hierarchy)
  bool force_amg_breakdown = false;  // test hook: the AMG setup reports a breakdown
};

struct PcgResult {
  int64_t iterations = 0;
  double rel_residual = 0.0;
  bool converged = false;
};

class PcgGpu {
 public:
  virtual ~PcgGpu() = default;
  // x = A^-1 b from x0 = 0; b and x in host or device memory
  virtual PcgResult solve(const double* b, double* x, double tol, int64_t maxit, bool on_device) = 0;
  // [0] setup ms (host + GPU), [1] last solve ms, [2] levels (incl. coarsest),
  // [3] operator complexity, [4] fine-matrix SpMV ms (one product, measured),
  // [5] fine-matrix bytes per SpMV, [6] coarsest size, [7] iterations of the last solve,
  // [8..19] rows per level (finest first), [20] IC(0) diagonal shift,
  // [21] IC(0) forward levels, [22] IC(0) backward levels, [23] IC(0) factor attempts,
  // [24] 1 when the last solve fell back from the learned smoother to plain AMG-PCG,
  // [25] true relative residual of the last solve (learned smoother set; else -1),
  // [26] iterations of the abandoned learned-smoother attempt, [27] learned sweeps,
  // [28..39] DOFs per block of the level matrices (finest first; 1 = scalar CSR)
  virtual void stats(double* out, size_t len) const = 0;
  // z = M r, one application of the preconditioner (Precond::apply,
  // src/solver/iterative/csr.rs:193-196); r and z in host or device memory
  virtual void apply(const double* r, double* z, bool on_device) = 0;
  virtual std::vector<int64_t> level_sizes() const = 0;
  // AMG only, 3 DOFs per node: the finest level's smoother becomes the learned
  // one of mpnn.hpp (x = M b, then x += M (b - A x), `sweeps` steps before and
  // after the coarse correction), M computed on the GPU from the finest
  // matrix and the network parameters (MPNN_NPARAMS doubles, host).  Every
  // solve then checks the true FP64 residual of its result; a breakdown, a
  // residual that has not halved over 40 iterations, or a true residual above
  // the tolerance makes the solve start again with plain AMG-PCG (Chebyshev
  // smoothing, standard PCG), whose result is returned.
  virtual void set_learned_smoother(const double* params, int sweeps) = 0;
  // the host AMG hierarchy (aggregates, tentative and smoothed prolongators,
  // level matrices), when the options asked to keep it; else nullptr
  virtual const AmgHost* hierarchy() const { return nullptr; }
  // true when the AMG hierarchy was built on the GPU (else on the host)
  virtual bool setup_on_device() const { return false; }
};

// kind: 0 AMG (smoothed aggregation, rigid body modes), 1 IC(0), 2 Jacobi.  coords /
// ndof describe the node blocks (natural numbering: node v owns ndof[v]
// consecutive equations).  Throws AmgBreakdown (amg.hpp) when the AMG setup
// breaks down (src/solver/iterative/amg.rs:289-291, 328-330: Err(())).
std::unique_ptr<PcgGpu> make_pcg_gpu(int64_t n, const int64_t* colptr, const int64_t* rowidx, const double* val,
                                     int kind, int64_t nn, const double* coords, const int64_t* ndof,
                                     const PcgOptions& opt);

}  // namespace rf
