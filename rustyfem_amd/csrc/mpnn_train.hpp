// Training handle of the learned smoother (see mpnn.hpp): one problem's matrix
// as 3x3 blocks on the device, the smoothing loss and its gradient.
#pragma once
#include <cstdint>
#include <memory>

namespace rf {

class MpnnTrainer {
 public:
  virtual ~MpnnTrainer() = default;
  // L = mean_v e_nu^T A e_nu / e_0^T A e_0 with e_{k+1} = (I - M A) e_k, for
  // the nvec columns of e0 (n x nvec, column-major, host); grad (host,
  // MPNN_NPARAMS) receives dL/dparams when not null
  virtual double loss(const double* params, const double* e0, int nvec, int nu, double* grad) = 0;
  // M's blocks (nnzb x 9, row-major 3x3) and its block pattern (nb + 1 row
  // pointers, nnzb sorted block columns); any output may be null
  virtual void smoother(const double* params, double* blocks, int64_t* ptr, int32_t* idx) = 0;
  virtual int64_t nnzb() const = 0;
};

// number of network parameters (MPNN_NPARAMS of mpnn.hpp)
int64_t mpnn_nparams();

// lower-triangle CSC in natural node-major numbering, 3 DOFs per node
std::unique_ptr<MpnnTrainer> make_mpnn_trainer(int64_t n, const int64_t* colptr, const int64_t* rowidx,
                                               const double* val);

}  // namespace rf
