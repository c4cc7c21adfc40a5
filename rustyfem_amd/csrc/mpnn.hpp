// Learned smoother for 3x3-block (3 DOFs per node) stiffness matrices: a small
// message-passing network maps the matrix's local blocks to a symmetric block
// matrix M on the matrix's own block pattern; the smoother is x += M (b - A x).
//
// Network (FP64, H = 32 hidden units, F = 10 edge features), for every block
// e = (i, j) of the pattern (including the diagonal block):
//   s_i   = sqrt(3 / trace(A_ii))                       (node scale)
//   x_e   = [s_i s_j vec(A_ij) (9, row-major), [i == j]] (edge features)
//   m_e   = relu(W1 x_e + b1)                            (edge message)
//   n_i   = mean over the blocks e of row i of m_e       (node state)
//   h_e   = relu(W2 [m_e; n_i; n_j] + b2)
//   o_e   = W3 h_e + b3                                  (a 3x3 block, row-major)
//   G_e   = (O_e + O_rev(e)^T) / 2                       (rev(e) = (j, i))
//   M_ij  = s_i s_j (G_e + omega [i == j] I),  omega = 2/3
// so that the network starts from scaled damped Jacobi when W3, b3 are small.
// Parameters, concatenated: W1 (H x F, row-major), b1 (H), W2 (H x 3H), b2 (H),
// W3 (9 x H), b3 (9).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace rf {

constexpr int MPNN_F = 10;
constexpr int MPNN_H = 32;
constexpr double MPNN_OMEGA = 2.0 / 3.0;
constexpr int64_t MPNN_NPARAMS = MPNN_H * MPNN_F + MPNN_H + MPNN_H * 3 * MPNN_H + MPNN_H + 9 * MPNN_H + 9;

// Device-side network state for one block pattern.  The pattern (block row
// pointers, sorted block column indices) and the 3x3 row-major block values
// are device arrays owned by the caller; the pattern must be structurally
// symmetric and hold every diagonal block.
class MpnnGraph {
 public:
  MpnnGraph(int64_t nb, int64_t nnzb, const int64_t* d_ptr, const int32_t* d_idx, const double* d_val,
            hipStream_t s);
  ~MpnnGraph();
  MpnnGraph(const MpnnGraph&) = delete;
  MpnnGraph& operator=(const MpnnGraph&) = delete;
  // M's blocks (nnzb x 9, the pattern's order) for the parameters d_params
  void forward(const double* d_params, double* d_m);
  // d_grad (MPNN_NPARAMS) = dL/dparams from d_gm = dL/dM (nnzb x 9), using the
  // activations of the last forward
  void backward(const double* d_params, const double* d_gm, double* d_grad);
  int64_t nb() const { return nb_; }
  int64_t nnzb() const { return nnzb_; }
  const int32_t* rows() const { return row_; }
  const int32_t* rev() const { return rev_; }

 private:
  int64_t nb_, nnzb_;
  const int64_t* ptr_;
  const int32_t* idx_;
  hipStream_t s_;
  int32_t *row_ = nullptr, *rev_ = nullptr;
  double *scale_ = nullptr, *feat_ = nullptr, *pre1_ = nullptr, *node_ = nullptr, *pre2_ = nullptr, *out_ = nullptr;
  double *go_ = nullptr, *dz_ = nullptr, *dnode_ = nullptr, *dpre1_ = nullptr, *dpre2_ = nullptr;
};

}  // namespace rf
