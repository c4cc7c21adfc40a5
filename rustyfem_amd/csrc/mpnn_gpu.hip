// HIP kernels (gfx950) of the learned smoother's network: forward (the
// smoother's blocks from the matrix) and backward (parameter gradients from
// dL/dM), and a training handle that evaluates the smoothing loss
//   L = mean_v  e_nu^T A e_nu / e_0^T A e_0,   e_{k+1} = (I - M A) e_k,
// and its gradient by the adjoint recursion
//   g_nu = 2 A e_nu / (nvec e_0^T A e_0),
//   dL/dM += -g_{k+1} (A e_k)^T   (on M's block pattern),
//   g_k = g_{k+1} - A M g_{k+1}   (A and M symmetric).
// All reductions run in a fixed order (no atomics), so results are
// reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "amg.hpp"
#include "mpnn.hpp"
#include "mpnn_train.hpp"

namespace rf {

#define MCHK(expr)                                                                                   \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) {                                                                          \
      char b_[512];                                                                                  \
      snprintf(b_, sizeof b_, "HIP error %s at %s:%d: %s", hipGetErrorString(e_), __FILE__, __LINE__, #expr); \
      throw std::runtime_error(b_);                                                                  \
    }                                                                                                \
  } while (0)

namespace {

constexpr int F = MPNN_F, H = MPNN_H;
constexpr int OW1 = 0, OB1 = H * F, OW2 = OB1 + H, OB2 = OW2 + H * 3 * H, OW3 = OB2 + H, OB3 = OW3 + 9 * H;
static_assert(OB3 + 9 == MPNN_NPARAMS, "parameter layout");
constexpr int TB = 256;

unsigned blocks_for(int64_t n) { return static_cast<unsigned>(std::max<int64_t>(1, (n + TB - 1) / TB)); }

template <typename T>
T* dnew(int64_t count) {
  T* p = nullptr;
  MCHK(hipMalloc(&p, static_cast<size_t>(std::max<int64_t>(count, 1)) * sizeof(T)));
  return p;
}

__device__ inline double relu(double v) { return v > 0.0 ? v : 0.0; }

// row of every block, node scales s_i = sqrt(3 / trace(A_ii))
__global__ void k_mg_rows(int64_t nb, const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                          const double* __restrict__ val, int32_t* __restrict__ row, double* __restrict__ scale) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (i >= nb) return;
  double tr = 0.0;
  for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
    row[e] = static_cast<int32_t>(i);
    if (idx[e] == i) tr = val[9 * e] + val[9 * e + 4] + val[9 * e + 8];
  }
  scale[i] = tr > 0.0 ? sqrt(3.0 / tr) : 1.0;
}

// rev(e): the position of block (j, i) for block e = (i, j); -1 if absent
__global__ void k_mg_rev(int64_t nnzb, const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                         const int32_t* __restrict__ row, int32_t* __restrict__ rev) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (e >= nnzb) return;
  const int32_t i = row[e], j = idx[e];
  int64_t lo = ptr[j], hi = ptr[j + 1] - 1, found = -1;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) / 2;
    if (idx[mid] == i) {
      found = mid;
      break;
    }
    if (idx[mid] < i) lo = mid + 1;
    else hi = mid - 1;
  }
  rev[e] = static_cast<int32_t>(found);
}

__global__ void k_mg_feat(int64_t nnzb, const int32_t* __restrict__ row, const int32_t* __restrict__ idx,
                          const double* __restrict__ val, const double* __restrict__ scale, double* __restrict__ feat) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (e >= nnzb) return;
  const int32_t i = row[e], j = idx[e];
  const double s = scale[i] * scale[j];
  for (int q = 0; q < 9; ++q) feat[F * e + q] = s * val[9 * e + q];
  feat[F * e + 9] = i == j ? 1.0 : 0.0;
}

__global__ void k_mg_l1(int64_t nnzb, const double* __restrict__ prm, const double* __restrict__ feat,
                        double* __restrict__ pre1) {
  __shared__ double w[H * F + H];
  for (int t = threadIdx.x; t < H * F + H; t += TB) w[t] = prm[OW1 + t];
  __syncthreads();
  const int64_t e = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (e >= nnzb) return;
  double x[F];
#pragma unroll
  for (int q = 0; q < F; ++q) x[q] = feat[F * e + q];
#pragma unroll
  for (int a = 0; a < H; ++a) {
    double v = w[H * F + a];
#pragma unroll
    for (int q = 0; q < F; ++q) v += w[a * F + q] * x[q];
    pre1[H * e + a] = v;
  }
}

__global__ void k_mg_agg(int64_t nb, const int64_t* __restrict__ ptr, const double* __restrict__ pre1,
                         double* __restrict__ node) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (i >= nb) return;
  double acc[H];
#pragma unroll
  for (int a = 0; a < H; ++a) acc[a] = 0.0;
  for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e)
#pragma unroll
    for (int a = 0; a < H; ++a) acc[a] += relu(pre1[H * e + a]);
  const double inv = 1.0 / static_cast<double>(max<int64_t>(1, ptr[i + 1] - ptr[i]));
#pragma unroll
  for (int a = 0; a < H; ++a) node[H * i + a] = acc[a] * inv;
}

__global__ void k_mg_l2(int64_t nnzb, const double* __restrict__ prm, const int32_t* __restrict__ row,
                        const int32_t* __restrict__ idx, const double* __restrict__ pre1,
                        const double* __restrict__ node, double* __restrict__ pre2, double* __restrict__ out) {
  constexpr int NW = H * 3 * H + H + 9 * H + 9;
  __shared__ double w[NW];
  for (int t = threadIdx.x; t < NW; t += TB) w[t] = prm[OW2 + t];
  __syncthreads();
  const double* W2 = w;
  const double* b2 = w + H * 3 * H;
  const double* W3 = b2 + H;
  const double* b3 = W3 + 9 * H;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (e >= nnzb) return;
  const int32_t i = row[e], j = idx[e];
  // h = W2 [m_e; n_i; n_j] + b2, the input streamed entry by entry (each
  // output still sums b = 0 .. 3H-1 in order), so only h stays in registers
  double h[H];
#pragma unroll
  for (int a = 0; a < H; ++a) h[a] = b2[a];
  for (int part = 0; part < 3; ++part) {
    const double* src = part == 0 ? pre1 + H * e : node + H * static_cast<int64_t>(part == 1 ? i : j);
    for (int b = 0; b < H; ++b) {
      const double zb = part == 0 ? relu(src[b]) : src[b];
#pragma unroll
      for (int a = 0; a < H; ++a) h[a] += W2[a * 3 * H + part * H + b] * zb;
    }
  }
#pragma unroll
  for (int a = 0; a < H; ++a) {
    pre2[H * e + a] = h[a];
    h[a] = relu(h[a]);
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    double v = b3[q];
#pragma unroll
    for (int a = 0; a < H; ++a) v += W3[q * H + a] * h[a];
    out[9 * e + q] = v;
  }
}

__global__ void k_mg_out(int64_t nnzb, const int32_t* __restrict__ row, const int32_t* __restrict__ idx,
                         const int32_t* __restrict__ rev, const double* __restrict__ scale,
                         const double* __restrict__ out, double* __restrict__ m) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (e >= nnzb) return;
  const int32_t i = row[e], j = idx[e];
  const int64_t t = rev[e];
  const double s = scale[i] * scale[j];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double g = 0.5 * (out[9 * e + 3 * r + c] + out[9 * t + 3 * c + r]);
      m[9 * e + 3 * r + c] = s * (g + ((i == j && r == c) ? MPNN_OMEGA : 0.0));
    }
}

// dL/dO_e = (s_i s_j) (dL/dM_e + dL/dM_rev(e)^T) / 2
__global__ void k_mg_bout(int64_t nnzb, const int32_t* __restrict__ row, const int32_t* __restrict__ idx,
                          const int32_t* __restrict__ rev, const double* __restrict__ scale,
                          const double* __restrict__ gm, double* __restrict__ go) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (e >= nnzb) return;
  const double s = scale[row[e]] * scale[idx[e]];
  const int64_t t = rev[e];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) go[9 * e + 3 * r + c] = 0.5 * s * (gm[9 * e + 3 * r + c] + gm[9 * t + 3 * c + r]);
}

__global__ void k_mg_b2(int64_t nnzb, const double* __restrict__ prm, const double* __restrict__ pre2,
                        const double* __restrict__ go, double* __restrict__ dpre2, double* __restrict__ dz) {
  __shared__ double W2[H * 3 * H];
  __shared__ double W3[9 * H];
  for (int t = threadIdx.x; t < H * 3 * H; t += TB) W2[t] = prm[OW2 + t];
  for (int t = threadIdx.x; t < 9 * H; t += TB) W3[t] = prm[OW3 + t];
  __syncthreads();
  const int64_t e = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (e >= nnzb) return;
  double d2[H];
#pragma unroll
  for (int a = 0; a < H; ++a) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) v += W3[q * H + a] * go[9 * e + q];
    d2[a] = pre2[H * e + a] > 0.0 ? v : 0.0;
    dpre2[H * e + a] = d2[a];
  }
#pragma unroll
  for (int b = 0; b < 3 * H; ++b) {
    double v = 0.0;
#pragma unroll
    for (int a = 0; a < H; ++a) v += W2[a * 3 * H + b] * d2[a];
    dz[3 * H * e + b] = v;
  }
}

// dL/dn_i: n_i enters z of the blocks of row i (as n_i) and of their reverses (as n_j)
__global__ void k_mg_bnode(int64_t nb, const int64_t* __restrict__ ptr, const int32_t* __restrict__ rev,
                           const double* __restrict__ dz, double* __restrict__ dnode) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (i >= nb) return;
  double acc[H];
#pragma unroll
  for (int a = 0; a < H; ++a) acc[a] = 0.0;
  for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
    const int64_t t = rev[e];
#pragma unroll
    for (int a = 0; a < H; ++a) acc[a] += dz[3 * H * e + H + a] + dz[3 * H * t + 2 * H + a];
  }
#pragma unroll
  for (int a = 0; a < H; ++a) dnode[H * i + a] = acc[a];
}

__global__ void k_mg_b1(int64_t nnzb, const int64_t* __restrict__ ptr, const int32_t* __restrict__ row,
                        const double* __restrict__ pre1, const double* __restrict__ dz,
                        const double* __restrict__ dnode, double* __restrict__ dpre1) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (e >= nnzb) return;
  const int32_t i = row[e];
  const double inv = 1.0 / static_cast<double>(max<int64_t>(1, ptr[i + 1] - ptr[i]));
#pragma unroll
  for (int a = 0; a < H; ++a) {
    const double dm = dz[3 * H * e + a] + dnode[H * i + a] * inv;
    dpre1[H * e + a] = pre1[H * e + a] > 0.0 ? dm : 0.0;
  }
}

// one parameter per workgroup: its gradient summed over the blocks in a fixed order
__global__ void k_mg_wgrad(int64_t nnzb, const int32_t* __restrict__ row, const int32_t* __restrict__ idx,
                           const double* __restrict__ feat, const double* __restrict__ pre1,
                           const double* __restrict__ node, const double* __restrict__ pre2,
                           const double* __restrict__ go, const double* __restrict__ dpre1,
                           const double* __restrict__ dpre2, double* __restrict__ grad) {
  __shared__ double red[TB];
  const int p = blockIdx.x;
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < nnzb; e += TB) {
    double v;
    if (p < OB1) {
      v = dpre1[H * e + p / F] * feat[F * e + p % F];
    } else if (p < OW2) {
      v = dpre1[H * e + (p - OB1)];
    } else if (p < OB2) {
      const int a = (p - OW2) / (3 * H), b = (p - OW2) % (3 * H);
      const double z = b < H ? relu(pre1[H * e + b])
                             : (b < 2 * H ? node[H * static_cast<int64_t>(row[e]) + b - H]
                                          : node[H * static_cast<int64_t>(idx[e]) + b - 2 * H]);
      v = dpre2[H * e + a] * z;
    } else if (p < OW3) {
      v = dpre2[H * e + (p - OB2)];
    } else if (p < OB3) {
      const int q = (p - OW3) / H, a = (p - OW3) % H;
      v = go[9 * e + q] * relu(pre2[H * e + a]);
    } else {
      v = go[9 * e + (p - OB3)];
    }
    acc += v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) grad[p] = red[0];
}

// ------------------------------------------------------------ training loss

// Y = A X for nvec column vectors of length 3 nb (A: 3x3 blocks)
__global__ void k_bsr3_mv(int64_t nb, int nvec, const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                          const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (t >= nb * nvec) return;
  const int64_t i = t % nb, v = t / nb, n = 3 * nb;
  const double* xv = x + v * n;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
    const double* b = val + 9 * e;
    const double* xj = xv + 3 * static_cast<int64_t>(idx[e]);
    a0 += b[0] * xj[0] + b[1] * xj[1] + b[2] * xj[2];
    a1 += b[3] * xj[0] + b[4] * xj[1] + b[5] * xj[2];
    a2 += b[6] * xj[0] + b[7] * xj[1] + b[8] * xj[2];
  }
  y[v * n + 3 * i] = a0;
  y[v * n + 3 * i + 1] = a1;
  y[v * n + 3 * i + 2] = a2;
}

// z = x - y
__global__ void k_sub(int64_t n, const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ z) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (t < n) z[t] = x[t] - y[t];
}

// out[v] = x_v . y_v (one workgroup per vector, fixed order)
__global__ void k_dots(int64_t n, const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ out) {
  __shared__ double red[TB];
  const int v = blockIdx.x;
  double acc = 0.0;
  for (int64_t t = threadIdx.x; t < n; t += TB) acc += x[v * n + t] * y[v * n + t];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[v] = red[0];
}

// g_v = 2 (A e)_v / (nvec den_v)
__global__ void k_gscale(int64_t n, int nvec, const double* __restrict__ ae, const double* __restrict__ den,
                         double* __restrict__ g) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (t >= n * nvec) return;
  g[t] = 2.0 * ae[t] / (static_cast<double>(nvec) * den[t / n]);
}

// gm_e -= sum_v g[i] y[j]^T on the block pattern
__global__ void k_outer(int64_t nnzb, int64_t n, int nvec, const int32_t* __restrict__ row,
                        const int32_t* __restrict__ idx, const double* __restrict__ g, const double* __restrict__ y,
                        double* __restrict__ gm) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x;
  if (e >= nnzb) return;
  const int64_t i = row[e], j = idx[e];
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int v = 0; v < nvec; ++v)
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[3 * r + c] += g[v * n + 3 * i + r] * y[v * n + 3 * j + c];
#pragma unroll
  for (int q = 0; q < 9; ++q) gm[9 * e + q] -= acc[q];
}

}  // namespace

MpnnGraph::MpnnGraph(int64_t nb, int64_t nnzb, const int64_t* d_ptr, const int32_t* d_idx, const double* d_val,
                     hipStream_t s)
    : nb_(nb), nnzb_(nnzb), ptr_(d_ptr), idx_(d_idx), s_(s) {
  row_ = dnew<int32_t>(nnzb);
  rev_ = dnew<int32_t>(nnzb);
  scale_ = dnew<double>(nb);
  feat_ = dnew<double>(nnzb * F);
  pre1_ = dnew<double>(nnzb * H);
  node_ = dnew<double>(nb * H);
  pre2_ = dnew<double>(nnzb * H);
  out_ = dnew<double>(nnzb * 9);
  k_mg_rows<<<blocks_for(nb), TB, 0, s_>>>(nb, d_ptr, d_idx, d_val, row_, scale_);
  k_mg_rev<<<blocks_for(nnzb), TB, 0, s_>>>(nnzb, d_ptr, d_idx, row_, rev_);
  k_mg_feat<<<blocks_for(nnzb), TB, 0, s_>>>(nnzb, row_, d_idx, d_val, scale_, feat_);
  MCHK(hipGetLastError());
  std::vector<int32_t> rv(static_cast<size_t>(nnzb));
  MCHK(hipMemcpyAsync(rv.data(), rev_, nnzb * sizeof(int32_t), hipMemcpyDeviceToHost, s_));
  MCHK(hipStreamSynchronize(s_));
  for (int32_t v : rv)
    if (v < 0) throw std::runtime_error("learned smoother: the block pattern is not structurally symmetric");
}

MpnnGraph::~MpnnGraph() {
  for (void* p : {static_cast<void*>(row_), static_cast<void*>(rev_), static_cast<void*>(scale_),
                  static_cast<void*>(feat_), static_cast<void*>(pre1_), static_cast<void*>(node_),
                  static_cast<void*>(pre2_), static_cast<void*>(out_), static_cast<void*>(go_),
                  static_cast<void*>(dz_), static_cast<void*>(dnode_), static_cast<void*>(dpre1_),
                  static_cast<void*>(dpre2_)})
    if (p) (void)hipFree(p);
}

void MpnnGraph::forward(const double* d_params, double* d_m) {
  k_mg_l1<<<blocks_for(nnzb_), TB, 0, s_>>>(nnzb_, d_params, feat_, pre1_);
  k_mg_agg<<<blocks_for(nb_), TB, 0, s_>>>(nb_, ptr_, pre1_, node_);
  k_mg_l2<<<blocks_for(nnzb_), TB, 0, s_>>>(nnzb_, d_params, row_, idx_, pre1_, node_, pre2_, out_);
  k_mg_out<<<blocks_for(nnzb_), TB, 0, s_>>>(nnzb_, row_, idx_, rev_, scale_, out_, d_m);
  MCHK(hipGetLastError());
}

void MpnnGraph::backward(const double* d_params, const double* d_gm, double* d_grad) {
  if (!go_) {
    go_ = dnew<double>(nnzb_ * 9);
    dz_ = dnew<double>(nnzb_ * 3 * H);
    dnode_ = dnew<double>(nb_ * H);
    dpre1_ = dnew<double>(nnzb_ * H);
    dpre2_ = dnew<double>(nnzb_ * H);
  }
  k_mg_bout<<<blocks_for(nnzb_), TB, 0, s_>>>(nnzb_, row_, idx_, rev_, scale_, d_gm, go_);
  k_mg_b2<<<blocks_for(nnzb_), TB, 0, s_>>>(nnzb_, d_params, pre2_, go_, dpre2_, dz_);
  k_mg_bnode<<<blocks_for(nb_), TB, 0, s_>>>(nb_, ptr_, rev_, dz_, dnode_);
  k_mg_b1<<<blocks_for(nnzb_), TB, 0, s_>>>(nnzb_, ptr_, row_, pre1_, dz_, dnode_, dpre1_);
  k_mg_wgrad<<<static_cast<unsigned>(MPNN_NPARAMS), TB, 0, s_>>>(nnzb_, row_, idx_, feat_, pre1_, node_, pre2_, go_,
                                                                 dpre1_, dpre2_, d_grad);
  MCHK(hipGetLastError());
}

// ------------------------------------------------------------ training handle

namespace {

class TrainerImpl final : public MpnnTrainer {
 public:
  TrainerImpl(int64_t n, const int64_t* colptr, const int64_t* rowidx, const double* val) : n_(n) {
    if (n % 3 != 0) throw std::runtime_error("learned smoother: 3 DOFs per node expected");
    MCHK(hipStreamCreate(&s_));
    nb_ = n / 3;
    const Csr a = csr_from_lower(n, colptr, rowidx, val);
    // 3x3 blocks of nodes i = row / 3 (natural node-major numbering)
    std::vector<int64_t> bptr(nb_ + 1, 0);
    std::vector<int32_t> bidx;
    std::vector<double> bval;
    std::vector<int64_t> pos(nb_, -1);
    for (int64_t bi = 0; bi < nb_; ++bi) {
      std::vector<int32_t> cols;
      for (int r = 0; r < 3; ++r)
        for (int64_t p = a.ptr[3 * bi + r]; p < a.ptr[3 * bi + r + 1]; ++p) cols.push_back(a.idx[p] / 3);
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
      for (int32_t c : cols) {
        pos[c] = static_cast<int64_t>(bidx.size());
        bidx.push_back(c);
      }
      bval.resize(bidx.size() * 9, 0.0);
      for (int r = 0; r < 3; ++r)
        for (int64_t p = a.ptr[3 * bi + r]; p < a.ptr[3 * bi + r + 1]; ++p)
          bval[9 * pos[a.idx[p] / 3] + 3 * r + a.idx[p] % 3] = a.val[p];
      for (int32_t c : cols) pos[c] = -1;
      bptr[bi + 1] = static_cast<int64_t>(bidx.size());
    }
    nnzb_ = bptr[nb_];
    ptr_ = upload(bptr);
    idx_ = upload(bidx);
    val_ = upload(bval);
    hptr_ = bptr;
    hidx_ = bidx;
    graph_ = std::make_unique<MpnnGraph>(nb_, nnzb_, ptr_, idx_, val_, s_);
    prm_ = dnew<double>(MPNN_NPARAMS);
    grad_ = dnew<double>(MPNN_NPARAMS);
    m_ = dnew<double>(nnzb_ * 9);
    gm_ = dnew<double>(nnzb_ * 9);
    MCHK(hipStreamSynchronize(s_));
  }
  ~TrainerImpl() override {
    graph_.reset();
    for (void* p : {static_cast<void*>(ptr_), static_cast<void*>(idx_), static_cast<void*>(val_),
                    static_cast<void*>(prm_), static_cast<void*>(grad_), static_cast<void*>(m_),
                    static_cast<void*>(gm_), static_cast<void*>(vec_), static_cast<void*>(dot_)})
      if (p) (void)hipFree(p);
    if (s_) (void)hipStreamDestroy(s_);
  }

  int64_t nnzb() const override { return nnzb_; }

  void smoother(const double* params, double* blocks, int64_t* ptr, int32_t* idx) override {
    MCHK(hipMemcpyAsync(prm_, params, MPNN_NPARAMS * sizeof(double), hipMemcpyHostToDevice, s_));
    graph_->forward(prm_, m_);
    if (blocks) MCHK(hipMemcpyAsync(blocks, m_, nnzb_ * 9 * sizeof(double), hipMemcpyDeviceToHost, s_));
    MCHK(hipStreamSynchronize(s_));
    if (ptr) std::copy(hptr_.begin(), hptr_.end(), ptr);
    if (idx) std::copy(hidx_.begin(), hidx_.end(), idx);
  }

  double loss(const double* params, const double* e0, int nvec, int nu, double* grad) override {
    if (nvec < 1 || nu < 1) throw std::runtime_error("learned smoother loss: nvec and nu must be positive");
    const int64_t len = n_ * nvec;
    // vectors: E_0..E_nu, Y_0..Y_{nu-1}, T, G, AE
    const int64_t need = len * (2 * static_cast<int64_t>(nu) + 4);
    if (need > cap_) {
      if (vec_) MCHK(hipFree(vec_));
      if (dot_) MCHK(hipFree(dot_));
      vec_ = dnew<double>(need);
      dot_ = dnew<double>(2 * nvec);
      cap_ = need;
    }
    double* E = vec_;
    double* Y = E + len * (nu + 1);
    double* T = Y + len * nu;
    double* G = T + len;
    double* AE = G + len;
    MCHK(hipMemcpyAsync(prm_, params, MPNN_NPARAMS * sizeof(double), hipMemcpyHostToDevice, s_));
    MCHK(hipMemcpyAsync(E, e0, len * sizeof(double), hipMemcpyHostToDevice, s_));
    graph_->forward(prm_, m_);
    const unsigned gv = blocks_for(nb_ * nvec), gl = blocks_for(len);
    for (int k = 0; k < nu; ++k) {
      k_bsr3_mv<<<gv, TB, 0, s_>>>(nb_, nvec, ptr_, idx_, val_, E + k * len, Y + k * len);
      k_bsr3_mv<<<gv, TB, 0, s_>>>(nb_, nvec, ptr_, idx_, m_, Y + k * len, T);
      k_sub<<<gl, TB, 0, s_>>>(len, E + k * len, T, E + (k + 1) * len);
    }
    k_bsr3_mv<<<gv, TB, 0, s_>>>(nb_, nvec, ptr_, idx_, val_, E + nu * len, AE);
    k_dots<<<nvec, TB, 0, s_>>>(n_, E + nu * len, AE, dot_);
    k_dots<<<nvec, TB, 0, s_>>>(n_, E, Y, dot_ + nvec);
    MCHK(hipGetLastError());
    std::vector<double> dh(2 * nvec);
    MCHK(hipMemcpyAsync(dh.data(), dot_, 2 * nvec * sizeof(double), hipMemcpyDeviceToHost, s_));
    if (grad) {
      k_gscale<<<gl, TB, 0, s_>>>(n_, nvec, AE, dot_ + nvec, G);
      MCHK(hipMemsetAsync(gm_, 0, nnzb_ * 9 * sizeof(double), s_));
      for (int k = nu - 1; k >= 0; --k) {
        k_outer<<<blocks_for(nnzb_), TB, 0, s_>>>(nnzb_, n_, nvec, graph_->rows(), idx_, G, Y + k * len, gm_);
        if (k > 0) {
          k_bsr3_mv<<<gv, TB, 0, s_>>>(nb_, nvec, ptr_, idx_, m_, G, T);
          k_bsr3_mv<<<gv, TB, 0, s_>>>(nb_, nvec, ptr_, idx_, val_, T, AE);
          k_sub<<<gl, TB, 0, s_>>>(len, G, AE, G);
        }
      }
      graph_->backward(prm_, gm_, grad_);
      MCHK(hipMemcpyAsync(grad, grad_, MPNN_NPARAMS * sizeof(double), hipMemcpyDeviceToHost, s_));
    }
    MCHK(hipStreamSynchronize(s_));
    double l = 0.0;
    for (int v = 0; v < nvec; ++v) l += dh[v] / dh[nvec + v];
    return l / nvec;
  }

 private:
  template <typename T>
  T* upload(const std::vector<T>& h) {
    T* p = dnew<T>(static_cast<int64_t>(h.size()));
    if (!h.empty()) MCHK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s_));
    return p;
  }

  int64_t n_ = 0, nb_ = 0, nnzb_ = 0, cap_ = 0;
  hipStream_t s_ = nullptr;
  int64_t* ptr_ = nullptr;
  int32_t* idx_ = nullptr;
  double *val_ = nullptr, *prm_ = nullptr, *grad_ = nullptr, *m_ = nullptr, *gm_ = nullptr, *vec_ = nullptr,
         *dot_ = nullptr;
  std::vector<int64_t> hptr_;
  std::vector<int32_t> hidx_;
  std::unique_ptr<MpnnGraph> graph_;
};

}  // namespace

int64_t mpnn_nparams() { return MPNN_NPARAMS; }

std::unique_ptr<MpnnTrainer> make_mpnn_trainer(int64_t n, const int64_t* colptr, const int64_t* rowidx,
                                               const double* val) {
  return std::make_unique<TrainerImpl>(n, colptr, rowidx, val);
}

}  // namespace rf
