// Shift-invert block subspace iteration on one MI355X (src/solver/eigen.rs:
// 229-375): Op x = (1/theta) B x with Op (= K - sigma M, factored by the
// multifrontal solver) applied through its multi-right-hand-side device
// solve.  Per iteration: X~ = Op^-1 Y (one solve of the whole block),
// Z = B X~, the Gram matrices X~^T Y and X~^T Z, the Rayleigh-Ritz step on
// the host (Jacobi on m x m, m <= the block size), and the new block
// X = X~ W, Op X = Y W, B X = Z W.  Every device reduction sums in a fixed
// order (per-chunk partial sums, then the chunks in order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <stdexcept>
#include <string>

#include "eigen.hpp"

namespace rf {

namespace {

#define EHIPCHK(expr)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) {                                                                           \
      char b_[512];                                                                                   \
      snprintf(b_, sizeof b_, "HIP error %s at %s:%d: %s", hipGetErrorString(e_), __FILE__, __LINE__, #expr); \
      throw std::runtime_error(b_);                                                                   \
    }                                                                                                 \
  } while (0)

constexpr int EB = 256;     // threads per block
constexpr int ECH = 2048;   // rows per partial-sum chunk

// z[:, j] = B x[:, j], q columns of n (one thread per row and column)
__global__ __launch_bounds__(EB) void k_bmul(const int64_t* __restrict__ ptr, const int64_t* __restrict__ idx,
                                             const double* __restrict__ val, const double* __restrict__ x,
                                             double* __restrict__ z, int64_t n) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * EB + threadIdx.x;
  if (r >= n) return;
  const int64_t j = blockIdx.y;
  const double* xj = x + j * n;
  double s = 0.0;
  for (int64_t p = ptr[r]; p < ptr[r + 1]; ++p) s = fma(val[p], xj[idx[p]], s);
  z[j * n + r] = s;
}

// partial[chunk][i * mb + j] = sum over the chunk's rows of a[:, i] b[:, j]
__global__ __launch_bounds__(EB) void k_gram_partial(const double* __restrict__ a, const double* __restrict__ b,
                                                     int64_t n, int ma, int mb, double* __restrict__ partial) {
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * ECH, r1 = min(n, r0 + ECH);
  const int np = ma * mb;
  for (int p = threadIdx.x; p < np; p += EB) {
    const int i = p / mb, j = p % mb;
    const double* ai = a + static_cast<int64_t>(i) * n;
    const double* bj = b + static_cast<int64_t>(j) * n;
    double s = 0.0;
    for (int64_t r = r0; r < r1; ++r) s = fma(ai[r], bj[r], s);
    partial[static_cast<int64_t>(blockIdx.x) * np + p] = s;
  }
}

// out[p] = sum over chunks (in order) of partial[c][p]
__global__ __launch_bounds__(EB) void k_chunk_sum(const double* __restrict__ partial, int nchunk, int np,
                                                  double* __restrict__ out) {
  const int p = blockIdx.x * EB + threadIdx.x;
  if (p >= np) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += partial[static_cast<int64_t>(c) * np + p];
  out[p] = s;
}

// out[:, jj] = sum_i in[:, i] w[i * r + jj] (i increasing), jj < r
__global__ __launch_bounds__(EB) void k_combine(const double* __restrict__ in, int64_t n, int m,
                                                const double* __restrict__ w, int r, double* __restrict__ out) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * EB + threadIdx.x;
  if (row >= n) return;
  const int jj = blockIdx.y;
  double s = 0.0;
  for (int i = 0; i < m; ++i) s = fma(in[static_cast<int64_t>(i) * n + row], w[i * r + jj], s);
  out[static_cast<int64_t>(jj) * n + row] = s;
}

// partial[chunk][i] = sum over the chunk of (xb_i - th_i x_i)(y_i - th_i opx_i)
__global__ __launch_bounds__(EB) void k_resdot_partial(const double* __restrict__ xb, const double* __restrict__ x,
                                                       const double* __restrict__ y, const double* __restrict__ opx,
                                                       const double* __restrict__ th, int64_t n, int k,
                                                       double* __restrict__ partial) {
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * ECH, r1 = min(n, r0 + ECH);
  __shared__ double red[EB];
  for (int i = 0; i < k; ++i) {
    const int64_t o = static_cast<int64_t>(i) * n;
    double s = 0.0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += EB)
      s = fma(xb[o + r] - th[i] * x[o + r], y[o + r] - th[i] * opx[o + r], s);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = EB / 2; h > 0; h /= 2) {
      if (static_cast<int>(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[static_cast<int64_t>(blockIdx.x) * k + i] = red[0];
    __syncthreads();
  }
}

// deterministic start vectors (xorshift64, the reference's generator and seed,
// src/solver/eigen.rs:243-252)
struct XorShift {
  uint64_t s = 0x9E3779B97F4A7C15ull;
  double next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return static_cast<double>(s >> 11) / static_cast<double>(uint64_t{1} << 53) - 0.5;
  }
};

struct DevBuf {
  double* p = nullptr;
  explicit DevBuf(size_t count) { EHIPCHK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(double))); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

template <typename X>
struct DevArr {
  X* p = nullptr;
  DevArr(const X* h, size_t count) {
    EHIPCHK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(X)));
    if (count) EHIPCHK(hipMemcpy(p, h, count * sizeof(X), hipMemcpyHostToDevice));
  }
  ~DevArr() {
    if (p) (void)hipFree(p);
  }
  DevArr(const DevArr&) = delete;
  DevArr& operator=(const DevArr&) = delete;
};

}  // namespace

void jacobi_eigen(std::vector<double> a, int m, std::vector<double>& d, std::vector<double>& v) {
  v.assign(static_cast<size_t>(m) * m, 0.0);
  for (int i = 0; i < m; ++i) v[static_cast<size_t>(i) * m + i] = 1.0;
  double fro = 0.0;
  for (double x : a) fro += x * x;
  const double scale = std::max(std::sqrt(fro), 1e-300);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j)
        if (j != i) off += a[i * m + j] * a[i * m + j];
    if (std::sqrt(off) <= 1e-15 * scale) break;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = a[p * m + q];
        if (std::fabs(apq) <= 1e-300) continue;
        // rotation that zeroes a[p][q]: tan = t, t the smaller root of
        // t^2 + 2 theta t - 1 = 0 with theta = (a_qq - a_pp) / (2 a_pq)
        const double theta = (a[q * m + q] - a[p * m + p]) / (2.0 * apq);
        double t = (theta > 0 ? 1.0 : theta < 0 ? -1.0 : 0.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        if (theta == 0.0) t = 1.0;
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < m; ++k) {
          const double akp = a[k * m + p], akq = a[k * m + q];
          a[k * m + p] = c * akp - s * akq;
          a[k * m + q] = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {
          const double apk = a[p * m + k], aqk = a[q * m + k];
          a[p * m + k] = c * apk - s * aqk;
          a[q * m + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          const double vkp = v[k * m + p], vkq = v[k * m + q];
          v[k * m + p] = c * vkp - s * vkq;
          v[k * m + q] = s * vkp + c * vkq;
        }
      }
  }
  d.resize(m);
  for (int i = 0; i < m; ++i) d[i] = a[i * m + i];
}

EigenResult subspace_gpu(int64_t n, const int64_t* bptr, const int64_t* bidx, const double* bval,
                         const BlockSolve& solve, const EigenOptions& opt) {
  EigenResult res;
  if (n == 0 || opt.nev <= 0) return res;
  const int nev = static_cast<int>(std::min<int64_t>(opt.nev, n));
  int q = opt.vectors > 0 ? std::max(opt.vectors, nev) : std::min(2 * nev, nev + 8);
  q = static_cast<int>(std::min<int64_t>(q, n));
  const size_t nq = static_cast<size_t>(n) * q;
  DevArr<int64_t> dptr(bptr, n + 1), didx(bidx, bptr[n]);
  DevArr<double> dval(bval, bptr[n]);
  DevBuf y(nq), xb(nq), z(nq), x(nq), opx(nq), bx(nq), tmp(nq), dw(static_cast<size_t>(q) * q), dth(q);
  const int nchunk = static_cast<int>((n + ECH - 1) / ECH);
  DevBuf part(static_cast<size_t>(nchunk) * q * q), gram(static_cast<size_t>(q) * q);
  const unsigned rowblocks = static_cast<unsigned>((n + EB - 1) / EB);
  XorShift rng;
  std::vector<double> h(static_cast<size_t>(n));
  auto bmul = [&](const double* src, double* dst, int cols) {
    if (cols <= 0) return;
    hipLaunchKernelGGL(k_bmul, dim3(rowblocks, cols), dim3(EB), 0, nullptr, dptr.p, didx.p, dval.p, src, dst, n);
    EHIPCHK(hipGetLastError());
  };
  // column c of dst = B * (a fresh random vector)
  auto random_b = [&](double* dst_col) {
    for (int64_t i = 0; i < n; ++i) h[i] = rng.next();
    EHIPCHK(hipMemcpy(tmp.p, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
    bmul(tmp.p, dst_col, 1);
  };
  // full ma x mb Gram matrix a^T b (row-major, host)
  auto gram_host = [&](const double* a, int ma, const double* b, int mb) {
    const int np = ma * mb;
    hipLaunchKernelGGL(k_gram_partial, dim3(nchunk), dim3(EB), 0, nullptr, a, b, n, ma, mb, part.p);
    EHIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_chunk_sum, dim3((np + EB - 1) / EB), dim3(EB), 0, nullptr, part.p, nchunk, np, gram.p);
    EHIPCHK(hipGetLastError());
    std::vector<double> g(np);
    EHIPCHK(hipMemcpy(g.data(), gram.p, np * sizeof(double), hipMemcpyDeviceToHost));
    return g;
  };
  auto combine = [&](const double* in, int m, int r, double* out) {
    hipLaunchKernelGGL(k_combine, dim3(rowblocks, r), dim3(EB), 0, nullptr, in, n, m, dw.p, r, out);
    EHIPCHK(hipGetLastError());
  };
  for (int j = 0; j < q; ++j) random_b(y.p + static_cast<size_t>(j) * n);
  int my = q;           // columns of y
  int nx = 0;           // Ritz vectors held in x
  std::vector<double> theta;
  for (int it = 1; it <= opt.max_iter; ++it) {
    // X~ = Op^-1 Y
    EHIPCHK(hipMemcpy(xb.p, y.p, static_cast<size_t>(n) * my * sizeof(double), hipMemcpyDeviceToDevice));
    EHIPCHK(hipDeviceSynchronize());
    solve(xb.p, my);
    // convergence of the previous Ritz vectors, residual in the Op norm
    if (nx > 0) {
      const int k = std::min(nev, nx);
      EHIPCHK(hipMemcpy(dth.p, theta.data(), k * sizeof(double), hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_resdot_partial, dim3(nchunk), dim3(EB), 0, nullptr, xb.p, x.p, y.p, opx.p, dth.p, n, k,
                         part.p);
      EHIPCHK(hipGetLastError());
      hipLaunchKernelGGL(k_chunk_sum, dim3(1), dim3(EB), 0, nullptr, part.p, nchunk, k, gram.p);
      EHIPCHK(hipGetLastError());
      std::vector<double> rr(k);
      EHIPCHK(hipMemcpy(rr.data(), gram.p, k * sizeof(double), hipMemcpyDeviceToHost));
      bool ok = true;
      for (int i = 0; i < k; ++i) ok = ok && std::sqrt(std::fabs(rr[i])) / std::fabs(theta[i]) <= opt.tol;
      if (ok && (nx >= nev || it > 5)) {
        res.iterations = it - 1;
        res.converged = true;
        res.theta.assign(theta.begin(), theta.begin() + k);
        break;
      }
    }
    const int m = my;
    bmul(xb.p, z.p, m);
    const std::vector<double> gxy = gram_host(xb.p, m, y.p, m), gxz = gram_host(xb.p, m, z.p, m);
    std::vector<double> g(static_cast<size_t>(m) * m), a(static_cast<size_t>(m) * m);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j <= i; ++j) {
        const double gij = 0.5 * (gxy[i * m + j] + gxy[j * m + i]);
        const double aij = 0.5 * (gxz[i * m + j] + gxz[j * m + i]);
        g[i * m + j] = g[j * m + i] = gij;
        a[i * m + j] = a[j * m + i] = aij;
      }
    // unit diagonal of G (vectors of very different magnitudes)
    double gmax = 0.0;
    for (int i = 0; i < m; ++i) gmax = std::max(gmax, g[i * m + i]);
    std::vector<double> sc(m);
    for (int i = 0; i < m; ++i)
      sc[i] = g[i * m + i] > 1e-300 * std::max(gmax, 1e-300) ? 1.0 / std::sqrt(g[i * m + i]) : 0.0;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) {
        g[i * m + j] *= sc[i] * sc[j];
        a[i * m + j] *= sc[i] * sc[j];
      }
    // T = U D^-1/2 over the directions with D > 1e-13 max D
    std::vector<double> d, u;
    jacobi_eigen(g, m, d, u);
    double dmax = 0.0;
    for (double v : d) dmax = std::max(dmax, v);
    if (dmax <= 0.0 || gmax <= 0.0)
      throw std::runtime_error("the eigenvalue problem has no solutions (B is zero on the iteration vectors)");
    std::vector<int> keep;
    for (int j = 0; j < m; ++j)
      if (d[j] > 1e-13 * dmax) keep.push_back(j);
    const int r = static_cast<int>(keep.size());
    std::vector<double> t(static_cast<size_t>(m) * r), at(static_cast<size_t>(m) * r), c(static_cast<size_t>(r) * r);
    for (int jj = 0; jj < r; ++jj) {
      const double s = 1.0 / std::sqrt(d[keep[jj]]);
      for (int i = 0; i < m; ++i) t[i * r + jj] = u[i * m + keep[jj]] * s;
    }
    for (int i = 0; i < m; ++i)
      for (int jj = 0; jj < r; ++jj) {
        double s = 0.0;
        for (int k = 0; k < m; ++k) s += a[i * m + k] * t[k * r + jj];
        at[i * r + jj] = s;
      }
    for (int ii = 0; ii < r; ++ii)
      for (int jj = 0; jj < r; ++jj) {
        double s = 0.0;
        for (int k = 0; k < m; ++k) s += t[k * r + ii] * at[k * r + jj];
        c[ii * r + jj] = s;
      }
    for (int ii = 0; ii < r; ++ii)
      for (int jj = 0; jj < ii; ++jj) c[ii * r + jj] = c[jj * r + ii] = 0.5 * (c[ii * r + jj] + c[jj * r + ii]);
    std::vector<double> th, v;
    jacobi_eigen(c, r, th, v);
    std::vector<int> order(r);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int i, int j) { return std::fabs(th[j]) < std::fabs(th[i]); });
    // W = S T V (m x r, row-major); X = X~ W, Op X = Y W, B X = Z W
    std::vector<double> w(static_cast<size_t>(m) * r);
    for (int i = 0; i < m; ++i)
      for (int jj = 0; jj < r; ++jj) {
        double s = 0.0;
        for (int k = 0; k < r; ++k) s += t[i * r + k] * v[k * r + order[jj]];
        w[i * r + jj] = sc[i] * s;
      }
    EHIPCHK(hipMemcpy(dw.p, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
    combine(xb.p, m, r, x.p);
    combine(y.p, m, r, opx.p);
    combine(z.p, m, r, bx.p);
    nx = r;
    theta.resize(r);
    for (int jj = 0; jj < r; ++jj) theta[jj] = th[order[jj]];
    // next block: Y = B X, topped up with B times new random vectors
    EHIPCHK(hipMemcpy(y.p, bx.p, static_cast<size_t>(n) * r * sizeof(double), hipMemcpyDeviceToDevice));
    my = r;
    while (my < q) random_b(y.p + static_cast<size_t>(my++) * n);
    res.iterations = it;
  }
  if (!res.converged) {
    if (!opt.allow_unconverged || nx == 0) {
      char m[160];
      snprintf(m, sizeof m, "the eigenvalue solver did not converge in %d iterations", opt.max_iter);
      throw std::runtime_error(m);
    }
    const int k = std::min(nev, nx);
    res.theta.assign(theta.begin(), theta.begin() + k);
  }
  // x^T Op x = 1: scale by (x^T B x / theta)^-1/2
  const int k = static_cast<int>(res.theta.size());
  bmul(x.p, z.p, k);
  const std::vector<double> xbx = gram_host(x.p, k, z.p, k);
  res.vectors.resize(static_cast<size_t>(n) * k);
  EHIPCHK(hipMemcpy(res.vectors.data(), x.p, res.vectors.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int i = 0; i < k; ++i) {
    const double s = xbx[i * k + i] / res.theta[i];
    if (s > 0.0) {
      const double f = 1.0 / std::sqrt(s);
      for (int64_t r = 0; r < n; ++r) res.vectors[static_cast<size_t>(i) * n + r] *= f;
    }
  }
  return res;
}

}  // namespace rf
