// IC(0) on the GPU (gfx950): factorization and application of
// A ~ L D L^T with L unit lower triangular on the strictly lower pattern of A
// (src/solver/iterative/csr.rs:219-324).
//
// Both run level by level over the dependency graph of L: row i of L (and of
// the forward solve) needs the rows j < i of its pattern, so a row's level is
// one more than the highest level among them; the backward solve runs on the
// transpose U = L^T with the levels of the reversed graph.  Rows of one level
// are independent and go to one launch.
//
// Factorization (csr.rs:259-293): one wave per row.  The wave holds its row
// (columns and values) in LDS; for every entry j of the row, in increasing
// order, the lanes stride over row j of L, find each column k in the row's
// own columns by binary search, and sum L_ik d_k L_jk; L_ij = (A_ij - sum) /
// d_j.  Then d_i = A_ii (1 + shift) - sum_j L_ij^2 d_j, and a pivot that is
// not above 1e-12 |A_ii| marks the attempt failed; the host retries with the
// shift sequence of csr.rs:250-256 (1e-3, then 4x).
//
// Application (csr.rs:296-316): y = L^-1 r (forward levels), then
// z = L^-T D^-1 y (backward levels on U, the D^-1 folded in), one group of
// G = 16 lanes per row, products summed by lane shuffles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "ic0.hpp"

namespace rf {

#define ICHK(expr)                                                                                   \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) {                                                                          \
      char b_[512];                                                                                  \
      snprintf(b_, sizeof b_, "HIP error %s at %s:%d: %s", hipGetErrorString(e_), __FILE__, __LINE__, #expr); \
      throw std::runtime_error(b_);                                                                  \
    }                                                                                                \
  } while (0)

namespace {

constexpr int TB = 256;  // threads per block of the solve kernels
constexpr int G = 16;    // lanes per row in the solves

template <typename T>
struct Dev {
  T* p = nullptr;
  size_t n = 0;
  Dev() = default;
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = count;
    ICHK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
  }
  void put(const std::vector<T>& v) {
    alloc(v.size());
    if (!v.empty()) ICHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One wave per row of the level list[lb, lb + gridDim.x).
__global__ __launch_bounds__(64) void k_ic0_factor(int64_t lb, const int32_t* __restrict__ list,
                                                   const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                                   const double* __restrict__ aval, double* lval,
                                                   const double* __restrict__ diag, double* d, double shift,
                                                   int32_t* fail) {
  __shared__ int32_t ci[IC0_MAX_ROW];
  __shared__ double cv[IC0_MAX_ROW];
  const int lane = threadIdx.x;
  const int i = list[lb + blockIdx.x];
  const int64_t lo = ptr[i];
  const int len = static_cast<int>(ptr[i + 1] - lo);
  for (int q = lane; q < len; q += 64) {
    ci[q] = idx[lo + q];
    cv[q] = aval[lo + q];
  }
  __syncthreads();
  for (int p = 0; p < len; ++p) {
    const int j = ci[p];
    const int64_t jhi = ptr[j + 1];
    double s = 0.0;
    for (int64_t q = ptr[j] + lane; q < jhi; q += 64) {
      const int k = idx[q];
      // k < j for every entry of row j, so its partner in row i lies in ci[0, p)
      int a = 0, b = p;
      while (a < b) {
        const int m = (a + b) >> 1;
        if (ci[m] < k) a = m + 1;
        else b = m;
      }
      if (a < p && ci[a] == k) s += cv[a] * d[k] * lval[q];
    }
    s = wave_sum(s);
    if (lane == 0) cv[p] = (cv[p] - s) / d[j];
    __syncthreads();
  }
  double t = 0.0;
  for (int q = lane; q < len; q += 64) {
    const double v = cv[q];
    t += v * v * d[ci[q]];
    lval[lo + q] = v;
  }
  t = wave_sum(t);
  if (lane == 0) {
    const double di = diag[i] * (1.0 + shift) - t;
    d[i] = di;
    if (!(di > 1e-12 * fabs(diag[i]))) *fail = 1;
  }
}

// uval[q] = lval[tmap[q]]: the factor's values in transposed (row-of-U) order
__global__ __launch_bounds__(TB) void k_gather(int64_t n, const int64_t* __restrict__ tmap,
                                               const double* __restrict__ src, double* __restrict__ dst) {
  for (int64_t q = static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x; q < n; q += static_cast<int64_t>(gridDim.x) * TB)
    dst[q] = src[tmap[q]];
}

// y_i = r_i - sum_j L_ij y_j for the rows list[lb, le)
__global__ __launch_bounds__(TB) void k_ic0_fwd(int64_t lb, int64_t le, const int32_t* __restrict__ list,
                                                const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                                const double* __restrict__ val, const double* __restrict__ r, double* y,
                                                const int32_t* done) {
  if (done && *done) return;
  const int64_t t = lb + (static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x) / G;
  const int g = threadIdx.x & (G - 1);
  if (t >= le) return;  // whole groups leave together
  const int i = list[t];
  const int64_t hi = ptr[i + 1];
  double s = 0.0;
  for (int64_t p = ptr[i] + g; p < hi; p += G) s += val[p] * y[idx[p]];
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (g == 0) y[i] = r[i] - s;
}

// z_j = y_j / d_j - sum_i U_ji z_i for the rows list[lb, le) of U = L^T
__global__ __launch_bounds__(TB) void k_ic0_bwd(int64_t lb, int64_t le, const int32_t* __restrict__ list,
                                                const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                                const double* __restrict__ val, const double* __restrict__ y,
                                                const double* __restrict__ d, double* z, const int32_t* done) {
  if (done && *done) return;
  const int64_t t = lb + (static_cast<int64_t>(blockIdx.x) * TB + threadIdx.x) / G;
  const int g = threadIdx.x & (G - 1);
  if (t >= le) return;
  const int j = list[t];
  const int64_t hi = ptr[j + 1];
  double s = 0.0;
  for (int64_t p = ptr[j] + g; p < hi; p += G) s += val[p] * z[idx[p]];
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (g == 0) z[j] = y[j] / d[j] - s;
}

// rows grouped by level: list (rows in increasing order within a level) and
// the level offsets
void by_level(const std::vector<int32_t>& lev, int32_t nlev, std::vector<int32_t>& list, std::vector<int64_t>& off) {
  off.assign(static_cast<size_t>(nlev) + 1, 0);
  for (int32_t l : lev) ++off[l + 1];
  for (int32_t l = 0; l < nlev; ++l) off[l + 1] += off[l];
  list.resize(lev.size());
  std::vector<int64_t> pos(off.begin(), off.end() - 1);
  for (size_t i = 0; i < lev.size(); ++i) list[pos[lev[i]]++] = static_cast<int32_t>(i);
}

class Ic0Impl final : public Ic0Gpu {
 public:
  Ic0Impl(const Csr& a, hipStream_t s);
  void apply(const double* r, double* z, const int32_t* done) override;
  double shift() const override { return shift_; }
  int64_t forward_levels() const override { return static_cast<int64_t>(foff_.size()) - 1; }
  int64_t backward_levels() const override { return static_cast<int64_t>(boff_.size()) - 1; }
  double factor_ms() const override { return factor_ms_; }
  int attempts() const override { return attempts_; }

 private:
  int64_t n_ = 0;
  hipStream_t s_;
  std::vector<int64_t> foff_, boff_;
  Dev<int64_t> lptr_, uptr_, tmap_;
  Dev<int32_t> lidx_, uidx_, flist_, blist_, fail_;
  Dev<double> aval_, lval_, uval_, diag_, d_, y_;
  double shift_ = 0.0, factor_ms_ = 0.0;
  int attempts_ = 0;
};

Ic0Impl::Ic0Impl(const Csr& a, hipStream_t s) : n_(a.nrows), s_(s) {
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t n = n_;
  if (a.ncols != n) throw std::runtime_error("IC(0): square matrix required");
  if (n >= (int64_t(1) << 31)) throw std::runtime_error("IC(0): too many rows");
  // strictly lower L row-wise and the diagonal (csr.rs:231-247)
  std::vector<int64_t> lptr(n + 1, 0);
  std::vector<int32_t> lidx;
  std::vector<double> aval, diag(n, 0.0);
  lidx.reserve(static_cast<size_t>(a.nnz() / 2 + 1));
  aval.reserve(static_cast<size_t>(a.nnz() / 2 + 1));
  for (int64_t i = 0; i < n; ++i) {
    int32_t prev = -1;
    for (int64_t p = a.ptr[i]; p < a.ptr[i + 1]; ++p) {
      const int32_t j = a.idx[p];
      if (j < i) {
        if (j <= prev) throw std::runtime_error("IC(0): rows must be sorted without duplicates");
        prev = j;
        lidx.push_back(j);
        aval.push_back(a.val[p]);
      } else if (j == i) {
        diag[i] += a.val[p];
      }
    }
    lptr[i + 1] = static_cast<int64_t>(lidx.size());
    if (lptr[i + 1] - lptr[i] > IC0_MAX_ROW)
      throw std::runtime_error("IC(0): a row of L holds more than " + std::to_string(IC0_MAX_ROW) + " entries");
  }
  const int64_t nnz = lptr[n];
  // forward levels: 1 + the highest level of the row's columns
  std::vector<int32_t> lev(n, 0);
  int32_t nf = n ? 1 : 0;
  for (int64_t i = 0; i < n; ++i) {
    int32_t l = 0;
    for (int64_t p = lptr[i]; p < lptr[i + 1]; ++p) l = std::max(l, lev[lidx[p]] + 1);
    lev[i] = l;
    nf = std::max(nf, l + 1);
  }
  std::vector<int32_t> flist, blist;
  by_level(lev, nf, flist, foff_);
  // U = L^T row-wise (columns increasing) and the map into L's values
  std::vector<int64_t> uptr(n + 1, 0), tmap(nnz);
  std::vector<int32_t> uidx(nnz);
  for (int64_t p = 0; p < nnz; ++p) ++uptr[lidx[p] + 1];
  for (int64_t j = 0; j < n; ++j) uptr[j + 1] += uptr[j];
  {
    std::vector<int64_t> pos(uptr.begin(), uptr.end() - 1);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t p = lptr[i]; p < lptr[i + 1]; ++p) {
        const int64_t q = pos[lidx[p]]++;
        uidx[q] = static_cast<int32_t>(i);
        tmap[q] = p;
      }
  }
  // backward levels on U: row j needs the rows i > j of its pattern
  std::vector<int32_t> blev(n, 0);
  int32_t nbk = n ? 1 : 0;
  for (int64_t j = n - 1; j >= 0; --j) {
    int32_t l = 0;
    for (int64_t q = uptr[j]; q < uptr[j + 1]; ++q) l = std::max(l, blev[uidx[q]] + 1);
    blev[j] = l;
    nbk = std::max(nbk, l + 1);
  }
  by_level(blev, nbk, blist, boff_);

  lptr_.put(lptr);
  lidx_.put(lidx);
  aval_.put(aval);
  diag_.put(diag);
  uptr_.put(uptr);
  uidx_.put(uidx);
  tmap_.put(tmap);
  flist_.put(flist);
  blist_.put(blist);
  lval_.alloc(nnz);
  uval_.alloc(nnz);
  d_.alloc(n);
  y_.alloc(n);
  fail_.alloc(1);

  // factor, with the diagonal shift of csr.rs:250-256 on breakdown
  double shift = 0.0;
  for (;;) {
    ++attempts_;
    ICHK(hipMemsetAsync(fail_.p, 0, sizeof(int32_t), s_));
    for (size_t l = 0; l + 1 < foff_.size(); ++l) {
      const int64_t rows = foff_[l + 1] - foff_[l];
      k_ic0_factor<<<static_cast<unsigned>(rows), 64, 0, s_>>>(foff_[l], flist_.p, lptr_.p, lidx_.p, aval_.p, lval_.p,
                                                              diag_.p, d_.p, shift, fail_.p);
    }
    ICHK(hipGetLastError());
    int32_t fail = 0;
    ICHK(hipMemcpyAsync(&fail, fail_.p, sizeof fail, hipMemcpyDeviceToHost, s_));
    ICHK(hipStreamSynchronize(s_));
    if (!fail) break;
    if (attempts_ >= 40) throw std::runtime_error("IC(0): no positive factorization up to shift " + std::to_string(shift));
    shift = shift == 0.0 ? 1e-3 : shift * 4.0;
  }
  shift_ = shift;
  if (nnz) {
    const unsigned g = static_cast<unsigned>(std::min<int64_t>(1024, (nnz + TB - 1) / TB));
    k_gather<<<g, TB, 0, s_>>>(nnz, tmap_.p, lval_.p, uval_.p);
    ICHK(hipGetLastError());
  }
  ICHK(hipStreamSynchronize(s_));
  factor_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void Ic0Impl::apply(const double* r, double* z, const int32_t* done) {
  for (size_t l = 0; l + 1 < foff_.size(); ++l) {
    const int64_t rows = foff_[l + 1] - foff_[l];
    const unsigned g = static_cast<unsigned>((rows * G + TB - 1) / TB);
    k_ic0_fwd<<<g, TB, 0, s_>>>(foff_[l], foff_[l + 1], flist_.p, lptr_.p, lidx_.p, lval_.p, r, y_.p, done);
  }
  for (size_t l = 0; l + 1 < boff_.size(); ++l) {
    const int64_t rows = boff_[l + 1] - boff_[l];
    const unsigned g = static_cast<unsigned>((rows * G + TB - 1) / TB);
    k_ic0_bwd<<<g, TB, 0, s_>>>(boff_[l], boff_[l + 1], blist_.p, uptr_.p, uidx_.p, uval_.p, y_.p, d_.p, z, done);
  }
  ICHK(hipGetLastError());
}

}  // namespace

std::unique_ptr<Ic0Gpu> make_ic0_gpu(const Csr& a, hipStream_t s) { return std::make_unique<Ic0Impl>(a, s); }

}  // namespace rf
