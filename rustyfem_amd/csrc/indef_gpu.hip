// Indefinite LDL^T and unsymmetric-value LU on one MI355X.
//
// Restates src/sparse/indef.rs (choose, choose_forced, elim1, elim2,
// eliminate, trailing_update, factor_node) and src/sparse/lu.rs (eliminate,
// dense_lu, factor_node) with the same operation order per matrix entry, so
// the pivot decisions, D and L are the ones the CPU algorithm computes:
// * assembly adds the original entries first, then child slot 0, then slot 1;
// * elimination updates use a product and a separate subtraction (this file
//   is compiled without floating-point contraction);
// * the trailing update of each entry is one fused multiply-add chain over the
//   block's pivots in pivot order, subtracted at the end.
//
// Delayed pivots make front sizes depend on the values, so the host drives
// the tree level by level: it builds a level's fronts (candidate equations =
// own equations followed by the children's delayed ones), launches assembly
// and then rounds of (pivot block, trailing update, symmetric copy) until
// every front of the level has finished, and reads back the pivot counts and
// the final candidate order the next level needs.  Every front is a dense
// column-major m x m block (m = kf + u) in a per-level device buffer.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "indef.hpp"

namespace rf {

#define IHIPCHK(expr)                                                                                \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) {                                                                          \
      char b_[512];                                                                                  \
      snprintf(b_, sizeof b_, "HIP error %s at %s:%d: %s", hipGetErrorString(e_), __FILE__, __LINE__, #expr); \
      throw std::runtime_error(b_);                                                                  \
    }                                                                                                \
  } while (0)

namespace {

constexpr int IB = 64;          // pivot block (src/sparse/indef.rs:33, src/sparse/lu.rs:23)
constexpr int ITH = 256;        // threads of the per-front kernels
constexpr int NW = ITH / 64;    // waves per front kernel
constexpr int WCAP = 4096;      // LDS snapshot of an elimination row/column (doubles)
constexpr int TT = 64;          // tile edge (trailing update, symmetric copy, extend-add)
constexpr int QC = 16;          // pivots per LDS stage of the trailing update
constexpr double ZERO_TOL = 1e-12;                 // src/sparse/indef.rs:29, src/sparse/lu.rs:21
constexpr double ALPHA = 0.6403882032022079;       // src/sparse/indef.rs:31 ((1 + sqrt 17) / 8)

struct IRec {
  double* G;             // m x m column-major front
  double* d;             // per candidate position (kf): D diagonal
  double* e;             // 2x2 off-diagonal on the first of a pair
  int64_t* fs;           // equation of each candidate position (pivot order once eliminated)
  int32_t* kind;         // 1 = 1x1, 2 = first of a 2x2, 0 = second
  int32_t* perm;         // LU root: row permutation of the partial-pivoting block
  double* t;             // solve vector (m, plus kf at an LU root)
  const int64_t* rows;   // equations of the u structural rows
  int64_t smap;          // offset of this front's update-row map (into the parent's pivot order)
  int32_t m, kf, k, start, root, parent, child0, child1;
};

struct IState {
  int32_t p, width, done, t_p0, t_npb, t_b1, t_p, npiv;
  int64_t err;
};

struct ITile { int32_t rec, r0, c0, pad; };
struct IEaTile { int32_t crec, prec, x0, y0, nr, cnp; int64_t moff; };

// ---------------------------------------------------------------------------
// block-wide reductions (256 threads, 4 waves)

__device__ __forceinline__ double block_max(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < NW; ++w) r = fmax(r, red[w]);
  return r;
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < NW; ++w) r += red[w];
  return r;
}

// (value, key) maximum with the smallest key among equal values; strict ">"
// from the initial value, as a sequential scan in key order would pick
__device__ __forceinline__ void vk_better(double& v, int64_t& k, double ov, int64_t ok) {
  if (ov > v || (ov == v && ok < k)) {
    v = ov;
    k = ok;
  }
}

__device__ __forceinline__ int64_t block_argmax(double v, int64_t k, double* red, int64_t* redk) {
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(v, o, 64);
    const int64_t ok = __shfl_xor(k, o, 64);
    vk_better(v, k, ov, ok);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = v;
    redk[threadIdx.x >> 6] = k;
  }
  __syncthreads();
  double bv = red[0];
  int64_t bk = redk[0];
  for (int w = 1; w < NW; ++w) vk_better(bv, bk, red[w], redk[w]);
  return bk;
}

// max |c[i]| over i in [p, m), i != x1, i != x2 (src/sparse/indef.rs:277-285)
__device__ double colmax(const double* c, int m, int p, int x1, int x2, double* red) {
  double g = 0.0;
  for (int i = p + static_cast<int>(threadIdx.x); i < m; i += ITH)
    if (i != x1 && i != x2) g = fmax(g, fabs(c[i]));
  return block_max(g, red);
}

// ---------------------------------------------------------------------------
// LDL^T pivoting steps (one block of 256 threads per front)

// symmetric interchange of candidate positions p and q (src/sparse/indef.rs:262-273)
__device__ void iswap(double* G, int m, int kf, int64_t* fs, int p, int q) {
  if (p == q) return;
  const int tid = threadIdx.x;
  for (int c = tid; c < kf; c += ITH) {
    double* col = G + static_cast<size_t>(c) * m;
    const double t = col[p];
    col[p] = col[q];
    col[q] = t;
  }
  __syncthreads();
  double* cp = G + static_cast<size_t>(p) * m;
  double* cq = G + static_cast<size_t>(q) * m;
  for (int i = tid; i < m; i += ITH) {
    const double t = cp[i];
    cp[i] = cq[i];
    cq[i] = t;
  }
  if (tid == 0) {
    const int64_t t = fs[p];
    fs[p] = fs[q];
    fs[q] = t;
  }
  __syncthreads();
}

// 1x1 pivot at p, columns (p, bend) updated for every row (src/sparse/indef.rs:375-389)
__device__ double ielim1(double* G, int m, int p, int bend, double* wl) {
  const int tid = threadIdx.x;
  double* cp = G + static_cast<size_t>(p) * m;
  const double dp = cp[p];
  const int nw = bend > p + 1 ? bend - p - 1 : 0;
  for (int c0 = 0; c0 < nw; c0 += WCAP) {
    const int cn = min(WCAP, nw - c0);
    __syncthreads();
    for (int q = tid; q < cn; q += ITH) wl[q] = cp[p + 1 + c0 + q];
    __syncthreads();
    for (int i = p + 1 + tid; i < m; i += ITH) {
      const double li = cp[i] / dp;
      if (li != 0.0) {
        double* row = G + static_cast<size_t>(p + 1 + c0) * m + i;
        for (int q = 0; q < cn; ++q) row[static_cast<size_t>(q) * m] -= li * wl[q];
      }
    }
  }
  __syncthreads();
  for (int i = p + 1 + tid; i < m; i += ITH) cp[i] = cp[i] / dp;
  __syncthreads();
  return dp;
}

// 2x2 pivot at (p, p+1) (src/sparse/indef.rs:392-411); returns a, b, c
__device__ void ielim2(double* G, int m, int p, int bend, double* wl, double& oa, double& ob, double& oc) {
  const int tid = threadIdx.x;
  double* c0p = G + static_cast<size_t>(p) * m;
  double* c1p = G + static_cast<size_t>(p + 1) * m;
  const double a = c0p[p], b = c0p[p + 1], c = c1p[p + 1];
  const double det = a * c - b * b;
  const int nw = bend > p + 2 ? bend - p - 2 : 0;
  constexpr int H = WCAP / 2;
  for (int q0 = 0; q0 < nw; q0 += H) {
    const int cn = min(H, nw - q0);
    __syncthreads();
    for (int q = tid; q < cn; q += ITH) {
      wl[q] = c0p[p + 2 + q0 + q];
      wl[H + q] = c1p[p + 2 + q0 + q];
    }
    __syncthreads();
    for (int i = p + 2 + tid; i < m; i += ITH) {
      const double x0 = c0p[i], x1 = c1p[i];
      const double l0 = (x0 * c - x1 * b) / det;
      const double l1 = (x1 * a - x0 * b) / det;
      if (l0 != 0.0 || l1 != 0.0) {
        double* row = G + static_cast<size_t>(p + 2 + q0) * m + i;
        for (int q = 0; q < cn; ++q) row[static_cast<size_t>(q) * m] -= l0 * wl[q] + l1 * wl[H + q];
      }
    }
  }
  __syncthreads();
  for (int i = p + 2 + tid; i < m; i += ITH) {
    const double x0 = c0p[i], x1 = c1p[i];
    c0p[i] = (x0 * c - x1 * b) / det;
    c1p[i] = (x1 * a - x0 * b) / det;
  }
  __syncthreads();
  oa = a;
  ob = b;
  oc = c;
}

// threshold pivot search in [p, end) (src/sparse/indef.rs:287-324):
// 0 none, 1 one (j), 2 two (j, r)
__device__ int ichoose(const double* G, int m, int p, int end, const int64_t* fs, const double* scale, double thr,
                       int& oj, int& orr, double* red, int64_t* redk) {
  for (int j = p; j < end; ++j) {
    const double* cj = G + static_cast<size_t>(j) * m;
    const double ajj = cj[j];
    const double sj = scale[fs[j]];
    const double gj = colmax(cj, m, p, j, -1, red);
    if (fabs(ajj) > ZERO_TOL * sj && fabs(ajj) >= thr * gj) {
      oj = j;
      return 1;
    }
    double best = 0.0;
    int64_t bk = INT64_MAX;
    for (int i = p + static_cast<int>(threadIdx.x); i < end; i += ITH)
      if (i != j) {
        const double v = fabs(cj[i]);
        if (v > best) {
          best = v;
          bk = i;
        }
      }
    const int64_t r64 = block_argmax(best, bk, red, redk);
    if (r64 == INT64_MAX) continue;
    const int r = static_cast<int>(r64);
    const double a = ajj, b = cj[r], cc = G[static_cast<size_t>(r) * m + r];
    const double det = a * cc - b * b;
    const double sr = scale[fs[r]];
    if (!(fabs(det) > ZERO_TOL * sj * sr)) continue;
    const double gj2 = colmax(cj, m, p, j, r, red);
    const double gr2 = colmax(G + static_cast<size_t>(r) * m, m, p, r, j, red);
    const double ad = fabs(det);
    if ((fabs(cc) * gj2 + fabs(b) * gr2) / ad <= 1.0 / thr && (fabs(b) * gj2 + fabs(a) * gr2) / ad <= 1.0 / thr) {
      oj = j;
      orr = r;
      return 2;
    }
  }
  return 0;
}

// Bunch-Kaufman choice over the remaining candidates of a root
// (src/sparse/indef.rs:326-372): returns -1 with kind/j/r, or the singular equation
__device__ int64_t ichoose_forced(const double* G, int m, int p, int kf, const int64_t* fs, const double* scale,
                                  int& kind, int& oj, int& orr, double* red, int64_t* redk) {
  double big = -1.0;
  int64_t key = INT64_MAX;
  for (int j = p + static_cast<int>(threadIdx.x); j < kf; j += ITH) {
    const double* cj = G + static_cast<size_t>(j) * m;
    for (int i = j; i < kf; ++i) {
      const double v = fabs(cj[i]);
      if (v > big) {
        big = v;
        key = static_cast<int64_t>(j) * kf + i;
      }
    }
  }
  // the value of the maximum is needed too: reduce (value, key) and read it back
  double bv = big;
  int64_t bk = key;
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(bv, o, 64);
    const int64_t ok = __shfl_xor(bk, o, 64);
    vk_better(bv, bk, ov, ok);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = bv;
    redk[threadIdx.x >> 6] = bk;
  }
  __syncthreads();
  double lam = red[0];
  int64_t lk = redk[0];
  for (int w = 1; w < NW; ++w) vk_better(lam, lk, red[w], redk[w]);
  int bj = p, bi = p;
  if (lk != INT64_MAX) {
    bj = static_cast<int>(lk / kf);
    bi = static_cast<int>(lk % kf);
  } else {
    lam = -1.0;
  }
  auto diag = [&](int x) { return G[static_cast<size_t>(x) * m + x]; };
  auto one = [&](int x) -> int64_t {
    if (fabs(diag(x)) > ZERO_TOL * scale[fs[x]]) {
      kind = 1;
      oj = x;
      return -1;
    }
    return fs[x];
  };
  if (!(lam > ZERO_TOL * fmax(scale[fs[bj]], scale[fs[bi]]))) return fs[bj];
  if (bi == bj) return one(bj);
  const int j = bj, r = bi;
  const double ajj = fabs(diag(j));
  if (ajj >= ALPHA * lam) return one(j);
  double sig = 0.0;
  for (int i = p + static_cast<int>(threadIdx.x); i < kf; i += ITH) {
    if (i == r) continue;
    const double v = fabs(i >= r ? G[static_cast<size_t>(r) * m + i] : G[static_cast<size_t>(i) * m + r]);
    sig = fmax(sig, v);
  }
  sig = block_max(sig, red);
  if (ajj * sig >= ALPHA * lam * lam) return one(j);
  if (fabs(diag(r)) >= ALPHA * sig) return one(r);
  const double a = diag(j), b = G[static_cast<size_t>(j) * m + r], cc = diag(r);
  const double det = a * cc - b * b;
  if (!(fabs(det) > ZERO_TOL * scale[fs[j]] * scale[fs[r]])) return fs[j];
  kind = 2;
  oj = j;
  orr = r;
  return -1;
}

__device__ __forceinline__ void put_piv(const IRec& R, int p, double d, double e, int kind) {
  R.d[p] = d;
  R.e[p] = e;
  R.kind[p] = kind;
}

// One pivot block of every unfinished front of a level (src/sparse/indef.rs:416-500).
// A front finishes when every candidate is eliminated or a block makes no
// progress over the remaining candidates; a root then takes Bunch-Kaufman
// pivots for the rest.  The trailing update and the symmetric copy of the
// block run in the two kernels that follow.
__global__ __launch_bounds__(ITH) void k_ipanel(const IRec* __restrict__ recs, IState* st, int base,
                                                const double* __restrict__ scale, double thr) {
  __shared__ double wl[WCAP];
  __shared__ double red[NW];
  __shared__ int64_t redk[NW];
  const IRec R = recs[base + blockIdx.x];
  IState* S = st + base + blockIdx.x;
  const int tid = threadIdx.x;
  const int done0 = S->done;
  __syncthreads();
  if (done0) {
    if (tid == 0) S->t_npb = 0;
    return;
  }
  double* G = R.G;
  const int m = R.m, kf = R.kf;
  int p = S->p, width = S->width;
  const int p0 = p;
  const int b1 = min(p + width, kf);
  while (p < b1) {
    int j = 0, r = 0;
    const int kd = ichoose(G, m, p, b1, R.fs, scale, thr, j, r, red, redk);
    if (kd == 1) {
      iswap(G, m, kf, R.fs, p, j);
      const double dp = ielim1(G, m, p, b1, wl);
      if (tid == 0) put_piv(R, p, dp, 0.0, 1);
      p += 1;
    } else if (kd == 2) {
      const int rr = (r == p) ? j : r;
      iswap(G, m, kf, R.fs, p, j);
      iswap(G, m, kf, R.fs, p + 1, rr);
      double a, b, c;
      ielim2(G, m, p, b1, wl, a, b, c);
      if (tid == 0) {
        put_piv(R, p, a, b, 2);
        put_piv(R, p + 1, c, 0.0, 0);
      }
      p += 2;
    } else {
      break;
    }
  }
  const int npb = p - p0;
  const bool trail = npb > 0 && b1 < m;
  bool fin = false;
  if (npb == 0) {
    if (b1 == kf) fin = true;
    else width *= 2;
  } else {
    width = IB;
  }
  if (p >= kf) fin = true;
  int64_t err = -1;
  if (fin && R.root && p < kf) {
    while (p < kf) {
      int kd = 0, j = 0, r = 0;
      const int64_t e = ichoose_forced(G, m, p, kf, R.fs, scale, kd, j, r, red, redk);
      if (e >= 0) {
        err = e;
        break;
      }
      if (kd == 1) {
        iswap(G, m, kf, R.fs, p, j);
        const double dp = ielim1(G, m, p, kf, wl);
        if (tid == 0) put_piv(R, p, dp, 0.0, 1);
        p += 1;
      } else {
        const int rr = (r == p) ? j : r;
        iswap(G, m, kf, R.fs, p, j);
        iswap(G, m, kf, R.fs, p + 1, rr);
        double a, b, c;
        ielim2(G, m, p, kf, wl, a, b, c);
        if (tid == 0) {
          put_piv(R, p, a, b, 2);
          put_piv(R, p + 1, c, 0.0, 0);
        }
        p += 2;
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    S->p = p;
    S->width = width;
    S->t_p0 = p0;
    S->t_npb = trail ? npb : 0;
    S->t_b1 = b1;
    S->t_p = p;
    if (fin) {
      S->done = 1;
      S->npiv = p;
      S->err = err;
    }
  }
}

// F[b1:, b1:] -= (L D) L^T of the block's pivots, lower triangle
// (src/sparse/indef.rs:504-532); one fused multiply-add chain per entry
__global__ __launch_bounds__(256) void k_itrail(const ITile* __restrict__ tiles, const IRec* __restrict__ recs,
                                                const IState* __restrict__ st) {
  const ITile T = tiles[blockIdx.x];
  const IState S = st[T.rec];
  const int npb = S.t_npb;
  if (npb == 0) return;
  const int b1 = S.t_b1, p0 = S.t_p0;
  if (T.r0 + TT <= b1 || T.c0 + TT <= b1) return;
  const IRec R = recs[T.rec];
  const int m = R.m;
  double* G = R.G;
  __shared__ double Ws[QC][TT];
  __shared__ double Ls[QC][TT];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  double s[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) s[a][b] = 0.0;
  const double* d = R.d + p0;
  const double* e = R.e + p0;
  const int32_t* kind = R.kind + p0;
  for (int q0 = 0; q0 < npb; q0 += QC) {
    const int qn = min(QC, npb - q0);
    __syncthreads();
    for (int idx = tid; idx < QC * TT; idx += 256) {
      const int qq = idx / TT, rr = idx % TT, q = q0 + qq;
      double w = 0.0, l = 0.0;
      if (qq < qn) {
        const int i = T.r0 + rr, j = T.c0 + rr;
        if (i < m) {
          const double* Li = G + static_cast<size_t>(p0) * m + i;  // Li[q * m] = L(i, p0 + q)
          const int kd = kind[q];
          if (kd == 1) w = Li[static_cast<size_t>(q) * m] * d[q];
          else if (kd == 2) w = Li[static_cast<size_t>(q) * m] * d[q] + Li[static_cast<size_t>(q + 1) * m] * e[q];
          else w = Li[static_cast<size_t>(q - 1) * m] * e[q - 1] + Li[static_cast<size_t>(q) * m] * d[q];
        }
        if (j < m) l = G[static_cast<size_t>(p0 + q) * m + j];
      }
      Ws[qq][rr] = w;
      Ls[qq][rr] = l;
    }
    __syncthreads();
    for (int qq = 0; qq < qn; ++qq) {
      double wa[4], lb[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) wa[a] = Ws[qq][tx * 4 + a];
#pragma unroll
      for (int b = 0; b < 4; ++b) lb[b] = Ls[qq][ty * 4 + b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) s[a][b] = fma(wa[a], lb[b], s[a][b]);
    }
  }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int j = T.c0 + ty * 4 + b;
    if (j < b1 || j >= m) continue;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int i = T.r0 + tx * 4 + a;
      if (i < b1 || i >= m || j > i) continue;
      G[static_cast<size_t>(j) * m + i] -= s[a][b];
    }
  }
}

// upper <- lower for candidate rows [b1, kf), columns [p, i)
// (src/sparse/indef.rs:486-490); init: the whole candidate block (:229-233)
__global__ __launch_bounds__(256) void k_imirror(const ITile* __restrict__ tiles, const IRec* __restrict__ recs,
                                                 const IState* __restrict__ st, int init) {
  const ITile T = tiles[blockIdx.x];
  int b1 = 0, pp = 0;
  if (!init) {
    const IState S = st[T.rec];
    if (S.t_npb == 0) return;
    b1 = S.t_b1;
    pp = S.t_p;
  }
  const IRec R = recs[T.rec];
  const int m = R.m, kf = R.kf;
  if (T.r0 + TT <= b1 || T.r0 >= kf) return;
  double* G = R.G;
  __shared__ double Ts[TT][TT + 1];
  const int tid = threadIdx.x, tx = tid & 63, ty = tid >> 6;
  for (int c = ty; c < TT; c += 4) {
    const int jj = T.c0 + c, i = T.r0 + tx;
    Ts[c][tx] = (jj < kf && i < kf) ? G[static_cast<size_t>(jj) * m + i] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < TT; r += 4) {
    const int i = T.r0 + r, jj = T.c0 + tx;
    if (i >= b1 && i < kf && jj >= pp && jj < i) G[static_cast<size_t>(i) * m + jj] = Ts[tx][r];
  }
}

// ---------------------------------------------------------------------------
// LU pivoting steps

// symmetric interchange, whole rows and columns (src/sparse/lu.rs:279-290)
__device__ void lswap(double* G, int m, int64_t* fs, int p, int q) {
  if (p == q) return;
  const int tid = threadIdx.x;
  for (int c = tid; c < m; c += ITH) {
    double* col = G + static_cast<size_t>(c) * m;
    const double t = col[p];
    col[p] = col[q];
    col[q] = t;
  }
  __syncthreads();
  double* cp = G + static_cast<size_t>(p) * m;
  double* cq = G + static_cast<size_t>(q) * m;
  for (int i = tid; i < m; i += ITH) {
    const double t = cp[i];
    cp[i] = cq[i];
    cq[i] = t;
  }
  if (tid == 0) {
    const int64_t t = fs[p];
    fs[p] = fs[q];
    fs[q] = t;
  }
  __syncthreads();
}

// pivot p: rows inside the block update every later column, rows below the
// block only the block's columns (src/sparse/lu.rs:319-335)
__device__ void lelim(double* G, int m, int p, int b1, double* wl) {
  const int tid = threadIdx.x;
  double* cp = G + static_cast<size_t>(p) * m;
  const double dp = cp[p];
  // block columns (p, b1) for every row below p
  const int nb = b1 - p - 1;
  for (int c0 = 0; c0 < nb; c0 += WCAP) {
    const int cn = min(WCAP, nb - c0);
    __syncthreads();
    for (int q = tid; q < cn; q += ITH) wl[q] = G[static_cast<size_t>(p + 1 + c0 + q) * m + p];
    __syncthreads();
    for (int i = p + 1 + tid; i < m; i += ITH) {
      const double li = cp[i] / dp;
      if (li == 0.0) continue;
      double* row = G + static_cast<size_t>(p + 1 + c0) * m + i;
      for (int q = 0; q < cn; ++q) row[static_cast<size_t>(q) * m] -= li * wl[q];
    }
  }
  // later columns [b1, m) for the block's rows (p, b1): one column per thread
  const int nr = b1 - p - 1;
  for (int r0 = 0; r0 < nr; r0 += WCAP) {
    const int rn = min(WCAP, nr - r0);
    __syncthreads();
    for (int q = tid; q < rn; q += ITH) wl[q] = cp[p + 1 + r0 + q] / dp;
    __syncthreads();
    for (int c = b1 + tid; c < m; c += ITH) {
      double* col = G + static_cast<size_t>(c) * m;
      const double u = col[p];
      double* x = col + p + 1 + r0;
      for (int q = 0; q < rn; ++q)
        if (wl[q] != 0.0) x[q] -= wl[q] * u;
    }
  }
  __syncthreads();
  for (int i = p + 1 + tid; i < m; i += ITH) cp[i] = cp[i] / dp;
  __syncthreads();
}

// partial-pivoting LU of the r x r block left at a root (src/sparse/lu.rs:246-277);
// returns -1 or the singular equation
__device__ int64_t ldense(double* G, int m, int p0, int r, const int64_t* fs, int32_t* perm, const double* scale,
                          double* wl, double* red, int64_t* redk) {
  const int tid = threadIdx.x;
  for (int i = tid; i < r; i += ITH) perm[i] = i;
  __syncthreads();
  auto at = [&](int i, int c) -> double& { return G[static_cast<size_t>(p0 + c) * m + p0 + i]; };
  for (int p = 0; p < r; ++p) {
    double big = 0.0;
    int64_t bk = INT64_MAX;
    for (int i = p + tid; i < r; i += ITH) {
      const double v = fabs(at(i, p));
      if (v > big) {
        big = v;
        bk = i;
      }
    }
    // the winning value, not only its index
    double bv = big;
    int64_t kk = bk;
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(bv, o, 64);
      const int64_t ok = __shfl_xor(kk, o, 64);
      vk_better(bv, kk, ov, ok);
    }
    __syncthreads();
    if ((tid & 63) == 0) {
      red[tid >> 6] = bv;
      redk[tid >> 6] = kk;
    }
    __syncthreads();
    double gv = red[0];
    int64_t gk = redk[0];
    for (int w = 1; w < NW; ++w) vk_better(gv, gk, red[w], redk[w]);
    const int bi = gk == INT64_MAX ? p : static_cast<int>(gk);
    if (gk == INT64_MAX) gv = 0.0;
    if (!(gv > ZERO_TOL * scale[fs[p0 + p]])) return fs[p0 + p];
    if (bi != p) {
      for (int c = tid; c < r; c += ITH) {
        const double t = at(p, c);
        at(p, c) = at(bi, c);
        at(bi, c) = t;
      }
      if (tid == 0) {
        const int32_t t = perm[p];
        perm[p] = perm[bi];
        perm[bi] = t;
      }
      __syncthreads();
    }
    const double dp = at(p, p);
    const int nc = r - p - 1;
    for (int c0 = 0; c0 < nc; c0 += WCAP) {
      const int cn = min(WCAP, nc - c0);
      __syncthreads();
      for (int q = tid; q < cn; q += ITH) wl[q] = at(p, p + 1 + c0 + q);
      __syncthreads();
      for (int i = p + 1 + tid; i < r; i += ITH) {
        const double l = at(i, p) / dp;
        if (l == 0.0) continue;
        double* row = &at(i, p + 1 + c0);
        for (int q = 0; q < cn; ++q) row[static_cast<size_t>(q) * m] -= l * wl[q];
      }
    }
    __syncthreads();
    for (int i = p + 1 + tid; i < r; i += ITH) at(i, p) = at(i, p) / dp;
    __syncthreads();
  }
  return -1;
}

// one pivot block of every unfinished front (src/sparse/lu.rs:292-365)
__global__ __launch_bounds__(ITH) void k_lpanel(const IRec* __restrict__ recs, IState* st, int base,
                                                const double* __restrict__ scale, double thr) {
  __shared__ double wl[WCAP];
  __shared__ double red[NW];
  __shared__ int64_t redk[NW];
  const IRec R = recs[base + blockIdx.x];
  IState* S = st + base + blockIdx.x;
  const int tid = threadIdx.x;
  const int done0 = S->done;
  __syncthreads();
  if (done0) {
    if (tid == 0) S->t_npb = 0;
    return;
  }
  double* G = R.G;
  const int m = R.m, kf = R.kf;
  int p = S->p, width = S->width;
  const int p0 = p;
  const int b1 = min(p + width, kf);
  while (p < b1) {
    int pick = -1;
    for (int j = p; j < b1; ++j) {
      const double* cj = G + static_cast<size_t>(j) * m;
      const double ajj = fabs(cj[j]);
      if (!(ajj > ZERO_TOL * scale[R.fs[j]])) continue;
      const double g = colmax(cj, m, p, j, -1, red);
      if (ajj >= thr * g) {
        pick = j;
        break;
      }
    }
    if (pick < 0) break;
    lswap(G, m, R.fs, p, pick);
    if (tid == 0) put_piv(R, p, G[static_cast<size_t>(p) * m + p], 0.0, 1);
    lelim(G, m, p, b1, wl);
    p += 1;
  }
  const int npb = p - p0;
  const bool trail = npb > 0 && b1 < m;
  bool fin = false;
  if (npb == 0) {
    if (b1 == kf) fin = true;
    else width *= 2;
  } else {
    width = IB;
  }
  if (p >= kf) fin = true;
  int64_t err = -1;
  if (fin && R.root && p < kf) {
    err = ldense(G, m, p, kf - p, R.fs, R.perm, scale, wl, red, redk);
    if (err < 0)
      for (int i = tid; i < kf - p; i += ITH) put_piv(R, p + i, G[static_cast<size_t>(p + i) * m + p + i], 0.0, 3);
  }
  __syncthreads();
  if (tid == 0) {
    S->p = p;
    S->width = width;
    S->t_p0 = p0;
    S->t_npb = trail ? npb : 0;
    S->t_b1 = b1;
    S->t_p = p;
    if (fin) {
      S->done = 1;
      S->npiv = p;
      S->err = err;
    }
  }
}

// F[b1:, b1:] -= L[b1:, block] U[block, b1:] (src/sparse/lu.rs:337-354)
__global__ __launch_bounds__(256) void k_ltrail(const ITile* __restrict__ tiles, const IRec* __restrict__ recs,
                                                const IState* __restrict__ st) {
  const ITile T = tiles[blockIdx.x];
  const IState S = st[T.rec];
  const int npb = S.t_npb;
  if (npb == 0) return;
  const int b1 = S.t_b1, p0 = S.t_p0;
  if (T.r0 + TT <= b1 || T.c0 + TT <= b1) return;
  const IRec R = recs[T.rec];
  const int m = R.m;
  double* G = R.G;
  __shared__ double As[QC][TT];
  __shared__ double Bs[QC][TT + 1];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  double s[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) s[a][b] = 0.0;
  for (int q0 = 0; q0 < npb; q0 += QC) {
    const int qn = min(QC, npb - q0);
    __syncthreads();
    for (int idx = tid; idx < QC * TT; idx += 256) {
      const int qq = idx / TT, rr = idx % TT;
      const int i = T.r0 + rr;
      As[qq][rr] = (qq < qn && i < m) ? G[static_cast<size_t>(p0 + q0 + qq) * m + i] : 0.0;
      const int cc = idx / QC, q2 = idx % QC, j = T.c0 + cc;
      Bs[q2][cc] = (q2 < qn && j < m) ? G[static_cast<size_t>(j) * m + p0 + q0 + q2] : 0.0;
    }
    __syncthreads();
    for (int qq = 0; qq < qn; ++qq) {
      double wa[4], lb[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) wa[a] = As[qq][tx * 4 + a];
#pragma unroll
      for (int b = 0; b < 4; ++b) lb[b] = Bs[qq][ty * 4 + b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) s[a][b] = fma(wa[a], lb[b], s[a][b]);
    }
  }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int j = T.c0 + ty * 4 + b;
    if (j < b1 || j >= m) continue;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int i = T.r0 + tx * 4 + a;
      if (i < b1 || i >= m) continue;
      G[static_cast<size_t>(j) * m + i] -= s[a][b];
    }
  }
}

// ---------------------------------------------------------------------------
// assembly

__global__ void k_iscatter(const int64_t* __restrict__ dst, const int64_t* __restrict__ src,
                           const double* __restrict__ vals, double* G0, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) G0[dst[i]] += vals[src[i]];
}

// child update block -> parent (src/sparse/indef.rs:217-227, src/sparse/lu.rs:205-214)
template <bool LU>
__global__ __launch_bounds__(256) void k_iea(const IEaTile* __restrict__ tiles, const IRec* __restrict__ recs,
                                             const int32_t* __restrict__ emap) {
  const IEaTile T = tiles[blockIdx.x];
  const IRec C = recs[T.crec];
  const IRec P = recs[T.prec];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int x = T.x0 + tx;
  if (x >= T.nr) return;
  const int32_t* mp = emap + T.moff;
  const int a = mp[x];
  const double* src = C.G + static_cast<size_t>(T.cnp) * C.m + T.cnp + x;
  for (int yy = ty; yy < TT; yy += 4) {
    const int y = T.y0 + yy;
    if (y >= T.nr || (!LU && y > x)) continue;
    const double v = src[static_cast<size_t>(y) * C.m];
    const int b = mp[y];
    if (LU) {
      P.G[static_cast<size_t>(b) * P.m + a] += v;
    } else {
      const int ii = max(a, b), jj = min(a, b);
      P.G[static_cast<size_t>(jj) * P.m + ii] += v;
    }
  }
}

// inertia and 2x2 count of the LDL^T pivots (src/sparse/indef.rs:125-154)
__global__ __launch_bounds__(256) void k_inertia(const IRec* __restrict__ recs, const IState* __restrict__ st,
                                                 unsigned long long* cnt) {
  const IRec R = recs[blockIdx.x];
  const int np = st[blockIdx.x].npiv;
  unsigned long long neg = 0, zero = 0, pos = 0, two = 0;
  for (int i = threadIdx.x; i < np; i += 256) {
    const int kd = R.kind[i];
    if (kd == 2) {
      ++two;
      const double a = R.d[i], b = R.e[i], c = R.d[i + 1];
      const double det = a * c - b * b;
      if (det < 0.0) {
        ++neg;
        ++pos;
      } else if (a + c < 0.0) {
        neg += 2;
      } else {
        pos += 2;
      }
    } else if (kd == 1) {
      const double d = R.d[i];
      if (d < 0.0) ++neg;
      else if (d > 0.0) ++pos;
      else ++zero;
    }
  }
  if (neg) atomicAdd(cnt + 0, neg);
  if (zero) atomicAdd(cnt + 1, zero);
  if (pos) atomicAdd(cnt + 2, pos);
  if (two) atomicAdd(cnt + 3, two);
}

// ---------------------------------------------------------------------------
// solves (one block per front, level by level).  Forward: the front vector in
// pivot order gathers the right-hand side of its own equations and the
// children's update vectors, is eliminated with the unit lower factor, and
// leaves its update rows for the parent; the pivots' values go to x after
// D^-1 (LDL^T).  Backward: gathers x of the update rows and back-substitutes.

template <bool LU>
__global__ __launch_bounds__(ITH) void k_ifwd(const IRec* __restrict__ recs, const IState* __restrict__ st,
                                              int base, const int32_t* __restrict__ smap, double* x) {
  const int r = base + blockIdx.x;
  const IRec R = recs[r];
  const int tid = threadIdx.x;
  const int m = R.m, kf = R.kf, np = st[r].npiv;
  double* t = R.t;
  for (int q = tid; q < m; q += ITH) {
    double v = 0.0;
    if (q < kf) {
      const int64_t g = R.fs[q];
      if (g >= R.start && g < R.start + R.k) v = x[g];
    }
    t[q] = v;
  }
  __syncthreads();
  for (int s = 0; s < 2; ++s) {
    const int c = s ? R.child1 : R.child0;
    if (c < 0) continue;
    const IRec C = recs[c];
    const int cnp = st[c].npiv, nr = C.m - cnp;
    const int32_t* mp = smap + C.smap;
    for (int q = tid; q < nr; q += ITH) t[mp[q]] += C.t[cnp + q];
    __syncthreads();
  }
  const double* G = R.G;
  for (int i = 0; i < np; ++i) {
    const double ti = t[i];
    const int first = (!LU && R.kind[i] == 2) ? i + 2 : i + 1;
    if (ti != 0.0) {
      const double* col = G + static_cast<size_t>(i) * m;
      for (int q = first + tid; q < m; q += ITH) t[q] -= col[q] * ti;
    }
    __syncthreads();
  }
  if (LU) {
    if (R.root && np < kf) {
      const int rd = kf - np;
      double* t2 = t + m;
      for (int i = tid; i < rd; i += ITH) t2[i] = t[np + R.perm[i]];
      __syncthreads();
      for (int i = 0; i < rd; ++i) {
        const double ti = t2[i];
        if (ti != 0.0) {
          const double* col = G + static_cast<size_t>(np + i) * m + np;
          for (int q = i + 1 + tid; q < rd; q += ITH) t2[q] -= col[q] * ti;
        }
        __syncthreads();
      }
      for (int i = tid; i < rd; i += ITH) x[R.fs[np + i]] = t2[i];
    }
    for (int i = tid; i < np; i += ITH) x[R.fs[i]] = t[i];
  } else {
    for (int i = tid; i < np; i += ITH) {
      const int kd = R.kind[i];
      if (kd == 1) {
        x[R.fs[i]] = t[i] / R.d[i];
      } else if (kd == 2) {
        const double a = R.d[i], b = R.e[i], c = R.d[i + 1];
        const double det = a * c - b * b;
        const double y0 = t[i], y1 = t[i + 1];
        x[R.fs[i]] = (c * y0 - b * y1) / det;
        x[R.fs[i + 1]] = (a * y1 - b * y0) / det;
      }
    }
  }
}

template <bool LU>
__global__ __launch_bounds__(ITH) void k_ibwd(const IRec* __restrict__ recs, const IState* __restrict__ st,
                                              int base, double* x) {
  __shared__ double red[NW];
  const int r = base + blockIdx.x;
  const IRec R = recs[r];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m = R.m, kf = R.kf, np = st[r].npiv;
  double* t = R.t;
  const double* G = R.G;
  if (LU && R.root && np < kf) {
    const int rd = kf - np;
    double* t2 = t + m;
    for (int i = tid; i < rd; i += ITH) t2[i] = x[R.fs[np + i]];
    __syncthreads();
    for (int i = rd - 1; i >= 0; --i) {
      const double* col = G + static_cast<size_t>(np + i) * m + np;
      const double ti = t2[i] / col[i];
      __syncthreads();
      if (tid == 0) t2[i] = ti;
      for (int q = tid; q < i; q += ITH) t2[q] -= col[q] * ti;
      __syncthreads();
    }
    for (int i = tid; i < rd; i += ITH) x[R.fs[np + i]] = t2[i];
    __syncthreads();
  }
  for (int q = tid; q < m; q += ITH) {
    const int64_t g = q < kf ? R.fs[q] : R.rows[q - kf];
    t[q] = x[g];
  }
  __syncthreads();
  if (LU) {
    // t[i] -= U[i, np:m] t[np:m]
    for (int i = tid; i < np; i += ITH) {
      double s = 0.0;
      for (int q = np; q < m; ++q) s += G[static_cast<size_t>(q) * m + i] * t[q];
      t[i] -= s;
    }
    __syncthreads();
    for (int i = np - 1; i >= 0; --i) {
      const double* col = G + static_cast<size_t>(i) * m;
      const double ti = t[i] / col[i];
      __syncthreads();
      if (tid == 0) t[i] = ti;
      for (int q = tid; q < i; q += ITH) t[q] -= col[q] * ti;
      __syncthreads();
    }
  } else {
    // t[i] -= L[np:m, i]^T t[np:m], one wave per column
    for (int i = wave; i < np; i += NW) {
      const double* col = G + static_cast<size_t>(i) * m;
      double s = 0.0;
      for (int q = np + lane; q < m; q += 64) s += col[q] * t[q];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      if (lane == 0) t[i] -= s;
    }
    __syncthreads();
    for (int i = np - 1; i >= 0; --i) {
      const double* col = G + static_cast<size_t>(i) * m;
      const int skip = R.kind[i] == 2 ? i + 1 : -1;
      double s = 0.0;
      for (int q = i + 1 + tid; q < np; q += ITH)
        if (q != skip) s += col[q] * t[q];
      s = block_sum(s, red);
      if (tid == 0) t[i] -= s;
      __syncthreads();
    }
  }
  for (int i = tid; i < np; i += ITH) x[R.fs[i]] = t[i];
}

// ---------------------------------------------------------------------------
// host side

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  explicit DBuf(size_t n) { alloc(n); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t n) {
    release();
    if (n) IHIPCHK(hipMalloc(&p, n));
    bytes = n;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

template <typename T>
DBuf upload(const std::vector<T>& v, hipStream_t s) {
  DBuf b(v.size() * sizeof(T));
  if (!v.empty()) IHIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return b;
}

struct HFront {
  int32_t rec = -1, level = 0;
  int32_t k = 0, start = 0, u = 0, m = 0, kf = 0, npiv = 0;
  int32_t child[2] = {-1, -1};
  int32_t parent = -1;
  bool root = false;
  std::vector<int64_t> fs;  // candidate equations (initial order, then the pivot order)
};

class IndefImpl final : public IndefGpu {
 public:
  IndefImpl(Symbolic sym, const std::vector<int64_t>& colptr, const std::vector<int64_t>& rowidx, bool lu)
      : sym_(std::move(sym)), colptr_(colptr), rowidx_(rowidx), lu_(lu) {
    IHIPCHK(hipStreamCreate(&stream_));
    const int64_t nf = static_cast<int64_t>(sym_.fronts.size());
    hf_.resize(nf);
    int32_t maxlev = -1;
    for (int64_t f = 0; f < nf; ++f) {
      const Front& F = sym_.fronts[f];
      HFront& H = hf_[f];
      H.k = static_cast<int32_t>(F.k);
      H.start = static_cast<int32_t>(F.start);
      H.u = static_cast<int32_t>(F.rows.size());
      H.parent = static_cast<int32_t>(F.parent);
      H.root = F.parent < 0;
      int32_t lev = 0;
      for (size_t c = 0; c < F.children.size() && c < 2; ++c) {
        H.child[c] = static_cast<int32_t>(F.children[c]);
        lev = std::max(lev, hf_[F.children[c]].level + 1);
      }
      if (F.children.size() > 2) throw std::runtime_error("internal: more than two children");
      H.level = lev;
      maxlev = std::max(maxlev, lev);
    }
    levels_.assign(maxlev + 1, {});
    for (int64_t f = 0; f < nf; ++f) levels_[hf_[f].level].push_back(static_cast<int32_t>(f));
    int32_t r = 0;
    for (auto& lv : levels_)
      for (int32_t f : lv) hf_[f].rec = r++;
    // structural rows of every front, for the backward solve
    std::vector<int64_t> rows_all;
    row_off_.resize(nf);
    for (int64_t f = 0; f < nf; ++f) {
      row_off_[f] = static_cast<int64_t>(rows_all.size());
      rows_all.insert(rows_all.end(), sym_.fronts[f].rows.begin(), sym_.fronts[f].rows.end());
    }
    d_rows_ = upload(rows_all, stream_);
    recs_h_.resize(nf);
    d_recs_.alloc(std::max<int64_t>(nf, 1) * sizeof(IRec));
    d_state_.alloc(std::max<int64_t>(nf, 1) * sizeof(IState));
    IHIPCHK(hipStreamSynchronize(stream_));
  }

  ~IndefImpl() override {
    lvbuf_.clear();
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  void set_caller_map(const std::vector<int64_t>& eq, const std::vector<double>& iso_diag) override {
    eq_ = eq;
    iso_ = iso_diag;
  }

  int64_t factor(const double* vals) override;
  void solve(double* b, int nrhs) override;
  void inertia(int64_t out[3]) const override { std::memcpy(out, inertia_, sizeof inertia_); }
  int64_t factor_size() const override { return factor_size_; }
  void stats(int64_t out[6]) const override {
    out[0] = inertia_[0];
    out[1] = inertia_[1];
    out[2] = inertia_[2];
    out[3] = delayed_;
    out[4] = two_;
    out[5] = factor_size_;
  }
  int64_t pivots(int64_t* eq, double* d, double* e, int32_t* kind) const override;

 private:
  struct Level {
    DBuf mem;         // fronts, then d, e, fs, kind, perm
    size_t d_off = 0, e_off = 0, fs_off = 0, kind_off = 0, perm_off = 0;
    std::vector<size_t> goff;   // per front of the level: element offset of G
    std::vector<size_t> poff;   // per front: offset of its candidate positions
    size_t sum_kf = 0;
  };

  int32_t local_index(const HFront& H, const Front& F, const std::unordered_map<int64_t, int32_t>& dpos,
                      int64_t g) const {
    if (g >= H.start && g < H.start + H.k) return static_cast<int32_t>(g - H.start);
    if (g < H.start) {
      auto it = dpos.find(g);
      if (it == dpos.end()) throw std::runtime_error("internal: delayed variable missing");
      return it->second;
    }
    auto it = std::lower_bound(F.rows.begin(), F.rows.end(), g);
    if (it == F.rows.end() || *it != g) throw std::runtime_error("internal: row missing in the front");
    return H.kf + static_cast<int32_t>(it - F.rows.begin());
  }

  void build_level_tiles(const std::vector<int32_t>& lv, std::vector<ITile>& trail, std::vector<ITile>& mirror) const {
    for (int32_t f : lv) {
      const HFront& H = hf_[f];
      for (int r0 = 0; r0 < H.m; r0 += TT)
        for (int c0 = 0; c0 < H.m; c0 += TT) {
          if (!lu_ && c0 > r0) continue;
          trail.push_back({H.rec, r0, c0, 0});
        }
      if (!lu_)
        for (int r0 = 0; r0 < H.kf; r0 += TT)
          for (int c0 = 0; c0 <= r0; c0 += TT) mirror.push_back({H.rec, r0, c0, 0});
    }
  }

  Symbolic sym_;
  std::vector<int64_t> colptr_, rowidx_;
  bool lu_;
  hipStream_t stream_ = nullptr;
  std::vector<HFront> hf_;
  std::vector<std::vector<int32_t>> levels_;
  std::vector<int64_t> row_off_;
  std::vector<IRec> recs_h_;
  DBuf d_rows_, d_recs_, d_state_, d_scale_, d_vals_, d_smap_, d_tvec_;
  std::vector<Level> lvbuf_;
  std::vector<int64_t> eq_;
  std::vector<double> iso_;
  int64_t inertia_[3] = {0, 0, 0};
  int64_t delayed_ = 0, two_ = 0, factor_size_ = 0;
  bool factored_ = false;
};

int64_t IndefImpl::factor(const double* vals) {
  factored_ = false;
  lvbuf_.clear();
  d_tvec_.release();
  d_smap_.release();
  const int64_t n = sym_.n;
  const size_t nl = rowidx_.size();
  const double thr = lu_ ? std::min(std::max(0.01, 1e-8), 1.0) : std::min(std::max(0.01, 1e-8), 0.5);
  // column scales (src/sparse/indef.rs:98-108, src/sparse/lu.rs:71-87)
  std::vector<double> scale(std::max<int64_t>(n, 1), 0.0);
  for (int t = 0; t < (lu_ ? 2 : 1); ++t)
    for (int64_t c = 0; c < n; ++c)
      for (int64_t p = colptr_[c]; p < colptr_[c + 1]; ++p) {
        const int64_t r = rowidx_[p];
        if (t && r == c) continue;
        const double v = std::fabs(vals[t * nl + p]);
        scale[c] = std::fmax(scale[c], v);
        scale[r] = std::fmax(scale[r], v);
      }
  d_scale_ = upload(scale, stream_);
  d_vals_.alloc(std::max<size_t>((lu_ ? 2 : 1) * nl, 1) * sizeof(double));
  if (nl) IHIPCHK(hipMemcpyAsync(d_vals_.p, vals, (lu_ ? 2 : 1) * nl * sizeof(double), hipMemcpyHostToDevice, stream_));
  IRec* d_recs = d_recs_.as<IRec>();
  IState* d_state = d_state_.as<IState>();
  delayed_ = 0;
  std::vector<int64_t> ferr(hf_.size(), -1);
  for (size_t L = 0; L < levels_.size(); ++L) {
    const std::vector<int32_t>& lv = levels_[L];
    const int nlv = static_cast<int>(lv.size());
    const int base = hf_[lv[0]].rec;
    // candidate lists: own equations, then the children's delayed ones in child order
    for (int32_t f : lv) {
      HFront& H = hf_[f];
      H.fs.clear();
      for (int32_t i = 0; i < H.k; ++i) H.fs.push_back(H.start + i);
      for (int c = 0; c < 2; ++c) {
        if (H.child[c] < 0) continue;
        const HFront& C = hf_[H.child[c]];
        for (int32_t q = C.npiv; q < C.kf; ++q) H.fs.push_back(C.fs[q]);
      }
      H.kf = static_cast<int32_t>(H.fs.size());
      H.m = H.kf + H.u;
      delayed_ += H.kf - H.k;
    }
    // level buffer
    Level LB;
    size_t gsum = 0, ksum = 0;
    for (int32_t f : lv) {
      LB.goff.push_back(gsum);
      LB.poff.push_back(ksum);
      gsum += static_cast<size_t>(hf_[f].m) * hf_[f].m;
      gsum = (gsum + 31) / 32 * 32;
      ksum += hf_[f].kf;
    }
    LB.sum_kf = ksum;
    const size_t kpad = (ksum + 31) / 32 * 32;
    LB.d_off = gsum * 8;
    LB.e_off = LB.d_off + kpad * 8;
    LB.fs_off = LB.e_off + kpad * 8;
    LB.kind_off = LB.fs_off + kpad * 8;
    LB.perm_off = LB.kind_off + kpad * 4;
    LB.mem.alloc(LB.perm_off + kpad * 4 + 256);
    char* mem = LB.mem.as<char>();
    IHIPCHK(hipMemsetAsync(mem, 0, LB.perm_off + kpad * 4, stream_));
    std::vector<int64_t> fs_all(ksum);
    std::vector<IState> st0(nlv);
    for (int i = 0; i < nlv; ++i) {
      HFront& H = hf_[lv[i]];
      std::copy(H.fs.begin(), H.fs.end(), fs_all.begin() + LB.poff[i]);
      IRec& R = recs_h_[H.rec];
      R.G = reinterpret_cast<double*>(mem) + LB.goff[i];
      R.d = reinterpret_cast<double*>(mem + LB.d_off) + LB.poff[i];
      R.e = reinterpret_cast<double*>(mem + LB.e_off) + LB.poff[i];
      R.fs = reinterpret_cast<int64_t*>(mem + LB.fs_off) + LB.poff[i];
      R.kind = reinterpret_cast<int32_t*>(mem + LB.kind_off) + LB.poff[i];
      R.perm = reinterpret_cast<int32_t*>(mem + LB.perm_off) + LB.poff[i];
      R.t = nullptr;
      R.rows = d_rows_.as<int64_t>() + row_off_[lv[i]];
      R.smap = 0;
      R.m = H.m;
      R.kf = H.kf;
      R.k = H.k;
      R.start = H.start;
      R.root = H.root ? 1 : 0;
      R.parent = H.parent >= 0 ? hf_[H.parent].rec : -1;
      R.child0 = H.child[0] >= 0 ? hf_[H.child[0]].rec : -1;
      R.child1 = H.child[1] >= 0 ? hf_[H.child[1]].rec : -1;
      IState s{};
      s.width = IB;
      s.err = -1;
      st0[i] = s;
    }
    if (ksum)
      IHIPCHK(hipMemcpyAsync(mem + LB.fs_off, fs_all.data(), ksum * 8, hipMemcpyHostToDevice, stream_));
    IHIPCHK(hipMemcpyAsync(d_recs + base, recs_h_.data() + base, nlv * sizeof(IRec), hipMemcpyHostToDevice, stream_));
    IHIPCHK(hipMemcpyAsync(d_state + base, st0.data(), nlv * sizeof(IState), hipMemcpyHostToDevice, stream_));
    // original entries and the children's update blocks
    std::vector<int64_t> dst, src;
    std::vector<int32_t> emap;
    std::vector<IEaTile> ea[2];
    for (int i = 0; i < nlv; ++i) {
      const int32_t f = lv[i];
      const HFront& H = hf_[f];
      const Front& F = sym_.fronts[f];
      std::unordered_map<int64_t, int32_t> dpos;
      for (int32_t q = H.k; q < H.kf; ++q) dpos.emplace(H.fs[q], q);
      const size_t g0 = LB.goff[i], m = H.m;
      for (int32_t cc = 0; cc < H.k; ++cc) {
        const int64_t col = H.start + cc;
        for (int64_t p = colptr_[col]; p < colptr_[col + 1]; ++p) {
          const size_t lr = static_cast<size_t>(local_index(H, F, dpos, rowidx_[p]));
          dst.push_back(static_cast<int64_t>(g0 + cc * m + lr));
          src.push_back(p);
          if (lu_ && lr != static_cast<size_t>(cc)) {
            dst.push_back(static_cast<int64_t>(g0 + lr * m + cc));
            src.push_back(static_cast<int64_t>(nl) + p);
          }
        }
      }
      for (int c = 0; c < 2; ++c) {
        if (H.child[c] < 0) continue;
        const HFront& C = hf_[H.child[c]];
        const Front& CF = sym_.fronts[H.child[c]];
        const int32_t nd = C.kf - C.npiv, nr = nd + C.u;
        const int64_t moff = static_cast<int64_t>(emap.size());
        for (int32_t x = 0; x < nr; ++x) {
          const int64_t g = x < nd ? C.fs[C.npiv + x] : CF.rows[x - nd];
          emap.push_back(local_index(H, F, dpos, g));
        }
        for (int32_t x0 = 0; x0 < nr; x0 += TT)
          for (int32_t y0 = 0; y0 < nr; y0 += TT) {
            if (!lu_ && y0 > x0) continue;
            ea[c].push_back({C.rec, H.rec, x0, y0, nr, C.npiv, moff});
          }
      }
    }
    DBuf d_dst = upload(dst, stream_), d_src = upload(src, stream_), d_emap = upload(emap, stream_);
    if (!dst.empty()) {
      const int64_t cnt = static_cast<int64_t>(dst.size());
      k_iscatter<<<static_cast<unsigned>((cnt + 255) / 256), 256, 0, stream_>>>(
          d_dst.as<int64_t>(), d_src.as<int64_t>(), d_vals_.as<double>(), reinterpret_cast<double*>(mem), cnt);
    }
    for (int c = 0; c < 2; ++c) {
      if (ea[c].empty()) continue;
      DBuf dt = upload(ea[c], stream_);
      if (lu_)
        k_iea<true><<<static_cast<unsigned>(ea[c].size()), 256, 0, stream_>>>(dt.as<IEaTile>(), d_recs, d_emap.as<int32_t>());
      else
        k_iea<false><<<static_cast<unsigned>(ea[c].size()), 256, 0, stream_>>>(dt.as<IEaTile>(), d_recs, d_emap.as<int32_t>());
      IHIPCHK(hipStreamSynchronize(stream_));
    }
    IHIPCHK(hipGetLastError());
    std::vector<ITile> trail, mirror;
    build_level_tiles(lv, trail, mirror);
    DBuf d_trail = upload(trail, stream_), d_mirror = upload(mirror, stream_);
    if (!lu_ && !mirror.empty())
      k_imirror<<<static_cast<unsigned>(mirror.size()), 256, 0, stream_>>>(d_mirror.as<ITile>(), d_recs, d_state, 1);
    // rounds of pivot blocks until every front of the level has finished
    int32_t maxkf = 0;
    for (int32_t f : lv) maxkf = std::max(maxkf, hf_[f].kf);
    const int64_t cap = 4 * static_cast<int64_t>(maxkf) + 64;
    int64_t rounds = 0;
    std::vector<IState> sh(nlv);
    int64_t batch = (maxkf + IB - 1) / IB;
    if (batch < 1) batch = 1;
    for (;;) {
      for (int64_t it = 0; it < batch; ++it) {
        if (lu_) {
          k_lpanel<<<nlv, ITH, 0, stream_>>>(d_recs, d_state, base, d_scale_.as<double>(), thr);
          if (!trail.empty())
            k_ltrail<<<static_cast<unsigned>(trail.size()), 256, 0, stream_>>>(d_trail.as<ITile>(), d_recs, d_state);
        } else {
          k_ipanel<<<nlv, ITH, 0, stream_>>>(d_recs, d_state, base, d_scale_.as<double>(), thr);
          if (!trail.empty())
            k_itrail<<<static_cast<unsigned>(trail.size()), 256, 0, stream_>>>(d_trail.as<ITile>(), d_recs, d_state);
          if (!mirror.empty())
            k_imirror<<<static_cast<unsigned>(mirror.size()), 256, 0, stream_>>>(d_mirror.as<ITile>(), d_recs, d_state, 0);
        }
      }
      rounds += batch;
      IHIPCHK(hipGetLastError());
      IHIPCHK(hipMemcpyAsync(sh.data(), d_state + base, nlv * sizeof(IState), hipMemcpyDeviceToHost, stream_));
      IHIPCHK(hipStreamSynchronize(stream_));
      int64_t left = 0;
      for (const IState& s : sh)
        if (!s.done) left = std::max<int64_t>(left, 1);
      if (!left) break;
      if (rounds > cap) throw std::runtime_error("internal: pivot rounds did not finish");
      batch = 1;
    }
    // the pivot order and counts the next level needs
    std::vector<int64_t> fs_fin(ksum);
    if (ksum)
      IHIPCHK(hipMemcpyAsync(fs_fin.data(), mem + LB.fs_off, ksum * 8, hipMemcpyDeviceToHost, stream_));
    IHIPCHK(hipStreamSynchronize(stream_));
    for (int i = 0; i < nlv; ++i) {
      HFront& H = hf_[lv[i]];
      H.npiv = sh[i].npiv;
      std::copy(fs_fin.begin() + LB.poff[i], fs_fin.begin() + LB.poff[i] + H.kf, H.fs.begin());
      ferr[lv[i]] = sh[i].err;
    }
    lvbuf_.push_back(std::move(LB));
  }
  // only roots can fail; report the first failing root in tree order, as the
  // CPU factorization, which factors one root's tree after the other, does
  for (int64_t r : sym_.roots)
    if (ferr[r] >= 0) return ferr[r];
  // inertia and sizes
  const int64_t nf = static_cast<int64_t>(hf_.size());
  inertia_[0] = inertia_[1] = inertia_[2] = 0;
  two_ = 0;
  factor_size_ = 0;
  for (const HFront& H : hf_) {
    const int64_t np = H.npiv, nr = H.m - H.npiv;
    if (lu_) {
      factor_size_ += np * np + 2 * nr * np;
      if (H.root && np < H.kf) factor_size_ += static_cast<int64_t>(H.kf - np) * (H.kf - np);
    } else {
      factor_size_ += np * (np > 0 ? np - 1 : 0) / 2 + nr * np;
    }
  }
  if (!lu_ && nf > 0) {
    DBuf cnt(4 * sizeof(unsigned long long));
    IHIPCHK(hipMemsetAsync(cnt.p, 0, cnt.bytes, stream_));
    k_inertia<<<static_cast<unsigned>(nf), 256, 0, stream_>>>(d_recs, d_state, cnt.as<unsigned long long>());
    unsigned long long c[4];
    IHIPCHK(hipMemcpyAsync(c, cnt.p, sizeof c, hipMemcpyDeviceToHost, stream_));
    IHIPCHK(hipStreamSynchronize(stream_));
    inertia_[0] = static_cast<int64_t>(c[0]);
    inertia_[1] = static_cast<int64_t>(c[1]);
    inertia_[2] = static_cast<int64_t>(c[2]);
    two_ = static_cast<int64_t>(c[3]);
  }
  // solve buffers: front vectors and the update-row maps into each parent's pivot order
  std::vector<int64_t> toff(nf);
  int64_t tsum = 0;
  for (int64_t f = 0; f < nf; ++f) {
    toff[f] = tsum;
    tsum += hf_[f].m + (lu_ && hf_[f].root ? hf_[f].kf : 0);
  }
  d_tvec_.alloc(std::max<int64_t>(tsum, 1) * sizeof(double));
  std::vector<int32_t> smap;
  for (int64_t f = 0; f < nf; ++f) {
    HFront& H = hf_[f];
    IRec& R = recs_h_[H.rec];
    R.t = d_tvec_.as<double>() + toff[f];
    if (H.parent < 0) continue;
    const HFront& P = hf_[H.parent];
    const Front& PF = sym_.fronts[H.parent];
    std::unordered_map<int64_t, int32_t> ppos;
    for (int32_t q = 0; q < P.kf; ++q) ppos.emplace(P.fs[q], q);
    R.smap = static_cast<int64_t>(smap.size());
    const int32_t nd = H.kf - H.npiv, nr = nd + H.u;
    const Front& F = sym_.fronts[f];
    for (int32_t x = 0; x < nr; ++x) {
      const int64_t g = x < nd ? H.fs[H.npiv + x] : F.rows[x - nd];
      auto it = ppos.find(g);
      if (it != ppos.end()) {
        smap.push_back(it->second);
      } else {
        auto jt = std::lower_bound(PF.rows.begin(), PF.rows.end(), g);
        if (jt == PF.rows.end() || *jt != g) throw std::runtime_error("internal: update row missing in the parent");
        smap.push_back(P.kf + static_cast<int32_t>(jt - PF.rows.begin()));
      }
    }
  }
  d_smap_ = upload(smap, stream_);
  if (nf) IHIPCHK(hipMemcpyAsync(d_recs, recs_h_.data(), nf * sizeof(IRec), hipMemcpyHostToDevice, stream_));
  IHIPCHK(hipStreamSynchronize(stream_));
  factored_ = true;
  return -1;
}

void IndefImpl::solve(double* b, int nrhs) {
  if (!factored_) throw std::runtime_error("solve before a successful factorization");
  const int64_t n = sym_.n, nc = static_cast<int64_t>(eq_.size());
  if (nc == 0 || nrhs <= 0) return;
  std::vector<double> xh(static_cast<size_t>(std::max<int64_t>(n, 1)) * nrhs, 0.0);
  for (int j = 0; j < nrhs; ++j)
    for (int64_t i = 0; i < nc; ++i)
      if (eq_[i] >= 0) xh[j * n + eq_[i]] = b[j * nc + i];
  if (n > 0) {
    DBuf dx = upload(xh, stream_);
    const IRec* d_recs = d_recs_.as<IRec>();
    const IState* d_state = d_state_.as<IState>();
    for (int j = 0; j < nrhs; ++j) {
      double* x = dx.as<double>() + static_cast<size_t>(j) * n;
      for (const auto& lv : levels_) {
        const int base = hf_[lv[0]].rec, cnt = static_cast<int>(lv.size());
        if (lu_) k_ifwd<true><<<cnt, ITH, 0, stream_>>>(d_recs, d_state, base, d_smap_.as<int32_t>(), x);
        else k_ifwd<false><<<cnt, ITH, 0, stream_>>>(d_recs, d_state, base, d_smap_.as<int32_t>(), x);
      }
      for (size_t L = levels_.size(); L-- > 0;) {
        const auto& lv = levels_[L];
        const int base = hf_[lv[0]].rec, cnt = static_cast<int>(lv.size());
        if (lu_) k_ibwd<true><<<cnt, ITH, 0, stream_>>>(d_recs, d_state, base, x);
        else k_ibwd<false><<<cnt, ITH, 0, stream_>>>(d_recs, d_state, base, x);
      }
    }
    IHIPCHK(hipGetLastError());
    IHIPCHK(hipMemcpyAsync(xh.data(), dx.p, xh.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    IHIPCHK(hipStreamSynchronize(stream_));
  }
  for (int j = 0; j < nrhs; ++j) {
    double* col = b + static_cast<size_t>(j) * nc;
    for (int64_t i = 0; i < nc; ++i) {
      if (eq_[i] >= 0) col[i] = xh[j * n + eq_[i]];
      else col[i] /= iso_[i];
    }
  }
}

int64_t IndefImpl::pivots(int64_t* eq, double* d, double* e, int32_t* kind) const {
  if (!factored_) return -1;
  int64_t c = 0;
  // per level: d, e, kind of every front, then emit in front (post)order
  std::vector<std::vector<double>> dl(lvbuf_.size()), el(lvbuf_.size());
  std::vector<std::vector<int32_t>> kl(lvbuf_.size()), pl(lvbuf_.size());
  for (size_t L = 0; L < lvbuf_.size(); ++L) {
    const Level& LB = lvbuf_[L];
    const size_t k = LB.sum_kf;
    dl[L].resize(k);
    el[L].resize(k);
    kl[L].resize(k);
    pl[L].resize(k);
    const char* mem = LB.mem.as<char>();
    if (k) {
      IHIPCHK(hipMemcpy(dl[L].data(), mem + LB.d_off, k * 8, hipMemcpyDeviceToHost));
      IHIPCHK(hipMemcpy(el[L].data(), mem + LB.e_off, k * 8, hipMemcpyDeviceToHost));
      IHIPCHK(hipMemcpy(kl[L].data(), mem + LB.kind_off, k * 4, hipMemcpyDeviceToHost));
      IHIPCHK(hipMemcpy(pl[L].data(), mem + LB.perm_off, k * 4, hipMemcpyDeviceToHost));
    }
  }
  for (size_t f = 0; f < hf_.size(); ++f) {
    const HFront& H = hf_[f];
    const size_t L = static_cast<size_t>(H.level);
    const auto& lv = levels_[L];
    const size_t i = static_cast<size_t>(std::find(lv.begin(), lv.end(), static_cast<int32_t>(f)) - lv.begin());
    const size_t po = lvbuf_[L].poff[i];
    const int32_t end = (lu_ && H.root) ? H.kf : H.npiv;
    for (int32_t q = 0; q < end; ++q, ++c) {
      eq[c] = H.fs[q];
      d[c] = dl[L][po + q];
      kind[c] = kl[L][po + q];
      e[c] = (lu_ && q >= H.npiv) ? static_cast<double>(pl[L][po + q - H.npiv]) : el[L][po + q];
    }
  }
  return c;
}

}  // namespace

std::unique_ptr<IndefGpu> make_indef_gpu(Symbolic sym, const std::vector<int64_t>& colptr,
                                         const std::vector<int64_t>& rowidx, bool lu) {
  return std::make_unique<IndefImpl>(std::move(sym), colptr, rowidx, lu);
}

}  // namespace rf
