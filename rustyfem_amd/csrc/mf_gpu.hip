// HIP kernels (gfx950) and device driver of the level-batched multifrontal
// LDL^T, its triangular solves and FP64 iterative refinement.
//
// Numerics per front (first k columns of the m x m lower front, column-major):
//   for each 64-wide pivot block B = [j0, j1):
//     F[j0:m, B] -= L[j0:m, 0:j0] D L[B, 0:j0]^T        (k_gemm, MFMA)
//     LDL^T of F[B, B] in FP64 with the pivot test        (k_diag)
//     L[j1:m, B] = F[j1:m, B] L[B,B]^-T D[B]^-1          (k_trsm, FP64 math)
//   F[k:m, k:m] -= L21 D L21^T (lower tiles)              (k_gemm, MFMA)
// T = double for the FP64 path (v_mfma_f64_16x16x4f64), T = float for the
// mixed path (v_mfma_f32_16x16x4f32, FP32 accumulation); D, the pivot test,
// the TRSM arithmetic and every solve vector stay FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>

#include "gpu.hpp"

namespace rf {

std::string& last_error() {
  static thread_local std::string e;
  return e;
}

#define HIPCHK(expr)                                                                                 \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) {                                                                          \
      char b_[512];                                                                                  \
      snprintf(b_, sizeof b_, "HIP error %s at %s:%d: %s", hipGetErrorString(e_), __FILE__, __LINE__, #expr); \
      throw std::runtime_error(b_);                                                                  \
    }                                                                                                \
  } while (0)

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

// v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16], B[l/16][l%16]; output
// register r of lane l is D[l/16 + 4r][l%16] (checked on gfx950).
template <>
struct Mfma<double> {
  static constexpr int KT = 16;
  using acc_t = v4d;
  __device__ static acc_t op(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int orow(int lane, int r) { return lane / 16 + 4 * r; }
};

// v_mfma_f32_16x16x4f32: same operand layout; output register r of lane l is
// D[4(l/16) + r][l%16] (checked on gfx950).
template <>
struct Mfma<float> {
  static constexpr int KT = 32;
  using acc_t = v4f;
  __device__ static acc_t op(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int orow(int lane, int r) { return 4 * (lane / 16) + r; }
};

// ---------------------------------------------------------------- assembly

template <typename T>
__global__ __launch_bounds__(256) void k_zero(const FrontDesc* __restrict__ fd, const ZeroTile* __restrict__ tiles,
                                              T* __restrict__ arena) {
  const ZeroTile t = tiles[blockIdx.x];
  const FrontDesc d = fd[t.front];
  const int cend = min(t.c0 + ZC, d.m);
  for (int c = t.c0; c < cend; ++c) {
    T* col = arena + d.off + static_cast<int64_t>(c) * d.m;
    for (int r = c + static_cast<int>(threadIdx.x); r < d.m; r += 256) col[r] = T(0);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_scatter(const ScatterTile* __restrict__ tiles, const int64_t* __restrict__ amap,
                                                 const double* __restrict__ val, T* __restrict__ arena) {
  const ScatterTile t = tiles[blockIdx.x];
  for (int64_t p = t.p0 + threadIdx.x; p < t.p1; p += 256) arena[amap[p]] = T(val[p]);
}

// F_parent[map[x], map[y]] += U_child[x, y] for y <= x in one 64 x 64 tile
template <typename T>
__global__ __launch_bounds__(256) void k_extend_add(const FrontDesc* __restrict__ fd, const EaTile* __restrict__ tiles,
                                                    const int32_t* __restrict__ cmap, T* __restrict__ arena) {
  const EaTile t = tiles[blockIdx.x];
  const FrontDesc c = fd[t.child];
  const FrontDesc p = fd[c.parent];
  __shared__ int32_t mx[EA_T], my[EA_T];
  const int32_t* cm = cmap + c.cmap_off;
  const int tid = threadIdx.x;
  if (tid < EA_T) {
    const int x = t.x0 + tid;
    mx[tid] = x < c.u ? cm[x] : 0;
  } else if (tid < 2 * EA_T) {
    const int y = t.y0 + tid - EA_T;
    my[tid - EA_T] = y < c.u ? cm[y] : 0;
  }
  __syncthreads();
  const int lx = tid % EA_T;
  const int x = t.x0 + lx;
  if (x >= c.u) return;
  const T* src = arena + c.off + (c.k + x) + static_cast<int64_t>(c.k) * c.m;
  T* dst = arena + p.off + mx[lx];
  for (int ly = tid / EA_T; ly < EA_T; ly += 256 / EA_T) {
    const int y = t.y0 + ly;
    if (y > x) break;
    dst[static_cast<int64_t>(my[ly]) * p.m] += src[static_cast<int64_t>(y) * c.m];
  }
}

// ------------------------------------------------------------ pivot blocks

// LDL^T of the jb x jb diagonal block in FP64 with the pivot test of
// src/sparse.rs:528-531 (the failing equation with the smallest index is
// kept), then W = L^-T D^-1 of the block in FP64 for the TRSM (column-major
// NB x NB in l11buf, zero outside the block).
template <typename T>
__global__ __launch_bounds__(256) void k_diag(const FrontDesc* __restrict__ fd, const DiagTask* __restrict__ tasks,
                                              T* __restrict__ arena, double* __restrict__ dpiv,
                                              const double* __restrict__ adiag, double* __restrict__ l11buf,
                                              unsigned long long* __restrict__ err) {
  const DiagTask t = tasks[blockIdx.x];
  const FrontDesc d = fd[t.front];
  const int jb = t.jb;
  __shared__ double A[NB][NB + 1];  // A[col][row]
  __shared__ double V[NB][NB + 1];  // V[col][row] = L^-1
  T* base = arena + d.off + t.j0 + static_cast<int64_t>(t.j0) * d.m;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int i = idx % NB, j = idx / NB;
    A[j][i] = (i < jb && j < jb && i >= j) ? static_cast<double>(base[i + static_cast<int64_t>(j) * d.m]) : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < jb; ++j) {
    const double dj = A[j][j];
    if (tid == 0) {
      const int64_t eq = static_cast<int64_t>(d.start) + t.j0 + j;
      const double orig = adiag[eq];
      if (!(orig > 0.0) || !(dj > 1e-12 * orig)) atomicMin(err, static_cast<unsigned long long>(eq));
    }
    __syncthreads();
    for (int i = j + 1 + tid; i < jb; i += 256) A[j][i] = A[j][i] / dj;
    __syncthreads();
    const int rem = jb - j - 1;
    for (int idx = tid; idx < rem * rem; idx += 256) {
      const int i = j + 1 + idx % rem, q = j + 1 + idx / rem;
      if (q <= i) A[q][i] -= A[j][i] * dj * A[j][q];
    }
    __syncthreads();
  }
  // column q of L^-1 by forward substitution (one thread per column)
  if (tid < NB) {
    const int q = tid;
    for (int i = 0; i < NB; ++i) V[q][i] = 0.0;
    if (q < jb) {
      V[q][q] = 1.0;
      for (int i = q + 1; i < jb; ++i) {
        double s = 0.0;
        for (int p = q; p < i; ++p) s = fma(A[p][i], V[q][p], s);
        V[q][i] = -s;
      }
    }
  }
  __syncthreads();
  double* lb = l11buf + static_cast<int64_t>(t.slot) * NB * NB;
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int p = idx % NB, q = idx / NB;
    // W[p][q] = L^-1[q][p] / d_q, stored column-major at p + q NB
    lb[idx] = (p < jb && q < jb && p <= q) ? V[p][q] / A[q][q] : 0.0;
    if (p < jb && q < jb && p > q) base[p + static_cast<int64_t>(q) * d.m] = T(A[q][p]);
  }
  if (tid < jb) dpiv[d.start + t.j0 + tid] = A[tid][tid];
}

// L[r, B] = F[r, B] W for the TRS rows of one tile (FP64 MFMA, in place)
template <typename T>
__global__ __launch_bounds__(256) void k_trsm(const FrontDesc* __restrict__ fd, const TrsmTile* __restrict__ tiles,
                                              T* __restrict__ arena, const double* __restrict__ l11buf) {
  using M = Mfma<double>;
  const TrsmTile t = tiles[blockIdx.x];
  const FrontDesc d = fd[t.front];
  __shared__ double Ws[NB][NB + 2];   // Ws[p][q]
  __shared__ double As[NB][TRS + 2];  // As[p][r]
  const double* lb = l11buf + static_cast<int64_t>(t.slot) * NB * NB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  T* F = arena + d.off;
  const int64_t m = d.m;
  for (int idx = tid; idx < NB * NB; idx += 256) Ws[idx % NB][idx / NB] = lb[idx];
  for (int idx = tid; idx < NB * TRS; idx += 256) {
    const int r = idx % TRS, p = idx / TRS;
    const int row = t.r0 + r;
    As[p][r] = (row < d.m && p < t.jb) ? static_cast<double>(F[row + (t.j0 + p) * m]) : 0.0;
  }
  __syncthreads();
  M::acc_t acc[4];
#pragma unroll
  for (int bj = 0; bj < 4; ++bj) acc[bj] = M::acc_t{0, 0, 0, 0};
#pragma unroll 4
  for (int kk = 0; kk < NB; kk += 4) {
    const double b = As[kk + lane / 16][wave * 16 + lane % 16];
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) acc[bj] = M::op(Ws[kk + lane / 16][bj * 16 + lane % 16], b, acc[bj]);
  }
  const int row = t.r0 + wave * 16 + lane % 16;
  if (row < d.m) {
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int q = bj * 16 + M::orow(lane, i);
        if (q < t.jb) F[row + (t.j0 + q) * m] = T(acc[bj][i]);
      }
  }
}

// C[r0:r0+128, c0:min(c0+TN, cend)) -= (L[r0:, k0:k1] D) L[c0:, k0:k1]^T on MFMA.
// Four waves tile 128 x TN (TN = 64: 4 x 1 waves of 32 x 64; TN = 128: 2 x 2
// waves of 64 x 64).  The MFMA row index runs over C's columns so that lanes
// 0..15 of an output register hold 16 consecutive rows of the column-major
// front (coalesced read-modify-write).  The next K tile is prefetched into
// registers while the current one is multiplied from LDS.
template <typename T, int TN>
__global__ __launch_bounds__(256) void k_gemm(const FrontDesc* __restrict__ fd, const GemmTile* __restrict__ tiles,
                                              T* __restrict__ arena, const double* __restrict__ dpiv) {
  using M = Mfma<T>;
  constexpr int KT = M::KT;
  constexpr int PAD = 4;
  constexpr int WN = TN / 64;
  constexpr int WM = 4 / WN;
  constexpr int BI = GT_M / WM / 16;
  constexpr int BJ = 4;
  constexpr int NA = GT_M * KT / 256;
  constexpr int NBL = TN * KT / 256;
  const GemmTile t = tiles[blockIdx.x];
  const FrontDesc d = fd[t.front];
  __shared__ T As[KT][GT_M + PAD];
  __shared__ T Bs[KT][TN + PAD];
  __shared__ T ds[KT];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  T* F = arena + d.off;
  const int64_t m = d.m;
  typename M::acc_t acc[BI][BJ];
#pragma unroll
  for (int bi = 0; bi < BI; ++bi)
#pragma unroll
    for (int bj = 0; bj < BJ; ++bj) acc[bi][bj] = typename M::acc_t{0, 0, 0, 0};
  T ra[NA], rb[NBL], rd = T(0);
  auto gload = [&](int kb) {
#pragma unroll
    for (int s = 0; s < NA; ++s) {
      const int idx = tid + 256 * s;
      const int i = idx % GT_M, kk = idx / GT_M;
      const int r = t.r0 + i, c = kb + kk;
      ra[s] = (r < d.m && c < t.k1) ? F[r + c * m] : T(0);
    }
#pragma unroll
    for (int s = 0; s < NBL; ++s) {
      const int idx = tid + 256 * s;
      const int j = idx % TN, kk = idx / TN;
      const int r = t.c0 + j, c = kb + kk;
      rb[s] = (r < t.cend && c < t.k1) ? F[r + c * m] : T(0);
    }
    if (tid < KT) rd = (kb + tid < t.k1) ? static_cast<T>(dpiv[d.start + kb + tid]) : T(0);
  };
  gload(t.k0);
  for (int kb = t.k0; kb < t.k1; kb += KT) {
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NA; ++s) {
      const int idx = tid + 256 * s;
      As[idx / GT_M][idx % GT_M] = ra[s];
    }
#pragma unroll
    for (int s = 0; s < NBL; ++s) {
      const int idx = tid + 256 * s;
      Bs[idx / TN][idx % TN] = rb[s];
    }
    if (tid < KT) ds[tid] = rd;
    __syncthreads();
    if (kb + KT < t.k1) gload(kb + KT);
#pragma unroll
    for (int kk = 0; kk < KT; kk += 4) {
      const int kr = kk + lane / 16;
      const T dk = ds[kr];
      T a[BJ], b[BI];
#pragma unroll
      for (int bj = 0; bj < BJ; ++bj) a[bj] = Bs[kr][wn * 64 + bj * 16 + lane % 16];
#pragma unroll
      for (int bi = 0; bi < BI; ++bi) b[bi] = As[kr][wm * (GT_M / WM) + bi * 16 + lane % 16] * dk;
#pragma unroll
      for (int bi = 0; bi < BI; ++bi)
#pragma unroll
        for (int bj = 0; bj < BJ; ++bj) acc[bi][bj] = M::op(a[bj], b[bi], acc[bi][bj]);
    }
  }
#pragma unroll
  for (int bi = 0; bi < BI; ++bi) {
    const int row = t.r0 + wm * (GT_M / WM) + bi * 16 + lane % 16;
    if (row >= d.m) continue;
#pragma unroll
    for (int bj = 0; bj < BJ; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = t.c0 + wn * 64 + bj * 16 + M::orow(lane, r);
        if (col < t.cend) F[row + col * m] -= acc[bi][bj][r];
      }
  }
}

// ------------------------------------------------------------------ solves
// Vectors t (per front, m entries, column per right-hand side) and x
// (internal numbering, column per right-hand side) are FP64.

constexpr int RB = 16;  // right-hand sides per pass in the solve kernels

__global__ __launch_bounds__(256) void k_fwd_init(const FrontDesc* __restrict__ fd, const int32_t* __restrict__ fronts,
                                                  const double* __restrict__ x, double* __restrict__ t, int64_t n,
                                                  int64_t tv, int nrhs) {
  const FrontDesc d = fd[fronts[blockIdx.x]];
  for (int j = 0; j < nrhs; ++j)
    for (int i = threadIdx.x; i < d.m; i += 256)
      t[j * tv + d.toff + i] = i < d.k ? x[j * n + d.start + i] : 0.0;
}

// t_parent[cmap[x]] += t_child[k_child + x] (forward: the child's update vector)
__global__ __launch_bounds__(256) void k_fwd_child(const FrontDesc* __restrict__ fd, const int32_t* __restrict__ children,
                                                   const int32_t* __restrict__ cmap, double* __restrict__ t, int64_t tv,
                                                   int nrhs) {
  const FrontDesc c = fd[children[blockIdx.x]];
  const FrontDesc p = fd[c.parent];
  const int32_t* cm = cmap + c.cmap_off;
  for (int j = 0; j < nrhs; ++j)
    for (int x = threadIdx.x; x < c.u; x += 256) t[j * tv + p.toff + cm[x]] += t[j * tv + c.toff + c.k + x];
}

// forward substitution with the unit lower diagonal block: y = L[B,B]^-1 t[B]
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_diag(const FrontDesc* __restrict__ fd, const DiagTask* __restrict__ tasks,
                                                  const T* __restrict__ arena, double* __restrict__ t, int64_t tv,
                                                  int nrhs) {
  const DiagTask task = tasks[blockIdx.x];
  const FrontDesc d = fd[task.front];
  __shared__ double L[NB][NB + 1];  // L[col][row]
  __shared__ double y[NB][RB + 1];
  const int jb = task.jb, tid = threadIdx.x;
  const T* base = arena + d.off + task.j0 + static_cast<int64_t>(task.j0) * d.m;
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int i = idx % NB, j = idx / NB;
    L[j][i] = (i < jb && j < jb && i > j) ? static_cast<double>(base[i + static_cast<int64_t>(j) * d.m]) : 0.0;
  }
  for (int j0 = 0; j0 < nrhs; j0 += RB) {
    const int nr = min(RB, nrhs - j0);
    __syncthreads();
    for (int idx = tid; idx < NB * RB; idx += 256) {
      const int i = idx % NB, j = idx / NB;
      y[i][j] = (i < jb && j < nr) ? t[(j0 + j) * tv + d.toff + task.j0 + i] : 0.0;
    }
    __syncthreads();
    const int i = tid % NB, jq = tid / NB;
    for (int p = 0; p < jb; ++p) {
      if (i > p && i < jb)
        for (int j = jq; j < nr; j += 4) y[i][j] -= L[p][i] * y[p][j];
      __syncthreads();
    }
    for (int idx = tid; idx < NB * RB; idx += 256) {
      const int ii = idx % NB, j = idx / NB;
      if (ii < jb && j < nr) t[(j0 + j) * tv + d.toff + task.j0 + ii] = y[ii][j];
    }
  }
}

// t[r] -= L[r, B] y[B] for the rows of one tile
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_update(const FrontDesc* __restrict__ fd, const TrsmTile* __restrict__ tiles,
                                                    const T* __restrict__ arena, double* __restrict__ t, int64_t tv,
                                                    int nrhs) {
  const TrsmTile tile = tiles[blockIdx.x];
  const FrontDesc d = fd[tile.front];
  __shared__ double y[RB][NB];
  const int jb = tile.jb, tid = threadIdx.x;
  const int r = tile.r0 + tid;
  const T* row = arena + d.off + r + static_cast<int64_t>(tile.j0) * d.m;
  for (int j0 = 0; j0 < nrhs; j0 += RB) {
    const int nr = min(RB, nrhs - j0);
    __syncthreads();
    for (int idx = tid; idx < NB * RB; idx += 256) {
      const int p = idx % NB, j = idx / NB;
      y[j][p] = (p < jb && j < nr) ? t[(j0 + j) * tv + d.toff + tile.j0 + p] : 0.0;
    }
    __syncthreads();
    if (r < d.m) {
      double acc[RB];
#pragma unroll
      for (int j = 0; j < RB; ++j) acc[j] = 0.0;
      for (int p = 0; p < jb; ++p) {
        const double l = static_cast<double>(row[static_cast<int64_t>(p) * d.m]);
#pragma unroll
        for (int j = 0; j < RB; ++j) acc[j] = fma(l, y[j][p], acc[j]);
      }
      for (int j = 0; j < nr; ++j) t[(j0 + j) * tv + d.toff + r] -= acc[j];
    }
  }
}

__global__ __launch_bounds__(256) void k_fwd_finish(const FrontDesc* __restrict__ fd, const int32_t* __restrict__ fronts,
                                                    const double* __restrict__ t, const double* __restrict__ dpiv,
                                                    double* __restrict__ x, int64_t n, int64_t tv, int nrhs) {
  const FrontDesc d = fd[fronts[blockIdx.x]];
  for (int j = 0; j < nrhs; ++j)
    for (int i = threadIdx.x; i < d.k; i += 256)
      x[j * n + d.start + i] = t[j * tv + d.toff + i] / dpiv[d.start + i];
}

// backward: t[0:k] = x[pivots], t[k + r] = t_parent[cmap[r]] (final values of ancestors)
__global__ __launch_bounds__(256) void k_bwd_init(const FrontDesc* __restrict__ fd, const int32_t* __restrict__ fronts,
                                                  const int32_t* __restrict__ cmap, const double* __restrict__ x,
                                                  double* __restrict__ t, int64_t n, int64_t tv, int nrhs) {
  const FrontDesc d = fd[fronts[blockIdx.x]];
  for (int j = 0; j < nrhs; ++j)
    for (int i = threadIdx.x; i < d.m; i += 256) {
      double v;
      if (i < d.k) {
        v = x[j * n + d.start + i];
      } else {
        const FrontDesc p = fd[d.parent];
        v = t[j * tv + p.toff + cmap[d.cmap_off + i - d.k]];
      }
      t[j * tv + d.toff + i] = v;
    }
}

// partial[tile][j][p] = sum over the tile's rows r of L[r, j0+p] t[r]
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_partial(const FrontDesc* __restrict__ fd, const TrsmTile* __restrict__ tiles,
                                                     const T* __restrict__ arena, const double* __restrict__ t, int64_t tv,
                                                     int nrhs, double* __restrict__ partial) {
  const TrsmTile tile = tiles[blockIdx.x];
  const FrontDesc d = fd[tile.front];
  __shared__ double red[4][NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = tile.r0 + tid;
  const T* row = arena + d.off + r + static_cast<int64_t>(tile.j0) * d.m;
  double* out = partial + static_cast<int64_t>(blockIdx.x) * nrhs * NB;
  for (int j = 0; j < nrhs; ++j) {
    const double tr = r < d.m ? t[j * tv + d.toff + r] : 0.0;
    for (int p = 0; p < tile.jb; ++p) {
      double v = r < d.m ? static_cast<double>(row[static_cast<int64_t>(p) * d.m]) * tr : 0.0;
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[wave][p] = v;
    }
    __syncthreads();
    if (tid < tile.jb) out[j * NB + tid] = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
    __syncthreads();
  }
}

// t[B] = L[B,B]^-T (t[B] - sum of the partials), back substitution
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_diag(const FrontDesc* __restrict__ fd, const DiagTask* __restrict__ tasks,
                                                  const T* __restrict__ arena, double* __restrict__ t, int64_t tv,
                                                  int nrhs, const double* __restrict__ partial, int64_t tile_base) {
  const DiagTask task = tasks[blockIdx.x];
  const FrontDesc d = fd[task.front];
  __shared__ double L[NB][NB + 1];  // L[col][row]
  __shared__ double z[NB][RB + 1];
  const int jb = task.jb, tid = threadIdx.x;
  const T* base = arena + d.off + task.j0 + static_cast<int64_t>(task.j0) * d.m;
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int i = idx % NB, j = idx / NB;
    L[j][i] = (i < jb && j < jb && i > j) ? static_cast<double>(base[i + static_cast<int64_t>(j) * d.m]) : 0.0;
  }
  for (int j0 = 0; j0 < nrhs; j0 += RB) {
    const int nr = min(RB, nrhs - j0);
    __syncthreads();
    for (int idx = tid; idx < NB * RB; idx += 256) {
      const int i = idx % NB, j = idx / NB;
      double v = 0.0;
      if (i < jb && j < nr) {
        v = t[(j0 + j) * tv + d.toff + task.j0 + i];
        const double* pp = partial + (tile_base + task.tile0) * static_cast<int64_t>(nrhs) * NB;
        for (int q = 0; q < task.ntiles; ++q) v -= pp[(static_cast<int64_t>(q) * nrhs + j0 + j) * NB + i];
      }
      z[i][j] = v;
    }
    __syncthreads();
    const int i = tid % NB, jq = tid / NB;
    for (int p = jb - 1; p >= 0; --p) {
      // z[p] is final; remove its contribution from rows i < p: z[i] -= L[p][i] z[p]
      if (i < p)
        for (int j = jq; j < nr; j += 4) z[i][j] -= L[i][p] * z[p][j];
      __syncthreads();
    }
    for (int idx = tid; idx < NB * RB; idx += 256) {
      const int ii = idx % NB, j = idx / NB;
      if (ii < jb && j < nr) t[(j0 + j) * tv + d.toff + task.j0 + ii] = z[ii][j];
    }
  }
}

__global__ __launch_bounds__(256) void k_bwd_finish(const FrontDesc* __restrict__ fd, const int32_t* __restrict__ fronts,
                                                    const double* __restrict__ t, double* __restrict__ x, int64_t n,
                                                    int64_t tv, int nrhs) {
  const FrontDesc d = fd[fronts[blockIdx.x]];
  for (int j = 0; j < nrhs; ++j)
    for (int i = threadIdx.x; i < d.k; i += 256) x[j * n + d.start + i] = t[j * tv + d.toff + i];
}

// ------------------------------------------------------------- refinement

__global__ void k_gather(const int64_t* __restrict__ map, const double* __restrict__ src, double* __restrict__ dst,
                         int64_t cnt) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < cnt) dst[i] = src[map[i]];
}

// r = b - A x for a CSR matrix, 16 lanes per row, column per right-hand side
__global__ __launch_bounds__(256) void k_residual(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                  const double* __restrict__ av, const double* __restrict__ x,
                                                  const double* __restrict__ b, double* __restrict__ r, int64_t n,
                                                  int nrhs) {
  const int64_t row = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / 16;
  const int sub = threadIdx.x % 16;
  if (row >= n) return;
  for (int j = 0; j < nrhs; ++j) {
    const double* xj = x + j * n;
    double s = 0.0;
    for (int64_t p = rp[row] + sub; p < rp[row + 1]; p += 16) s = fma(av[p], xj[ci[p]], s);
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
    if (sub == 0) r[j * n + row] = b[j * n + row] - s;
  }
}

// per-block partial sums of squares of each column (fixed order), then a final pass
__global__ __launch_bounds__(256) void k_sumsq_partial(const double* __restrict__ v, int64_t n, int nrhs,
                                                       double* __restrict__ part) {
  __shared__ double red[256];
  for (int j = 0; j < nrhs; ++j) {
    double s = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256)
      s = fma(v[j * n + i], v[j * n + i], s);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[j * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

__global__ void k_sumsq_final(const double* __restrict__ part, int nblk, int nrhs, double* __restrict__ out) {
  for (int j = threadIdx.x; j < nrhs; j += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[j * nblk + b];
    out[j] = s;
  }
}

__global__ void k_axpy(double* __restrict__ x, const double* __restrict__ c, int64_t cnt) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < cnt) x[i] += c[i];
}

// caller numbering <-> internal numbering (ffi/src/lib.rs:344-355)
__global__ void k_perm_in(const int64_t* __restrict__ eq, const double* __restrict__ bc, double* __restrict__ xi,
                          int64_t nc, int64_t n, int nrhs) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= nc * nrhs) return;
  const int64_t j = g / nc, i = g % nc;
  const int64_t q = eq[i];
  if (q >= 0) xi[j * n + q] = bc[g];
}

__global__ void k_perm_out(const int64_t* __restrict__ eq, const double* __restrict__ iso, const double* __restrict__ xi,
                           double* __restrict__ bc, int64_t nc, int64_t n, int nrhs) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= nc * nrhs) return;
  const int64_t j = g / nc, i = g % nc;
  const int64_t q = eq[i];
  bc[g] = q >= 0 ? xi[j * n + q] : bc[g] / iso[i];
}

// ----------------------------------------------------------------- driver

template <typename X>
X* dalloc(size_t count, size_t& total) {
  X* p = nullptr;
  if (count == 0) count = 1;
  HIPCHK(hipMalloc(&p, count * sizeof(X)));
  total += count * sizeof(X);
  return p;
}

template <typename X>
X* dupload(const std::vector<X>& v, size_t& total, hipStream_t s) {
  X* p = dalloc<X>(v.size(), total);
  if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice, s));
  return p;
}

inline unsigned grid(int64_t count) { return static_cast<unsigned>(count); }

template <typename T>
class Impl final : public GpuFactor {
 public:
  Impl(Plan plan, const std::vector<int64_t>& colptr, const std::vector<int64_t>& rowidx) : P_(std::move(plan)) {
    HIPCHK(hipStreamCreate(&s_));
    HIPCHK(hipEventCreate(&e0_));
    HIPCHK(hipEventCreate(&e1_));
    n_ = P_.n;
    nlower_ = static_cast<int64_t>(rowidx.size());
    d_fd_ = dupload(P_.fd, bytes_, s_);
    d_cmap_ = dupload(P_.cmap, bytes_, s_);
    d_amap_ = dupload(P_.amap, bytes_, s_);
    d_zero_ = dupload(P_.zero, bytes_, s_);
    d_scatter_ = dupload(P_.scatter, bytes_, s_);
    d_ea_ = dupload(P_.ea, bytes_, s_);
    d_gemm_ = dupload(P_.gemm, bytes_, s_);
    d_diag_ = dupload(P_.diag, bytes_, s_);
    d_trsm_ = dupload(P_.trsm, bytes_, s_);
    d_srow_ = dupload(P_.srow, bytes_, s_);
    d_lf_ = dupload(P_.level_fronts, bytes_, s_);
    for (int sl = 0; sl < 2; ++sl) d_cs_[sl] = dupload(P_.child_slot_fronts[sl], bytes_, s_);
    d_arena_ = dalloc<T>(static_cast<size_t>(P_.arena_elems), bytes_);
    d_dpiv_ = dalloc<double>(n_, bytes_);
    d_adiag_ = dalloc<double>(n_, bytes_);
    d_lval_ = dalloc<double>(nlower_, bytes_);
    d_l11_ = dalloc<double>(static_cast<size_t>(std::max(P_.max_slots, 1)) * NB * NB, bytes_);
    d_err_ = dalloc<unsigned long long>(1, bytes_);
    // diagonal positions of the Lower pattern (first entry of every column)
    std::vector<int64_t> dpos(n_);
    for (int64_t c = 0; c < n_; ++c) dpos[c] = colptr[c];
    d_dpos_ = dupload(dpos, bytes_, s_);
    // full symmetric CSR (both triangles) with a gather map into the Lower values
    std::vector<int64_t> cnt(n_ + 1, 0);
    for (int64_t c = 0; c < n_; ++c)
      for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p) {
        cnt[rowidx[p] + 1]++;
        if (rowidx[p] != c) cnt[c + 1]++;
      }
    for (int64_t i = 0; i < n_; ++i) cnt[i + 1] += cnt[i];
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    std::vector<int32_t> ci(cnt[n_]);
    std::vector<int64_t> gmap(cnt[n_]);
    for (int64_t c = 0; c < n_; ++c)
      for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p) {
        const int64_t r = rowidx[p];
        ci[pos[r]] = static_cast<int32_t>(c);
        gmap[pos[r]++] = p;
      }
    for (int64_t c = 0; c < n_; ++c)
      for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p) {
        const int64_t r = rowidx[p];
        if (r != c) {
          ci[pos[c]] = static_cast<int32_t>(r);
          gmap[pos[c]++] = p;
        }
      }
    nfull_ = cnt[n_];
    d_rp_ = dupload(cnt, bytes_, s_);
    d_ci_ = dupload(ci, bytes_, s_);
    d_gmap_ = dupload(gmap, bytes_, s_);
    d_av_ = dalloc<double>(nfull_, bytes_);
    HIPCHK(hipStreamSynchronize(s_));
  }

  ~Impl() override {
    (void)hipStreamSynchronize(s_);
    void* ptrs[] = {d_fd_, d_cmap_, d_amap_, d_zero_, d_scatter_, d_ea_, d_gemm_, d_diag_, d_trsm_, d_srow_, d_lf_, d_cs_[0],
                    d_cs_[1], d_arena_, d_dpiv_, d_adiag_, d_lval_, d_l11_, d_err_, d_dpos_, d_rp_, d_ci_, d_gmap_,
                    d_av_, d_t_, d_x_, d_part_, d_b_, d_r_, d_red_, d_nrm_, d_bc_, d_eq_, d_iso_};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    for (hipEvent_t e : evs_) (void)hipEventDestroy(e);
    (void)hipEventDestroy(e0_);
    (void)hipEventDestroy(e1_);
    (void)hipStreamDestroy(s_);
  }

  int64_t factor(const double* vals, bool on_device) override {
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (vals) HIPCHK(hipMemcpyAsync(d_lval_, vals, nlower_ * sizeof(double), kind, s_));
    HIPCHK(hipMemsetAsync(d_err_, 0xFF, sizeof(unsigned long long), s_));
    HIPCHK(hipEventRecord(e0_, s_));
    launch(k_gather, (n_ + 255) / 256, 256, d_dpos_, d_lval_, d_adiag_, n_);
    launch(k_gather, (nfull_ + 255) / 256, 256, d_gmap_, d_lval_, d_av_, nfull_);
    nev_ = 0;
    tm_.gemm_flops = 0;
    for (const LevelPlan& lp : P_.levels) {
      launch(k_zero<T>, lp.zero.count, 256, d_fd_, d_zero_ + lp.zero.begin, d_arena_);
      launch(k_scatter<T>, lp.scatter.count, 256, d_scatter_ + lp.scatter.begin, d_amap_, d_lval_, d_arena_);
      for (int sl = 0; sl < 2; ++sl)
        launch(k_extend_add<T>, lp.ea[sl].count, 256, d_fd_, d_ea_ + lp.ea[sl].begin, d_cmap_, d_arena_);
      for (const Phase& ph : lp.phases) {
        switch (ph.kind) {
          case PhaseKind::Gemm64:
          case PhaseKind::Gemm128:
            gemm(ph);
            break;
          case PhaseKind::Diag:
            launch(k_diag<T>, ph.r.count, 256, d_fd_, d_diag_ + ph.r.begin, d_arena_, d_dpiv_, d_adiag_, d_l11_,
                   d_err_);
            break;
          case PhaseKind::Trsm:
            launch(k_trsm<T>, ph.r.count, 256, d_fd_, d_trsm_ + ph.r.begin, d_arena_, d_l11_);
            break;
        }
      }
    }
    HIPCHK(hipEventRecord(e1_, s_));
    unsigned long long err = 0;
    HIPCHK(hipMemcpyAsync(&err, d_err_, sizeof err, hipMemcpyDeviceToHost, s_));
    HIPCHK(hipStreamSynchronize(s_));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0_, e1_));
    tm_.factor_ms = ms;
    if (ktime_) {
      tm_.gemm_ms = 0;
      for (int64_t i = 0; i < nev_; ++i) {
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, evs_[2 * i], evs_[2 * i + 1]));
        tm_.gemm_ms += t;
      }
      tm_.gemm_launches = static_cast<double>(nev_);
    }
    return err == ULLONG_MAX ? -1 : static_cast<int64_t>(err);
  }

  void set_caller_map(const std::vector<int64_t>& eq, const std::vector<double>& iso_diag) override {
    nc_ = static_cast<int64_t>(eq.size());
    std::vector<double> idg(nc_, 0.0);
    for (int64_t i = 0; i < nc_; ++i) idg[i] = eq[i] < 0 ? iso_diag[i] : 0.0;
    if (d_eq_) HIPCHK(hipFree(d_eq_));
    if (d_iso_) HIPCHK(hipFree(d_iso_));
    d_eq_ = dupload(eq, bytes_, s_);
    d_iso_ = dupload(idg, bytes_, s_);
    HIPCHK(hipStreamSynchronize(s_));
  }

  void solve(double* b, int nrhs, bool on_device, double tol, int max_steps) override {
    if (nrhs <= 0 || nc_ == 0) return;
    ensure_vectors(nrhs);
    const int64_t cc = nc_ * nrhs;
    if (!on_device) HIPCHK(hipMemcpyAsync(d_bc_, b, cc * sizeof(double), hipMemcpyHostToDevice, s_));
    double* bc = on_device ? b : d_bc_;
    HIPCHK(hipEventRecord(e0_, s_));
    if (n_ > 0) {
      launch(k_perm_in, (cc + 255) / 256, 256, d_eq_, bc, d_b_, nc_, n_, nrhs);
      if (max_steps > 0) {
        refine(nrhs, tol, max_steps);
      } else {
        HIPCHK(hipMemcpyAsync(d_x_, d_b_, n_ * nrhs * sizeof(double), hipMemcpyDeviceToDevice, s_));
        solve_device(d_x_, nrhs);
        tm_.refine_steps = 0;
        tm_.rel_residual = -1.0;
      }
    }
    launch(k_perm_out, (cc + 255) / 256, 256, d_eq_, d_iso_, d_x_, bc, nc_, n_, nrhs);
    HIPCHK(hipEventRecord(e1_, s_));
    if (!on_device) HIPCHK(hipMemcpyAsync(b, d_bc_, cc * sizeof(double), hipMemcpyDeviceToHost, s_));
    HIPCHK(hipStreamSynchronize(s_));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0_, e1_));
    tm_.solve_ms = ms;
  }

  void pivots(double* d) override {
    HIPCHK(hipMemcpyAsync(d, d_dpiv_, n_ * sizeof(double), hipMemcpyDeviceToHost, s_));
    HIPCHK(hipStreamSynchronize(s_));
  }
  const Timings& timings() const override { return tm_; }
  double flops() const override { return P_.flops; }
  size_t device_bytes() const override { return bytes_; }
  void sync() override { HIPCHK(hipStreamSynchronize(s_)); }
  void set_kernel_timing(bool on) override { ktime_ = on; }

 private:
  template <typename K, typename... A>
  void launch(K kernel, int64_t blocks, int threads, A... args) {
    if (blocks <= 0) return;
    hipLaunchKernelGGL(kernel, dim3(grid(blocks)), dim3(threads), 0, s_, args...);
    HIPCHK(hipGetLastError());
  }

  // one launch of the Schur-update GEMM over a tile range
  void gemm(const Phase& ph) {
    const Range& r = ph.r;
    const double flops = ph.flops;
    if (r.count <= 0) return;
    if (ktime_) {
      if (nev_ >= static_cast<int64_t>(evs_.size()) / 2) {
        for (int i = 0; i < 2; ++i) {
          hipEvent_t e;
          HIPCHK(hipEventCreate(&e));
          evs_.push_back(e);
        }
      }
      HIPCHK(hipEventRecord(evs_[2 * nev_], s_));
    }
    if (ph.kind == PhaseKind::Gemm64)
      launch(k_gemm<T, 64>, r.count, 256, d_fd_, d_gemm_ + r.begin, d_arena_, d_dpiv_);
    else
      launch(k_gemm<T, 128>, r.count, 256, d_fd_, d_gemm_ + r.begin, d_arena_, d_dpiv_);
    tm_.gemm_flops += flops;
    if (ktime_) {
      HIPCHK(hipEventRecord(evs_[2 * nev_ + 1], s_));
      ++nev_;
    }
  }

  void ensure_vectors(int nrhs) {
    if (nrhs <= cap_rhs_) return;
    for (void* p : {static_cast<void*>(d_t_), static_cast<void*>(d_x_), static_cast<void*>(d_part_),
                    static_cast<void*>(d_b_), static_cast<void*>(d_r_), static_cast<void*>(d_red_),
                    static_cast<void*>(d_nrm_), static_cast<void*>(d_bc_)})
      if (p) HIPCHK(hipFree(p));
    cap_rhs_ = nrhs;
    d_t_ = dalloc<double>(static_cast<size_t>(P_.tvec_elems) * nrhs, bytes_);
    d_x_ = dalloc<double>(static_cast<size_t>(n_) * nrhs, bytes_);
    d_b_ = dalloc<double>(static_cast<size_t>(n_) * nrhs, bytes_);
    d_r_ = dalloc<double>(static_cast<size_t>(n_) * nrhs, bytes_);
    d_part_ = dalloc<double>(static_cast<size_t>(std::max<int64_t>(P_.max_partials, 1)) * nrhs * NB, bytes_);
    d_red_ = dalloc<double>(static_cast<size_t>(kRedBlocks) * nrhs, bytes_);
    d_nrm_ = dalloc<double>(nrhs, bytes_);
    d_bc_ = dalloc<double>(static_cast<size_t>(std::max<int64_t>(nc_, 1)) * nrhs, bytes_);
  }

  // x = A^-1 b with FP64 refinement; b in d_b_ (internal numbering), result in d_x_
  void refine(int nrhs, double tol, int max_steps) {
    const int64_t cnt = n_ * nrhs;
    HIPCHK(hipMemcpyAsync(d_x_, d_b_, cnt * sizeof(double), hipMemcpyDeviceToDevice, s_));
    solve_device(d_x_, nrhs);
    std::vector<double> bn(nrhs), rn(nrhs);
    sumsq(d_b_, nrhs, bn.data());
    int steps = 0;
    double worst = 0.0;
    for (;;) {
      launch(k_residual, (n_ * 16 + 255) / 256, 256, d_rp_, d_ci_, d_av_, d_x_, d_b_, d_r_, n_, nrhs);
      sumsq(d_r_, nrhs, rn.data());
      worst = 0.0;
      bool done = true;
      for (int j = 0; j < nrhs; ++j) {
        const double rel = bn[j] > 0 ? std::sqrt(rn[j] / bn[j]) : 0.0;
        worst = std::max(worst, rel);
        if (!(rel <= tol)) done = false;
      }
      if (done || steps >= max_steps) break;
      solve_device(d_r_, nrhs);
      launch(k_axpy, (cnt + 255) / 256, 256, d_x_, d_r_, cnt);
      ++steps;
    }
    tm_.refine_steps = steps;
    tm_.rel_residual = worst;
  }

  void sumsq(const double* v, int nrhs, double* host_out) {
    launch(k_sumsq_partial, kRedBlocks, 256, v, n_, nrhs, d_red_);
    launch(k_sumsq_final, 1, 64, d_red_, kRedBlocks, nrhs, d_nrm_);
    HIPCHK(hipMemcpyAsync(host_out, d_nrm_, nrhs * sizeof(double), hipMemcpyDeviceToHost, s_));
    HIPCHK(hipStreamSynchronize(s_));
  }

  void solve_device(double* dx, int nrhs) {
    const int64_t tv = P_.tvec_elems;
    for (int64_t L = 0; L < static_cast<int64_t>(P_.levels.size()); ++L) {
      const LevelPlan& lp = P_.levels[L];
      const int32_t* fr = d_lf_ + lp.fronts_r.begin;
      launch(k_fwd_init, lp.fronts_r.count, 256, d_fd_, fr, dx, d_t_, n_, tv, nrhs);
      for (int sl = 0; sl < 2; ++sl) {
        const Range& cr = P_.child_slot_range[sl][L];
        launch(k_fwd_child, cr.count, 256, d_fd_, d_cs_[sl] + cr.begin, d_cmap_, d_t_, tv, nrhs);
      }
      for (const SolveStep& sp : lp.solve) {
        launch(k_fwd_diag<T>, sp.diag.count, 256, d_fd_, d_diag_ + sp.diag.begin, d_arena_, d_t_, tv, nrhs);
        launch(k_fwd_update<T>, sp.rows.count, 256, d_fd_, d_srow_ + sp.rows.begin, d_arena_, d_t_, tv, nrhs);
      }
      launch(k_fwd_finish, lp.fronts_r.count, 256, d_fd_, fr, d_t_, d_dpiv_, dx, n_, tv, nrhs);
    }
    for (int64_t L = static_cast<int64_t>(P_.levels.size()) - 1; L >= 0; --L) {
      const LevelPlan& lp = P_.levels[L];
      const int32_t* fr = d_lf_ + lp.fronts_r.begin;
      launch(k_bwd_init, lp.fronts_r.count, 256, d_fd_, fr, d_cmap_, dx, d_t_, n_, tv, nrhs);
      for (int64_t s = static_cast<int64_t>(lp.solve.size()) - 1; s >= 0; --s) {
        const SolveStep& sp = lp.solve[s];
        launch(k_bwd_partial<T>, sp.rows.count, 256, d_fd_, d_srow_ + sp.rows.begin, d_arena_, d_t_, tv, nrhs,
               d_part_);
        launch(k_bwd_diag<T>, sp.diag.count, 256, d_fd_, d_diag_ + sp.diag.begin, d_arena_, d_t_, tv, nrhs,
               d_part_, static_cast<int64_t>(0));
      }
      launch(k_bwd_finish, lp.fronts_r.count, 256, d_fd_, fr, d_t_, dx, n_, tv, nrhs);
    }
  }

  static constexpr int kRedBlocks = 1024;
  Plan P_;
  hipStream_t s_ = nullptr;
  hipEvent_t e0_ = nullptr, e1_ = nullptr;
  int64_t n_ = 0, nlower_ = 0, nfull_ = 0;
  size_t bytes_ = 0;
  Timings tm_;
  bool ktime_ = false;
  std::vector<hipEvent_t> evs_;
  int64_t nev_ = 0;
  int cap_rhs_ = 0;
  FrontDesc* d_fd_ = nullptr;
  int32_t* d_cmap_ = nullptr;
  int64_t* d_amap_ = nullptr;
  ZeroTile* d_zero_ = nullptr;
  ScatterTile* d_scatter_ = nullptr;
  EaTile* d_ea_ = nullptr;
  GemmTile* d_gemm_ = nullptr;
  DiagTask* d_diag_ = nullptr;
  TrsmTile* d_trsm_ = nullptr;
  TrsmTile* d_srow_ = nullptr;
  int32_t* d_lf_ = nullptr;
  int32_t* d_cs_[2] = {nullptr, nullptr};
  T* d_arena_ = nullptr;
  double *d_dpiv_ = nullptr, *d_adiag_ = nullptr, *d_lval_ = nullptr, *d_l11_ = nullptr;
  unsigned long long* d_err_ = nullptr;
  int64_t* d_dpos_ = nullptr;
  int64_t* d_rp_ = nullptr;
  int32_t* d_ci_ = nullptr;
  int64_t* d_gmap_ = nullptr;
  double* d_av_ = nullptr;
  double *d_t_ = nullptr, *d_x_ = nullptr, *d_part_ = nullptr, *d_b_ = nullptr, *d_r_ = nullptr, *d_red_ = nullptr,
         *d_nrm_ = nullptr;
  double* d_bc_ = nullptr;
  int64_t* d_eq_ = nullptr;
  double* d_iso_ = nullptr;
  int64_t nc_ = 0;
};

}  // namespace

std::unique_ptr<GpuFactor> make_gpu_factor(Plan plan, const std::vector<int64_t>& colptr,
                                           const std::vector<int64_t>& rowidx, Precision prec) {
  if (prec == Precision::MIXED) return std::make_unique<Impl<float>>(std::move(plan), colptr, rowidx);
  return std::make_unique<Impl<double>>(std::move(plan), colptr, rowidx);
}

}  // namespace rf
