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
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <type_traits>

#include "gemm_kernel.hpp"
#include "panel_kernels.hpp"
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

// ---------------------------------------------------------------- assembly

// Clears the lower triangle of every front (column c: rows c..m-1, one
// contiguous segment of the column-major front) once per factorization, one
// wave per column, 16-byte stores between the unaligned ends.  The arena is
// not reused across levels, so nothing written later is cleared by it.
template <typename T>
__global__ __launch_bounds__(256) void k_zero(const FrontDesc* __restrict__ fd, const ZeroTile* __restrict__ tiles,
                                              T* __restrict__ arena) {
  constexpr int V = 16 / static_cast<int>(sizeof(T));
  const ZeroTile t = tiles[blockIdx.x];
  const FrontDesc d = fd[t.front];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cend = min(t.c0 + ZC, d.m);
  for (int c = t.c0 + wave; c < cend; c += 4) {
    const int64_t base = d.off + static_cast<int64_t>(c) * d.m;
    const int64_t s = base + c, e = base + d.m;
    const int64_t a = min(e, (s + V - 1) / V * V);  // first 16-byte aligned element
    const int64_t b = max(a, e / V * V);             // end of the aligned body
    if (lane < a - s) arena[s + lane] = T(0);
    if (lane < e - b) arena[b + lane] = T(0);
    float4* p = reinterpret_cast<float4*>(arena + a);
    const int64_t nv = (b - a) / V;
    for (int64_t i = lane; i < nv; i += 64) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// the original entries of the level's pivot columns, added to what child slot
// 0 wrote into the cleared fronts (each entry once per launch)
template <typename T>
__global__ __launch_bounds__(256) void k_scatter(const ScatterTile* __restrict__ tiles, const int64_t* __restrict__ amap,
                                                 const double* __restrict__ val, T* __restrict__ arena) {
  const ScatterTile t = tiles[blockIdx.x];
  for (int64_t p = t.p0 + threadIdx.x; p < t.p1; p += 256) arena[amap[p]] += T(val[p]);
}

// F_parent[map[x], map[y]] += U_child[x, y] for y <= x in one tile of TX rows
// x and 4096 / TX columns y (lane = row, so that a wave reads and writes
// column segments of TX rows; TX = 256 for large blocks).  ADD = false (child
// slot 0, into the cleared parent before its original entries are added):
// F_parent[map[x], map[y]] = U_child[x, y], without reading the parent.
template <typename T, int TX, bool ADD>
__device__ __forceinline__ void extend_add_tile(const EaTile& t, const FrontDesc& c, const FrontDesc& p,
                                                const int32_t* __restrict__ cm, T* __restrict__ arena,
                                                int32_t* mx, int32_t* my) {
  constexpr int TY = EA_T * EA_T / TX, NY = EA_T * EA_T / 256, YS = 256 / TX;
  static_assert(TX <= 256 && TY <= 256 && NY * YS == TY, "one row per lane, 16 entries per thread");
  const int tid = threadIdx.x;
  if (tid < TX) {
    const int x = t.x0 + tid;
    mx[tid] = x < c.u ? cm[x] : 0;
  }
  if (tid < TY) {
    const int y = t.y0 + tid;
    my[tid] = y < c.u ? cm[y] : 0;
  }
  __syncthreads();
  // all reads of the thread's 16 entries are issued before its writes, from
  // clamped addresses inside both fronts (0 <= y0 <= yc <= u - 1 and
  // xc <= u - 1); only entries with y <= x < u are written, each by one lane
  const int lx = tid % TX, ly0 = tid / TX;
  const int x = t.x0 + lx;
  const int xc = min(x, c.u - 1);
  const T* src = arena + c.off + (c.k + xc) + static_cast<int64_t>(c.k) * c.m;
  T* dst = arena + p.off + mx[lx];
  T sv[NY], dv[NY];
#pragma unroll
  for (int k = 0; k < NY; ++k) {
    const int ly = ly0 + k * YS;
    const int yc = max(min(t.y0 + ly, xc), t.y0);
    sv[k] = src[static_cast<int64_t>(yc) * c.m];
    if constexpr (ADD) dv[k] = dst[static_cast<int64_t>(my[yc - t.y0]) * p.m];
  }
  if (x >= c.u) return;
#pragma unroll
  for (int k = 0; k < NY; ++k) {
    const int ly = ly0 + k * YS;
    if (t.y0 + ly <= x) dst[static_cast<int64_t>(my[ly]) * p.m] = ADD ? dv[k] + sv[k] : sv[k];
  }
}

template <typename T, bool ADD>
__global__ __launch_bounds__(256) void k_extend_add(const FrontDesc* __restrict__ fd, const EaTile* __restrict__ tiles,
                                                    const int32_t* __restrict__ cmap, T* __restrict__ arena) {
  const EaTile t = tiles[blockIdx.x];
  const FrontDesc c = fd[t.child];
  const FrontDesc p = fd[c.parent];
  __shared__ int32_t mx[EA_TX], my[EA_T];
  const int32_t* cm = cmap + c.cmap_off;
  if (t.tall)
    extend_add_tile<T, EA_TX, ADD>(t, c, p, cm, arena, mx, my);
  else
    extend_add_tile<T, EA_T, ADD>(t, c, p, cm, arena, mx, my);
}

// ------------------------------------------------------------ pivot blocks

// LDL^T of the jb x jb diagonal block in FP64 with the pivot test of
// src/sparse.rs:528-531 (the failing equation with the smallest index is
// kept), then the block's inverse L^-1 in FP64 (column-major NB x NB in the
// persistent V buffer at the task's index, zero outside the block).
template <typename T>
__global__ __launch_bounds__(256) void k_diag(const FrontDesc* __restrict__ fd, const DiagTask* __restrict__ tasks,
                                              int64_t task0, T* __restrict__ arena, double* __restrict__ dpiv,
                                              const double* __restrict__ adiag, double* __restrict__ vbuf,
                                              unsigned long long* __restrict__ err) {
  const int64_t task = task0 + blockIdx.x;
  const DiagTask t = tasks[task];
  const FrontDesc d = fd[t.front];
  const int jb = t.jb;
  __shared__ double A[NB][NB + 1];  // A[col][row]
  __shared__ double V[NB][NB + 1];  // V[col][row] = L^-1
  __shared__ double ao[NB];         // original diagonal of the block's equations
  T* base = arena + d.off + t.j0 + static_cast<int64_t>(t.j0) * d.m;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int i = idx % NB, j = idx / NB;
    A[j][i] = (i < jb && j < jb && i >= j) ? static_cast<double>(base[i + static_cast<int64_t>(j) * d.m]) : 0.0;
  }
  if (tid < NB) ao[tid] = tid < jb ? adiag[static_cast<int64_t>(d.start) + t.j0 + tid] : 1.0;
  __syncthreads();
  // right-looking elimination with the unscaled column j: A[q][i] -= A[j][i] A[j][q] / d_j
  for (int j = 0; j < jb; ++j) {
    const double dj = A[j][j];
    if (tid == 0 && (!(ao[j] > 0.0) || !(dj > 1e-12 * ao[j])))
      atomicMin(err, static_cast<unsigned long long>(static_cast<int64_t>(d.start) + t.j0 + j));
    const int rem = jb - j - 1;
    for (int idx = tid; idx < rem * rem; idx += 256) {
      const int i = j + 1 + idx % rem, q = j + 1 + idx / rem;
      if (q <= i) A[q][i] -= A[j][i] * A[j][q] / dj;
    }
    __syncthreads();
  }
  // L = column-scaled A
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int i = idx % NB, j = idx / NB;
    if (i > j && i < jb) A[j][i] = A[j][i] / A[j][j];
  }
  __syncthreads();
  // L^-1 by forward substitution, four lanes per column (q = tid / 4), no block barriers
  {
    const int q = tid >> 2, g = tid & 3;
    for (int i = g; i < NB; i += 4) V[q][i] = (i == q && q < jb) ? 1.0 : 0.0;
    __builtin_amdgcn_wave_barrier();
    for (int i = 1; i < jb; ++i) {
      double part = 0.0;
      if (i > q)
        for (int p = q + g; p < i; p += 4) part = fma(A[p][i], V[q][p], part);
      part += __shfl_xor(part, 1, 64);
      part += __shfl_xor(part, 2, 64);
      if (g == 0 && i > q && q < jb) V[q][i] = -part;
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
  double* vb = vbuf + task * NB * NB;
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int i = idx % NB, q = idx / NB;
    vb[idx] = V[q][i];  // L^-1[i][q], column-major, zero outside the block
    if (i < jb && q < jb && i > q) base[i + static_cast<int64_t>(q) * d.m] = T(A[q][i]);
  }
  if (tid < jb) dpiv[d.start + t.j0 + tid] = A[tid][tid];
}

// L[r, B] = F[r, B] W with W = L_BB^-T D_B^-1, for the TRS rows of one tile
// (FP64 MFMA, in place)
template <typename T>
__global__ __launch_bounds__(256) void k_trsm(const FrontDesc* __restrict__ fd, const TrsmTile* __restrict__ tiles,
                                              T* __restrict__ arena, const double* __restrict__ vbuf,
                                              const double* __restrict__ dpiv) {
  using M = Mfma<double>;
  const TrsmTile t = tiles[blockIdx.x];
  const FrontDesc d = fd[t.front];
  __shared__ double Ws[NB][NB + 2];   // Ws[p][q]
  __shared__ double As[NB][TRS + 2];  // As[p][r]
  const double* vb = vbuf + static_cast<int64_t>(t.task) * NB * NB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  T* F = arena + d.off;
  const int64_t m = d.m;
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int q = idx % NB, p = idx / NB;  // vb[q + p NB] = L^-1[q][p]
    Ws[p][q] = q < t.jb ? vb[idx] / dpiv[d.start + t.j0 + q] : 0.0;
  }
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

// One inner pivot step of one front for one tile of rows (PanelTile):
//  1. left-looking update from the outer block's columns [J0, j0) of the
//     diagonal block (C_BB) and of the tile's rows (C_R), FP64 MFMA;
//  2. LDL^T of the 64 x 64 diagonal block and its inverse L^-1
//     (block_ldlt_blk, blocked by 16 columns), with the pivot test of
//     src/sparse.rs:528-531;
//  3. the diagonal tile writes D and L^-1 (the arena's pivot block is left as
//     assembled, because the other tiles of the launch read it); every other
//     tile applies L[r, B] = A_R L_BB^-T D^-1 to its rows (FP64 MFMA).
// Every tile factors the (identical) diagonal block itself, so one launch
// covers the whole inner step.
template <typename T>
__global__ __launch_bounds__(256) void k_panel(const FrontDesc* __restrict__ fd, const PanelTile* __restrict__ tiles,
                                               T* __restrict__ arena, double* __restrict__ dpiv,
                                               const double* __restrict__ adiag, double* __restrict__ vbuf,
                                               unsigned long long* __restrict__ err) {
  using M = Mfma<double>;
  constexpr int KS = 32;
  const PanelTile t = tiles[blockIdx.x];
  const FrontDesc d = fd[t.front];
  const int jb = t.jb, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool diag_tile = t.r0 == t.j0;
  T* F = arena + d.off;
  const int64_t m = d.m;
  __shared__ double Ad[NB][NB + 1];   // diagonal block [col][row]
  __shared__ double Ar[NB][NB + 1];   // the tile's rows [col][row]
  __shared__ __attribute__((aligned(16))) double Vs[NB][NB + 2];        // L^-1 [col][row]
  __shared__ __attribute__((aligned(16))) double YY[2][KS][NB + 2];     // Ys, Y0; later scratch
  __shared__ double Xs[KS][NB + 2];   // L[r0 + i, k]
  __shared__ double cbuf[NB], dd[NB];
  double(*Ys)[NB + 2] = YY[0];        // L[j0 + q, k] d_k
  double(*Y0)[NB + 2] = YY[1];        // L[j0 + i, k]
  static_assert(2 * KS * (NB + 2) >= 4 * 256, "the inverse's scratch fits in the update buffers");
  M::acc_t cb[4], cr[4];
#pragma unroll
  for (int bj = 0; bj < 4; ++bj) {
    cb[bj] = M::acc_t{0, 0, 0, 0};
    cr[bj] = M::acc_t{0, 0, 0, 0};
  }
  for (int k0 = t.J0; k0 < t.j0; k0 += KS) {
    __syncthreads();
    for (int e = tid; e < KS * NB; e += 256) {
      const int i = e % NB, kk = e / NB;
      const int k = k0 + kk;
      const bool kin = k < t.j0;
      const double y = (kin && i < jb) ? static_cast<double>(F[t.j0 + i + k * m]) : 0.0;
      Y0[kk][i] = y;
      Ys[kk][i] = kin ? y * dpiv[d.start + k] : 0.0;
      Xs[kk][i] = (kin && !diag_tile && t.r0 + i < d.m) ? static_cast<double>(F[t.r0 + i + k * m]) : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KS; kk += 4) {
      const int kr = kk + lane / 16;
      const double b0 = Y0[kr][wave * 16 + lane % 16];
      const double bx = Xs[kr][wave * 16 + lane % 16];
#pragma unroll
      for (int bj = 0; bj < 4; ++bj) {
        const double a = Ys[kr][bj * 16 + lane % 16];
        cb[bj] = M::op(a, b0, cb[bj]);
        cr[bj] = M::op(a, bx, cr[bj]);
      }
    }
  }
  // A_BB = F_BB - C_BB (both triangles from the stored lower triangle); A_R = F_R - C_R
  __syncthreads();
#pragma unroll
  for (int bj = 0; bj < 4; ++bj)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = bj * 16 + M::orow(lane, r);
      const int i = wave * 16 + lane % 16;
      double fb = 0.0;
      if (i < jb && q < jb) {
        const int hi = i > q ? i : q, lo = i > q ? q : i;
        fb = static_cast<double>(F[t.j0 + hi + static_cast<int64_t>(t.j0 + lo) * m]) - cb[bj][r];
      }
      Ad[q][i] = fb;
      if (!diag_tile) {
        const int row = t.r0 + i;
        Ar[q][i] = (row < d.m && q < jb) ? static_cast<double>(F[row + static_cast<int64_t>(t.j0 + q) * m]) - cr[bj][r]
                                          : 0.0;
      }
    }
  __syncthreads();
  {
    const double orig = lane < jb ? adiag[static_cast<int64_t>(d.start) + t.j0 + lane] : 1.0;
    // the update buffers are free now: the inverse's scratch
    block_ldlt_blk(Ad, Vs, &YY[0][0][0], cbuf, dd, jb, orig, static_cast<int64_t>(d.start) + t.j0,
                   diag_tile ? err : nullptr);
  }
  __syncthreads();
  if (diag_tile) {
    // The pivot block itself stays as assembled in the arena: the other tiles
    // of this launch read it, and everything after the factorization uses the
    // block's inverse (V) and D instead of L_BB.
    double* vb = vbuf + static_cast<int64_t>(t.task) * NB * NB;
    for (int idx = tid; idx < NB * NB; idx += 256) vb[idx] = Vs[idx / NB][idx % NB];
    if (tid < jb) dpiv[d.start + t.j0 + tid] = dd[tid];
    return;
  }
  // W[p][q] = L^-1[q][p] / d_q, kept in Ad's storage as Wd[p][q]
  double (*Wd)[NB + 1] = Ad;
  __shared__ double dinv[NB];
  if (tid < NB) dinv[tid] = tid < jb ? 1.0 / dd[tid] : 0.0;
  __syncthreads();
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int q = idx % NB, p = idx / NB;
    Wd[p][q] = q < jb ? Vs[p][q] : 0.0;  // Vs[p][q] = L^-1[q][p]
  }
  __syncthreads();
  M::acc_t o[4];
#pragma unroll
  for (int bj = 0; bj < 4; ++bj) o[bj] = M::acc_t{0, 0, 0, 0};
#pragma unroll 4
  for (int kk = 0; kk < NB; kk += 4) {
    const int kr = kk + lane / 16;
    const double b = Ar[kr][wave * 16 + lane % 16];
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) o[bj] = M::op(Wd[kr][bj * 16 + lane % 16], b, o[bj]);
  }
  const int row = t.r0 + wave * 16 + lane % 16;
  if (row < d.m) {
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = bj * 16 + M::orow(lane, r);
        if (q < jb) F[row + static_cast<int64_t>(t.j0 + q) * m] = T(o[bj][r] * dinv[q]);
      }
  }
}

// ------------------------------------------------------------------ solves
// Vectors t (per front, m entries, column per right-hand side) and x
// (internal numbering, column per right-hand side) are FP64.  Every pivot
// block's triangular solve is a product with its FP64 inverse V = L_BB^-1.

constexpr int RB = 16;  // right-hand sides per pass in the solve kernels
// Entries per workgroup of the per-front vector kernels (k_fwd_init,
// k_fwd_child, k_*_finish, k_bwd_init): their grids are (front, right-hand
// side, chunk of VCH entries), so the large fronts of the upper levels and
// many right-hand sides spread over the whole GPU.
constexpr int VCH = 2048;

__global__ __launch_bounds__(256) void k_fwd_init(const FrontDesc* __restrict__ fd, const int32_t* __restrict__ fronts,
                                                  const double* __restrict__ x, double* __restrict__ t, int64_t n,
                                                  int64_t tv, int nrhs) {
  const FrontDesc d = fd[fronts[blockIdx.x]];
  const int j = blockIdx.y, i0 = blockIdx.z * VCH, i1 = min(d.m, i0 + VCH);
  for (int i = i0 + threadIdx.x; i < i1; i += 256) t[j * tv + d.toff + i] = i < d.k ? x[j * n + d.start + i] : 0.0;
}

// t_parent[cmap[x]] += t_child[k_child + x] (forward: the child's update vector)
__global__ __launch_bounds__(256) void k_fwd_child(const FrontDesc* __restrict__ fd, const int32_t* __restrict__ children,
                                                   const int32_t* __restrict__ cmap, double* __restrict__ t, int64_t tv,
                                                   int nrhs) {
  const FrontDesc c = fd[children[blockIdx.x]];
  const FrontDesc p = fd[c.parent];
  const int32_t* cm = cmap + c.cmap_off;
  const int j = blockIdx.y, x0 = blockIdx.z * VCH, x1 = min(c.u, x0 + VCH);
  for (int x = x0 + threadIdx.x; x < x1; x += 256) t[j * tv + p.toff + cm[x]] += t[j * tv + c.toff + c.k + x];
}

// Vs <- V of one pivot block (global -> LDS)
__device__ inline void load_v(const double* __restrict__ vb, double (*Vs)[NB + 1]) {
  for (int idx = threadIdx.x; idx < NB * NB; idx += 256) Vs[idx / NB][idx % NB] = vb[idx];  // Vs[q][i] = L^-1[i][q]
}

// t[B] <- V t[B] for one pivot block, R right-hand sides per pass
template <int R>
__device__ void block_lower_solve(const double (*Vs)[NB + 1], double* __restrict__ tb, int64_t tv, int jb, int nrhs,
                                  double (*tmp)[R + 1]) {
  const int tid = threadIdx.x;
  for (int j0 = 0; j0 < nrhs; j0 += R) {
    const int nr = min(R, nrhs - j0);
    __syncthreads();
    for (int idx = tid; idx < NB * R; idx += 256) {
      const int i = idx % NB, j = idx / NB;
      tmp[i][j] = (i < jb && j < nr) ? tb[(j0 + j) * tv + i] : 0.0;
    }
    __syncthreads();
    for (int idx = tid; idx < NB * R; idx += 256) {
      const int i = idx % NB, j = idx / NB;
      if (i < jb && j < nr) {
        double s = 0.0;
        for (int q = 0; q <= i; ++q) s = fma(Vs[q][i], tmp[q][j], s);
        tb[(j0 + j) * tv + i] = s;
      }
    }
  }
}

// first pivot block of every front of a level
template <int R>
__global__ __launch_bounds__(256) void k_fwd_first(const FrontDesc* __restrict__ fd, const DiagTask* __restrict__ tasks,
                                                   int64_t task0, const double* __restrict__ vbuf,
                                                   double* __restrict__ t, int64_t tv, int nrhs) {
  const int64_t task = task0 + blockIdx.x;
  const DiagTask dt = tasks[task];
  const FrontDesc d = fd[dt.front];
  __shared__ double Vs[NB][NB + 1];
  __shared__ double tmp[NB][R + 1];
  load_v(vbuf + task * NB * NB, Vs);
  block_lower_solve<R>(Vs, t + d.toff + dt.j0, tv, dt.jb, nrhs, tmp);
}

// t[r] -= L[r, B] y[B] for the SR rows of one tile (4 rows per thread); the
// tile that starts at the next pivot block then solves that block (its rows
// are final at this point).
template <typename T, int R>
__global__ __launch_bounds__(256) void k_fwd_step(const FrontDesc* __restrict__ fd, const SolveTile* __restrict__ tiles,
                                                  const T* __restrict__ arena, const double* __restrict__ vbuf,
                                                  double* __restrict__ t, int64_t tv, int nrhs) {
  const SolveTile tile = tiles[blockIdx.x];
  const FrontDesc d = fd[tile.front];
  __shared__ double y[R][NB];
  __shared__ double Vs[NB][NB + 1];
  __shared__ double tmp[NB][R + 1];
  const int jb = tile.jb, tid = threadIdx.x;
  const bool solve_next = tile.local == 0 && tile.next >= 0;
  if (solve_next) load_v(vbuf + static_cast<int64_t>(tile.next) * NB * NB, Vs);
  const T* colb = arena + d.off + static_cast<int64_t>(tile.j0) * d.m;
  double* tf = t + d.toff;
  const int rend = min(tile.r0 + SR, d.m);
  for (int j0 = 0; j0 < nrhs; j0 += R) {
    const int nr = min(R, nrhs - j0);
    __syncthreads();
    for (int idx = tid; idx < NB * R; idx += 256) {
      const int p = idx % NB, j = idx / NB;
      y[j][p] = (p < jb && j < nr) ? tf[(j0 + j) * tv + tile.j0 + p] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int s4 = 0; s4 < SR / 256; ++s4) {
      const int r = tile.r0 + tid + 256 * s4;
      if (r < rend) {
        double acc[R];
#pragma unroll
        for (int j = 0; j < R; ++j) acc[j] = 0.0;
        const T* row = colb + r;
#pragma unroll 8
        for (int p = 0; p < jb; ++p) {
          const double l = static_cast<double>(row[static_cast<int64_t>(p) * d.m]);
#pragma unroll
          for (int j = 0; j < R; ++j) acc[j] = fma(l, y[j][p], acc[j]);
        }
#pragma unroll
        for (int j = 0; j < R; ++j)
          if (j < nr) tf[(j0 + j) * tv + r] -= acc[j];
      }
    }
  }
  if (solve_next) {
    __syncthreads();
    const int jbn = min(NB, d.k - (tile.j0 + jb));
    block_lower_solve<R>(Vs, tf + tile.j0 + jb, tv, jbn, nrhs, tmp);
  }
}

__global__ __launch_bounds__(256) void k_fwd_finish(const FrontDesc* __restrict__ fd, const int32_t* __restrict__ fronts,
                                                    const double* __restrict__ t, const double* __restrict__ dpiv,
                                                    double* __restrict__ x, int64_t n, int64_t tv, int nrhs) {
  const FrontDesc d = fd[fronts[blockIdx.x]];
  const int j = blockIdx.y, i0 = blockIdx.z * VCH, i1 = min(d.k, i0 + VCH);
  for (int i = i0 + threadIdx.x; i < i1; i += 256) x[j * n + d.start + i] = t[j * tv + d.toff + i] / dpiv[d.start + i];
}

// backward: t[0:k] = x[pivots], t[k + r] = t_parent[cmap[r]] (final values of ancestors)
__global__ __launch_bounds__(256) void k_bwd_init(const FrontDesc* __restrict__ fd, const int32_t* __restrict__ fronts,
                                                  const int32_t* __restrict__ cmap, const double* __restrict__ x,
                                                  double* __restrict__ t, int64_t n, int64_t tv, int nrhs) {
  const FrontDesc d = fd[fronts[blockIdx.x]];
  const int j = blockIdx.y, i0 = blockIdx.z * VCH, i1 = min(d.m, i0 + VCH);
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
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

// partial[tile][j][c] = sum over rows r of the tile of L21[r, c] t[r][j] for
// the tile's BC columns; L is staged through LDS in 32-row slabs loaded with
// lanes along the (contiguous) rows, then each thread sums one column.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_l21_partial(const FrontDesc* __restrict__ fd, const L21Tile* __restrict__ tiles,
                                                     const T* __restrict__ arena, const double* __restrict__ t,
                                                     int64_t tv, int nrhs, double* __restrict__ partial) {
  const L21Tile tile = tiles[blockIdx.x];
  const FrontDesc d = fd[tile.front];
  constexpr int SL = 32;
  __shared__ T Ls[BC][SL + 1];
  __shared__ double ts[SL][R];
  const int tid = threadIdx.x;
  const int rend = min(tile.r0 + SR, d.m);
  const int cend = min(tile.c0 + BC, d.k);
  const T* F = arena + d.off;
  const double* tf = t + d.toff;
  double* out = partial + static_cast<int64_t>(blockIdx.x) * nrhs * BC;
  for (int j0 = 0; j0 < nrhs; j0 += R) {
    const int nr = min(R, nrhs - j0);
    double acc[R];
#pragma unroll
    for (int j = 0; j < R; ++j) acc[j] = 0.0;
    for (int rs = tile.r0; rs < rend; rs += SL) {
      __syncthreads();
#pragma unroll 4
      for (int e = tid; e < SL * BC; e += 256) {
        const int rr = e % SL, cc = e / SL;
        const int r = rs + rr, c = tile.c0 + cc;
        Ls[cc][rr] = (r < rend && c < cend) ? F[r + static_cast<int64_t>(c) * d.m] : T(0);
      }
      for (int e = tid; e < SL * R; e += 256) {
        const int rr = e % SL, j = e / SL;
        ts[rr][j] = (rs + rr < rend && j < nr) ? tf[(j0 + j) * tv + rs + rr] : 0.0;
      }
      __syncthreads();
#pragma unroll 8
      for (int rr = 0; rr < SL; ++rr) {
        const double l = static_cast<double>(Ls[tid][rr]);
#pragma unroll
        for (int j = 0; j < R; ++j) acc[j] = fma(l, ts[rr][j], acc[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (j < nr) out[(j0 + j) * BC + tid] = acc[j];
  }
}

// t[c] -= sum of the L21 partials of column c (tile order) for BC columns of a front
__global__ __launch_bounds__(256) void k_l21_combine(const FrontDesc* __restrict__ fd, const L21Comb* __restrict__ comb,
                                                     const double* __restrict__ partial,
                                                     double* __restrict__ t, int64_t tv, int nrhs) {
  const L21Comb cb = comb[blockIdx.x];
  const FrontDesc d = fd[cb.front];
  const int c = cb.c0 + threadIdx.x;
  if (c >= d.k) return;
  for (int j = 0; j < nrhs; ++j) {
    const double* src = partial + (cb.t0 * static_cast<int64_t>(nrhs) + j) * BC + threadIdx.x;
    const int64_t stride = static_cast<int64_t>(nrhs) * BC;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int q = 0;
    for (; q + 4 <= cb.nt; q += 4) {
      a0 += src[(q + 0) * stride];
      a1 += src[(q + 1) * stride];
      a2 += src[(q + 2) * stride];
      a3 += src[(q + 3) * stride];
    }
    for (; q < cb.nt; ++q) a0 += src[q * stride];
    t[j * tv + d.toff + c] -= (a0 + a1) + (a2 + a3);
  }
}

// t[B] <- V^T t[B] for one pivot block, R right-hand sides per pass
template <int R>
__device__ void block_upper_solve(const double (*Vs)[NB + 1], double* __restrict__ tb, int64_t tv, int jb, int nrhs,
                                  double (*tmp)[R + 1]) {
  const int tid = threadIdx.x;
  for (int j0 = 0; j0 < nrhs; j0 += R) {
    const int nr = min(R, nrhs - j0);
    __syncthreads();
    for (int idx = tid; idx < NB * R; idx += 256) {
      const int i = idx % NB, j = idx / NB;
      tmp[i][j] = (i < jb && j < nr) ? tb[(j0 + j) * tv + i] : 0.0;
    }
    __syncthreads();
    for (int idx = tid; idx < NB * R; idx += 256) {
      const int p = idx % NB, j = idx / NB;
      if (p < jb && j < nr) {
        double s = 0.0;
        for (int i = p; i < jb; ++i) s = fma(Vs[p][i], tmp[i][j], s);
        tb[(j0 + j) * tv + p] = s;
      }
    }
  }
}

// the last pivot block of every front of a level: t[B] <- V^T t[B]
template <int R>
__global__ __launch_bounds__(256) void k_bwd_last(const FrontDesc* __restrict__ fd, const DiagTask* __restrict__ tasks,
                                                  const int32_t* __restrict__ last, const double* __restrict__ vbuf,
                                                  double* __restrict__ t, int64_t tv, int nrhs) {
  const int64_t task = last[blockIdx.x];
  const DiagTask dt = tasks[task];
  const FrontDesc d = fd[dt.front];
  __shared__ double Vs[NB][NB + 1];
  __shared__ double tmp[NB][R + 1];
  load_v(vbuf + task * NB * NB, Vs);
  block_upper_solve<R>(Vs, t + d.toff + dt.j0, tv, dt.jb, nrhs, tmp);
}

// t[c] -= sum_i L[j0 + i, c] z_i for the tile's columns (z = t[B], solved); the
// tile ending at the block then solves the previous block.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_bwd_step(const FrontDesc* __restrict__ fd, const BwdTile* __restrict__ tiles,
                                                  const T* __restrict__ arena, const double* __restrict__ vbuf,
                                                  double* __restrict__ t, int64_t tv, int nrhs) {
  const BwdTile tile = tiles[blockIdx.x];
  const FrontDesc d = fd[tile.front];
  constexpr int SL = 32;
  __shared__ T Ls[BC][SL + 1];
  __shared__ double zs[NB][R];
  __shared__ double Vs[NB][NB + 1];
  __shared__ double tmp[NB][R + 1];
  const int tid = threadIdx.x;
  const T* F = arena + d.off;
  double* tf = t + d.toff;
  const int c = tile.c0 + tid;
  for (int j0 = 0; j0 < nrhs; j0 += R) {
    const int nr = min(R, nrhs - j0);
    __syncthreads();
    for (int e = tid; e < NB * R; e += 256) {
      const int i = e % NB, j = e / NB;
      zs[i][j] = (i < tile.jb && j < nr) ? tf[(j0 + j) * tv + tile.j0 + i] : 0.0;
    }
    double acc[R];
#pragma unroll
    for (int j = 0; j < R; ++j) acc[j] = 0.0;
    for (int is = 0; is < tile.jb; is += SL) {
      __syncthreads();
#pragma unroll 4
      for (int e = tid; e < SL * BC; e += 256) {
        const int rr = e % SL, cc = e / SL;
        const int i = is + rr, cg = tile.c0 + cc;
        Ls[cc][rr] = (i < tile.jb && cg < tile.c1) ? F[tile.j0 + i + static_cast<int64_t>(cg) * d.m] : T(0);
      }
      __syncthreads();
#pragma unroll 8
      for (int rr = 0; rr < SL; ++rr) {
        const double l = static_cast<double>(Ls[tid][rr]);
#pragma unroll
        for (int j = 0; j < R; ++j) acc[j] = fma(l, zs[is + rr][j], acc[j]);
      }
    }
    if (c < tile.c1)
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (j < nr) tf[(j0 + j) * tv + c] -= acc[j];
  }
  if (tile.prev >= 0) {
    __syncthreads();
    load_v(vbuf + static_cast<int64_t>(tile.prev) * NB * NB, Vs);
    block_upper_solve<R>(Vs, tf + tile.c1 - tile.jbp, tv, tile.jbp, nrhs, tmp);
  }
}

__global__ __launch_bounds__(256) void k_bwd_finish(const FrontDesc* __restrict__ fd, const int32_t* __restrict__ fronts,
                                                    const double* __restrict__ t, double* __restrict__ x, int64_t n,
                                                    int64_t tv, int nrhs) {
  const FrontDesc d = fd[fronts[blockIdx.x]];
  const int j = blockIdx.y, i0 = blockIdx.z * VCH, i1 = min(d.k, i0 + VCH);
  for (int i = i0 + threadIdx.x; i < i1; i += 256) x[j * n + d.start + i] = t[j * tv + d.toff + i];
}

// ------------------------------------------------ super-block triangular solves
//
// A front's pivot columns are taken SBW at a time.  Forward: one workgroup
// solves the super-block's 64-wide blocks in sequence from LDS (V_b t_b, then
// the super-block's rows below block b), then one launch over SRB-row tiles
// applies the whole super-block to every row below it.  Backward: partial sums
// of L[rows below, J]^T t over PRT-row tiles, then one workgroup per
// super-block subtracts them in tile order and back-substitutes its blocks
// from the last (V_b^T t_b, then the super-block's earlier columns).

template <typename T, int R>
__global__ __launch_bounds__(256) void k_sfwd_chain(const FrontDesc* __restrict__ fd, const SbTask* __restrict__ tasks,
                                                    const T* __restrict__ arena, const double* __restrict__ vbuf,
                                                    const int32_t* __restrict__ btask, double* __restrict__ t,
                                                    int64_t tv, int nrhs) {
  const SbTask tk = tasks[blockIdx.x];
  const FrontDesc d = fd[tk.front];
  __shared__ double tJ[SBW][R];
  __shared__ double Vb[2][NB][NB + 1];
  __shared__ double tmp[NB][R + 1];
  const int W = tk.J1 - tk.J0, tid = threadIdx.x;
  double* tf = t + d.toff + tk.J0;
  const T* F = arena + d.off;
  for (int j0 = 0; j0 < nrhs; j0 += R) {
    const int nr = min(R, nrhs - j0);
    __syncthreads();
    for (int idx = tid; idx < W * R; idx += 256) {
      const int i = idx % W, j = idx / W;
      tJ[i][j] = j < nr ? tf[(j0 + j) * tv + i] : 0.0;
    }
    load_v(vbuf + static_cast<int64_t>(btask[tk.bt]) * NB * NB, Vb[0]);
    for (int b0 = 0; b0 < W; b0 += NB) {
      const int jb = min(NB, W - b0);
      double (*Vs)[NB + 1] = Vb[(b0 / NB) & 1];
      __syncthreads();
      // the next block's inverse streams into the other buffer meanwhile
      if (b0 + NB < W) load_v(vbuf + static_cast<int64_t>(btask[tk.bt + b0 / NB + 1]) * NB * NB, Vb[((b0 / NB) + 1) & 1]);
      for (int idx = tid; idx < NB * R; idx += 256) {
        const int i = idx % NB, j = idx / NB;
        tmp[i][j] = i < jb ? tJ[b0 + i][j] : 0.0;
      }
      __syncthreads();
      for (int idx = tid; idx < NB * R; idx += 256) {
        const int i = idx % NB, j = idx / NB;
        if (i < jb) {
          double s = 0.0;
          for (int q = 0; q <= i; ++q) s = fma(Vs[q][i], tmp[q][j], s);
          tJ[b0 + i][j] = s;
        }
      }
      __syncthreads();
      for (int r = b0 + jb + tid; r < W; r += 256) {
        double acc[R];
#pragma unroll
        for (int j = 0; j < R; ++j) acc[j] = 0.0;
        const T* row = F + static_cast<int64_t>(tk.J0 + b0) * d.m + tk.J0 + r;
#pragma unroll 8
        for (int p = 0; p < jb; ++p) {
          const double l = static_cast<double>(row[static_cast<int64_t>(p) * d.m]);
#pragma unroll
          for (int j = 0; j < R; ++j) acc[j] = fma(l, tJ[b0 + p][j], acc[j]);
        }
#pragma unroll
        for (int j = 0; j < R; ++j) tJ[r][j] -= acc[j];
      }
    }
    __syncthreads();
    for (int idx = tid; idx < W * R; idx += 256) {
      const int i = idx % W, j = idx / W;
      if (j < nr) tf[(j0 + j) * tv + i] = tJ[i][j];
    }
  }
}

// t[r] -= L[r, J0:J1] y[J0:J1] for SRB rows; the four waves split the columns
// Waves per workgroup of the super-block solve kernels (k_sfwd_rows,
// k_sfwd_inv, k_sbwd_comb, k_sbwd_inv): the columns of a tile are split among
// the waves and every wave issues all loads of its share before its first
// product, so that a launch waits for about one memory latency.  16 waves
// with up to 4 right-hand sides per pass, 8 with 16 (LDS); levels whose
// fronts have at most 128 pivots take 4 (many small tiles).
template <int R>
__host__ __device__ constexpr int sb_waves() {
  return R > 4 ? 8 : 16;
}
// widest super-block a workgroup of NW waves serves (the 4-wave kernels are
// launched only for levels whose fronts have at most 128 pivots)
template <int NW>
__host__ __device__ constexpr int sb_maxw() {
  return NW == 4 ? 128 : SBW;
}

// sum of red[0..NW)[lane][j] in a fixed pairwise order
template <int NW, int ROWS, int RP>
__device__ __forceinline__ double wave_tree_sum(const double (&red)[NW][ROWS][RP], int lane, int j) {
  double v[NW];
#pragma unroll
  for (int q = 0; q < NW; ++q) v[q] = red[q][lane][j];
#pragma unroll
  for (int h = NW / 2; h >= 1; h /= 2)
#pragma unroll
    for (int q = 0; q < h; ++q) v[q] = v[q] + v[q + h];
  return v[0];
}

template <typename T, int R, int NW>
__global__ __launch_bounds__(64 * NW) void k_sfwd_rows(const FrontDesc* __restrict__ fd,
                                                                 const SbRowTile* __restrict__ tiles,
                                                                 const T* __restrict__ arena,
                                                                 const double* __restrict__ ysrc,
                                                                 double* __restrict__ t, int64_t tv, int nrhs) {
  constexpr int CPW = sb_maxw<NW>() / NW;
  const SbRowTile tl = tiles[blockIdx.x];
  const FrontDesc d = fd[tl.front];
  __shared__ double y[SBW][R];
  __shared__ double red[NW][SRB][R + 1];
  const int W = tl.J1 - tl.J0, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r = tl.r0 + lane;
  const bool valid = r < d.m;
  const int cw = (W + NW - 1) / NW, cb = w * cw, ce = min(W, cb + cw);
  double* tf = t + d.toff;
  const double* yf = ysrc + d.toff;
  // the lane's row (clamped into the front) of the wave's columns
  const T* F = arena + d.off + static_cast<int64_t>(tl.J0) * d.m + min(r, d.m - 1);
  // all of the wave's loads of L first (one memory latency, once for every
  // chunk of right-hand sides the workgroup takes), then y
  T lv[CPW];
#pragma unroll
  for (int c = 0; c < CPW; ++c) lv[c] = F[static_cast<int64_t>(min(cb + c, W - 1)) * d.m];
  for (int j0 = static_cast<int>(blockIdx.y) * R; j0 < nrhs; j0 += R * static_cast<int>(gridDim.y)) {
    const int nr = min(R, nrhs - j0);
    __syncthreads();
    for (int idx = tid; idx < W * R; idx += 64 * NW) {
      const int i = idx % W, j = idx / W;
      y[i][j] = j < nr ? yf[(j0 + j) * tv + tl.J0 + i] : 0.0;
    }
    __syncthreads();
    double acc[R];
#pragma unroll
    for (int j = 0; j < R; ++j) acc[j] = 0.0;
#pragma unroll
    for (int c = 0; c < CPW; ++c)
      if (cb + c < ce) {
        const double l = static_cast<double>(lv[c]);
#pragma unroll
        for (int j = 0; j < R; ++j) acc[j] = fma(l, y[cb + c][j], acc[j]);
      }
#pragma unroll
    for (int j = 0; j < R; ++j) red[w][lane][j] = acc[j];
    __syncthreads();
    if (tid < SRB && valid)
      for (int j = 0; j < nr; ++j) tf[(j0 + j) * tv + r] -= wave_tree_sum<NW>(red, lane, j);
  }
}

// 32-bit DPP move of both halves of a double (CTRL: a DPP control code)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// sum over the 16 lanes of a DPP row, returned in every lane of the row:
// quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_ror:4, row_ror:8
__device__ __forceinline__ double row16_sum(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x124>(v);
  v += dpp_d<0x128>(v);
  return v;
}

// partial[tile][j][c] = sum over the tile's rows of L[r, c] t[r][j], c in [c0, c1)
// Wave w owns the tile's columns [64 w, 64 w + 64) in groups of four; in a
// group, lane group cg = lane / 16 takes one column and its 16 lanes read 16
// consecutive rows per load (four loads cover the 64 rows).  Every load of
// the tile is issued before the first product; rows past the tile or the
// front are clamped to a valid row and weighted by t = 0, columns past c1 are
// clamped and not stored (their sums are written as 0).  The 16 partial sums
// of a column are reduced across the DPP row.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_sbwd_part(const FrontDesc* __restrict__ fd, const SbPartTile* __restrict__ tiles,
                                                   const T* __restrict__ arena, const double* __restrict__ t,
                                                   int64_t tv, int nrhs, double* __restrict__ partial) {
  static_assert(BC == 256 && PRT == 64, "four waves of 64 columns, 64 rows per tile");
  const SbPartTile tl = tiles[blockIdx.x];
  const FrontDesc d = fd[tl.front];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int rl = lane & 15, cg = lane >> 4;
  const int rend = min(tl.r0 + PRT, d.m);
  const T* F = arena + d.off;
  const double* tf = t + d.toff;
  double* out = partial + static_cast<int64_t>(blockIdx.x) * nrhs * BC;
  int rowc[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) rowc[s] = min(tl.r0 + rl + 16 * s, rend - 1);
  // the wave's groups of four columns that hold a column of the tile: a wave
  // past the tile's columns ends here, and the groups past them are loaded
  // (clamped, so that all loads still go out together) but not multiplied or
  // stored (the combination reads only the tile's columns)
  const int nq = min(16, max(0, (tl.c1 - tl.c0 - 64 * wave + 3) / 4));
  if (nq == 0) return;
  // the tile of L once, for every chunk of right-hand sides the workgroup takes
  T lv[16][4];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int cc = min(tl.c0 + 64 * wave + 4 * q + cg, tl.c1 - 1);
    const T* col = F + static_cast<int64_t>(cc) * d.m;
#pragma unroll
    for (int s = 0; s < 4; ++s) lv[q][s] = col[rowc[s]];
  }
  for (int j0 = static_cast<int>(blockIdx.y) * R; j0 < nrhs; j0 += R * static_cast<int>(gridDim.y)) {
    const int nr = min(R, nrhs - j0);
    double tr[4][R];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int row = tl.r0 + rl + 16 * s;
        tr[s][j] = (row < rend && j < nr) ? tf[(j0 + j) * tv + row] : 0.0;
      }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      // a uniform predicate, not a break: the loop stays unrolled and lv in
      // registers
      if (q < nq) {
        double acc[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
          acc[j] = 0.0;
#pragma unroll
          for (int s = 0; s < 4; ++s) acc[j] = fma(static_cast<double>(lv[q][s]), tr[s][j], acc[j]);
          acc[j] = row16_sum(acc[j]);
        }
        const int cl = 64 * wave + 4 * q + cg;
        if (rl == 0) {
          const bool in = tl.c0 + cl < tl.c1;
#pragma unroll
          for (int j = 0; j < R; ++j)
            if (j < nr) out[(j0 + j) * BC + cl] = in ? acc[j] : 0.0;
        }
      }
    }
  }
}

template <typename T, int R>
__global__ __launch_bounds__(256) void k_sbwd_chain(const FrontDesc* __restrict__ fd, const SbTask* __restrict__ tasks,
                                                    const T* __restrict__ arena, const double* __restrict__ vbuf,
                                                    const int32_t* __restrict__ btask, const double* __restrict__ partial,
                                                    double* __restrict__ t, int64_t tv, int nrhs) {
  const SbTask tk = tasks[blockIdx.x];
  const FrontDesc d = fd[tk.front];
  __shared__ double tJ[SBW][R];
  __shared__ double Vb[2][NB][NB + 1];
  __shared__ double tmp[NB][R + 1];
  const int W = tk.J1 - tk.J0, tid = threadIdx.x;
  double* tf = t + d.toff + tk.J0;
  const T* F = arena + d.off;
  const int last = ((W - 1) / NB) * NB;
  for (int j0 = 0; j0 < nrhs; j0 += R) {
    const int nr = min(R, nrhs - j0);
    __syncthreads();
    load_v(vbuf + static_cast<int64_t>(btask[tk.bt + last / NB]) * NB * NB, Vb[(last / NB) & 1]);
    for (int idx = tid; idx < W * R; idx += 256) {
      const int i = idx % W, j = idx / W;
      double v = 0.0;
      if (j < nr) {
        v = tf[(j0 + j) * tv + i];
        const int cc = i / BC, cl = i % BC;
        const double* src = partial + ((static_cast<int64_t>(tk.p0) + static_cast<int64_t>(cc) * tk.nrc) * nrhs + j0 + j) * BC + cl;
        double s = 0.0;
        for (int rc = 0; rc < tk.nrc; ++rc) s += src[static_cast<int64_t>(rc) * nrhs * BC];
        v -= s;
      }
      tJ[i][j] = v;
    }
    for (int b0 = last; b0 >= 0; b0 -= NB) {
      const int jb = min(NB, W - b0);
      double (*Vs)[NB + 1] = Vb[(b0 / NB) & 1];
      __syncthreads();
      if (b0 > 0) load_v(vbuf + static_cast<int64_t>(btask[tk.bt + b0 / NB - 1]) * NB * NB, Vb[((b0 / NB) - 1) & 1]);
      for (int idx = tid; idx < NB * R; idx += 256) {
        const int i = idx % NB, j = idx / NB;
        tmp[i][j] = i < jb ? tJ[b0 + i][j] : 0.0;
      }
      __syncthreads();
      for (int idx = tid; idx < NB * R; idx += 256) {
        const int p = idx % NB, j = idx / NB;
        if (p < jb) {
          double s = 0.0;
          for (int i = p; i < jb; ++i) s = fma(Vs[p][i], tmp[i][j], s);
          tJ[b0 + p][j] = s;
        }
      }
      __syncthreads();
      for (int c = tid; c < b0; c += 256) {
        double acc[R];
#pragma unroll
        for (int j = 0; j < R; ++j) acc[j] = 0.0;
        const T* col = F + static_cast<int64_t>(tk.J0 + c) * d.m + tk.J0 + b0;
#pragma unroll 8
        for (int i = 0; i < jb; ++i) {
          const double l = static_cast<double>(col[i]);
#pragma unroll
          for (int j = 0; j < R; ++j) acc[j] = fma(l, tJ[b0 + i][j], acc[j]);
        }
#pragma unroll
        for (int j = 0; j < R; ++j) tJ[c][j] -= acc[j];
      }
    }
    __syncthreads();
    for (int idx = tid; idx < W * R; idx += 256) {
      const int i = idx % W, j = idx / W;
      if (j < nr) tf[(j0 + j) * tv + i] = tJ[i][j];
    }
  }
}

// ---------------------------- multi-right-hand-side super-block products (MFMA)
//
// With more than four right-hand sides the two products of the super-block
// solves that read the factor (the forward row update and the backward
// partial sums) run on FP64 MFMA (v_mfma_f64_16x16x4f64) with 16 right-hand
// sides as the N dimension: a tile of the factor is read once per 16
// right-hand sides and multiplied by the matrix cores instead of once per 4
// with VALU FMAs and lane reductions.  The sums are formed in another order
// than the VALU kernels' (to rounding; both are FP64).  Used at every level
// but those whose fronts fit the fused small-front kernels.  (The products
// with the super-block inverses stay on the VALU kernels, which hold a
// wave's share of an inverse in registers: measured faster than this
// chunked MFMA product, 1.45 against 1.92 ms per direction at 16 RHS on the
// 64^3 block.)
constexpr int MR = 16;  // right-hand sides per MFMA pass
// two doubles at an 8-byte aligned address (one 16-byte access)
typedef double v2du __attribute__((ext_vector_type(2), aligned(8)));

// acc += A[rows of the wave, c0:c1) B[c0:c1, j0:j0+16) for one workgroup of
// four waves, wave w owning 16 rows (the MFMA's M): arow points at the lane's
// row of A in column 0 (column stride lda), B[j * ldb + c] is column c of
// right-hand side j.  The columns are the K in steps of 4 (lane group g takes
// column 4 s + g), staged through LDS by chunks of YC columns; the next
// chunk's loads are in flight during a chunk's products.  Step s goes to
// accumulator s % 4 (independent MFMA chains).  c0 and c1 are uniform over
// the workgroup.
constexpr int YC = 32;
template <typename TA>
__device__ __forceinline__ void rows_times_rhs(const TA* __restrict__ arow, int64_t lda, int c0, int c1,
                                               const double* __restrict__ B, int64_t ldb, int j0, int nr,
                                               double (*y)[MR], Mfma<double>::acc_t (&acc)[4]) {
  using M = Mfma<double>;
  constexpr int NST = YC / 4, YPT = YC * MR / 256;
  const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, i = lane & 15;
  TA a[NST];
  double yv[YPT];
  auto load = [&](int cb) {
    const int cn = min(YC, c1 - cb);
#pragma unroll
    for (int st = 0; st < NST; ++st) a[st] = arow[static_cast<int64_t>(cb + min(4 * st + g, cn - 1)) * lda];
#pragma unroll
    for (int q = 0; q < YPT; ++q) {
      const int idx = tid + 256 * q, c = idx % YC, j = idx / YC;
      yv[q] = (c < cn && j < nr) ? B[static_cast<int64_t>(j0 + j) * ldb + cb + c] : 0.0;
    }
  };
  if (c0 < c1) load(c0);
  for (int cb = c0; cb < c1; cb += YC) {
    const int ns = (min(YC, c1 - cb) + 3) / 4;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < YPT; ++q) {
      const int idx = tid + 256 * q;
      y[idx % YC][idx / YC] = yv[q];
    }
    __syncthreads();
    double ac[NST];
#pragma unroll
    for (int st = 0; st < NST; ++st) ac[st] = static_cast<double>(a[st]);
    if (cb + YC < c1) load(cb + YC);
    // columns past the chunk meet y = 0 (their A is a clamped, finite entry)
#pragma unroll
    for (int st = 0; st < NST; ++st)
      if (st < ns) acc[st & 3] = M::op(ac[st], y[4 * st + g][i], acc[st & 3]);
  }
}

// forward: t[r][j] -= sum_c L[r, J0 + c] y[c][j] for the tile's 64 rows
template <typename T>
__global__ __launch_bounds__(256) void k_sfwd_rows_m(const FrontDesc* __restrict__ fd, const SbRowTile* __restrict__ tiles,
                                                     const T* __restrict__ arena, const double* __restrict__ ysrc,
                                                     double* __restrict__ t, int64_t tv, int nrhs) {
  using M = Mfma<double>;
  static_assert(SRB == 64, "four waves of 16 rows");
  __shared__ double y[YC][MR];
  const SbRowTile tl = tiles[blockIdx.x];
  const FrontDesc d = fd[tl.front];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane >> 4, i = lane & 15;
  const int64_t m = d.m;
  const T* arow = arena + d.off + static_cast<int64_t>(tl.J0) * m + min(tl.r0 + 16 * w + i, d.m - 1);
  double* tf = t + d.toff;
  for (int j0 = static_cast<int>(blockIdx.y) * MR; j0 < nrhs; j0 += MR * static_cast<int>(gridDim.y)) {
    const int nr = min(MR, nrhs - j0);
    M::acc_t acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = M::acc_t{0, 0, 0, 0};
    rows_times_rhs(arow, m, 0, tl.J1 - tl.J0, ysrc + d.toff + tl.J0, tv, j0, nr, y, acc);
    // register r of lane l: row 16 w + l/16 + 4 r, right-hand side j0 + l%16
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = tl.r0 + 16 * w + g + 4 * r;
      const double v = (acc[0][r] + acc[1][r]) + (acc[2][r] + acc[3][r]);
      if (row < d.m && i < nr) tf[static_cast<int64_t>(j0 + i) * tv + row] -= v;
    }
  }
}

// backward: partial[tile][j][c] = sum over the tile's 64 rows of L[r, c] t[r][j];
// the workgroup takes the tile's 64-column quarter blockIdx.z, wave w its 16
// columns c0 + 64 z + 16 w + [0, 16) (the MFMA's M); the rows are the K: lane
// group g reads rows 16 s + 4 g + [0, 4) of its column with one 16-byte
// access (two in FP64) and feeds them to steps (s, e = 0..3); the
// right-hand sides (staged t rows) are N.
template <typename T>
__global__ __launch_bounds__(256) void k_sbwd_part_m(const FrontDesc* __restrict__ fd, const SbPartTile* __restrict__ tiles,
                                                     const T* __restrict__ arena, const double* __restrict__ t,
                                                     int64_t tv, int nrhs, double* __restrict__ partial) {
  using M = Mfma<double>;
  static_assert(BC == 256 && PRT == 64, "four quarters of 64 columns, 64 rows per tile");
  constexpr int TP = MR + 4;  // row pitch of the staged t (the four lane groups on distinct banks)
  constexpr int TPT = PRT * MR / 256;
  __shared__ double ts[PRT][TP];
  const SbPartTile tl = tiles[blockIdx.x];
  const FrontDesc d = fd[tl.front];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int g = lane >> 4, i = lane & 15;
  const int rend = min(tl.r0 + PRT, d.m);
  const int span = tl.c1 - tl.c0;
  const int cq = 64 * static_cast<int>(blockIdx.z);
  if (cq >= span) return;
  const int64_t m = d.m;
  const double* tf = t + d.toff;
  double* out = partial + static_cast<int64_t>(blockIdx.x) * nrhs * BC;
  const bool active = cq + 16 * w < span;  // the wave holds a column of the tile
  // the first chunk's t rows (issued before the factor's loads)
  double tr[TPT];
  auto tload = [&](int j0) {
    const int nr = min(MR, nrhs - j0);
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
      const int idx = tid + 256 * q, r = idx % PRT, j = idx / PRT;
      tr[q] = (tl.r0 + r < rend && j < nr) ? tf[static_cast<int64_t>(j0 + j) * tv + tl.r0 + r] : 0.0;
    }
  };
  if (static_cast<int>(blockIdx.y) * MR < nrhs) tload(static_cast<int>(blockIdx.y) * MR);
  // the lane's rows of its column: quad s starts at row r0 + 16 s + 4 g; a
  // quad that crosses the tile's end is read entry by entry (clamped rows,
  // whose t is 0)
  T lv[4][4];  // [s][e]
  {
    const int c = min(tl.c0 + cq + 16 * w + i, tl.c1 - 1);
    const T* col = arena + d.off + static_cast<int64_t>(c) * m;
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
      const int q0 = tl.r0 + 16 * s4 + 4 * g;
      if (q0 + 3 < rend) {
        if constexpr (sizeof(T) == 4) {
          const f4u v = *reinterpret_cast<const f4u*>(col + q0);
#pragma unroll
          for (int e = 0; e < 4; ++e) lv[s4][e] = v[e];
        } else {
          const v2du v0 = *reinterpret_cast<const v2du*>(col + q0);
          const v2du v1 = *reinterpret_cast<const v2du*>(col + q0 + 2);
          lv[s4][0] = v0[0];
          lv[s4][1] = v0[1];
          lv[s4][2] = v1[0];
          lv[s4][3] = v1[1];
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) lv[s4][e] = col[min(q0 + e, rend - 1)];
      }
    }
  }
  for (int j0 = static_cast<int>(blockIdx.y) * MR; j0 < nrhs; j0 += MR * static_cast<int>(gridDim.y)) {
    const int nr = min(MR, nrhs - j0);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
      const int idx = tid + 256 * q;
      ts[idx % PRT][idx / PRT] = tr[q];
    }
    __syncthreads();
    // the next chunk's t in flight during this chunk's products
    const int jn = j0 + MR * static_cast<int>(gridDim.y);
    if (jn < nrhs) tload(jn);
    if (!active) continue;
    M::acc_t acc[2] = {M::acc_t{0, 0, 0, 0}, M::acc_t{0, 0, 0, 0}};
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        acc[e & 1] = M::op(static_cast<double>(lv[s4][e]), ts[16 * s4 + 4 * g + e][i], acc[e & 1]);
    // register r of lane l: column cq + 16 w + l/16 + 4 r, right-hand side j0 + l%16
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int cl = cq + 16 * w + g + 4 * r;
      if (cl < span && i < nr) out[static_cast<int64_t>(j0 + i) * BC + cl] = acc[0][r] + acc[1][r];
    }
  }
}

// ---------------------------------------- super-block inverses (explicit)
//
// X = L_JJ^-1 for every super-block, built once per factorization from the
// pivot-block inverses V_a and the off-diagonal 64x64 blocks of L:
// X_bb = V_b, X_ab = -V_a sum_{c=b}^{a-1} L_ac X_cb (a > b), one workgroup
// per 64-wide block column, FP64 MFMA.  The solves then apply X (forward) and
// X^T (backward) as parallel products instead of a sequence of 64-wide blocks.

// out[i][q] += sum_k Bt[i][k] Ak[k][q] for a 64x64 output (the k_prow layout:
// lane row i = wave*16 + lane%16, column q = bj*16 + orow(lane, r))
__device__ __forceinline__ void mm64(const double (*Ak)[NB + 1], const double (*Bt)[NB + 1], Mfma<double>::acc_t* o) {
  using M = Mfma<double>;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll 4
  for (int kk = 0; kk < NB; kk += 4) {
    const int kr = kk + lane / 16;
    const double b = Bt[wave * 16 + lane % 16][kr];
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) o[bj] = M::op(Ak[kr][bj * 16 + lane % 16], b, o[bj]);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_sbinv(const FrontDesc* __restrict__ fd, const SbInvTask* __restrict__ tasks,
                                               const T* __restrict__ arena, const double* __restrict__ vbuf,
                                               const int32_t* __restrict__ btask, const int64_t* __restrict__ sboff,
                                               double* __restrict__ xinv, double* __restrict__ xinvt) {
  using M = Mfma<double>;
  const SbInvTask v = tasks[blockIdx.x];
  const FrontDesc d = fd[v.front];
  __shared__ double As[NB][NB + 1];   // k-major operand (X_cb, then acc)
  __shared__ double Bs[NB][NB + 1];   // row-major operand (L_ac, then V_a)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int W = v.W, b = v.b, jb = min(NB, W - b);
  double* X = xinv + sboff[v.sb];
  double* Xt = xinvt + sboff[v.sb];   // X^T, column-major (the backward reads its columns' rows)
  const T* F = arena + d.off;
  const int64_t m = d.m;
  // X_bb = V_b; the rest of the block column above the diagonal is zero
  {
    const double* vb = vbuf + static_cast<int64_t>(btask[v.bt]) * NB * NB;  // column-major V
    for (int idx = tid; idx < NB * NB; idx += 256) {
      const int i = idx % NB, j = idx / NB;
      if (i < jb && j < jb) {
        const double x = vb[idx];
        X[(b + i) + static_cast<int64_t>(b + j) * W] = x;
        Xt[(b + j) + static_cast<int64_t>(b + i) * W] = x;
      }
    }
    for (int idx = tid; idx < b * jb; idx += 256) {
      const int i = idx % b, j = idx / b;
      X[i + static_cast<int64_t>(b + j) * W] = 0.0;
    }
  }
  for (int a = b + NB; a < W; a += NB) {
    const int ja = min(NB, W - a);
    M::acc_t o[4];
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) o[bj] = M::acc_t{0, 0, 0, 0};
    for (int c = b; c < a; c += NB) {
      const int jc = min(NB, W - c);
      __syncthreads();
      for (int idx = tid; idx < NB * NB; idx += 256) {
        const int i = idx % NB, j = idx / NB;
        // As[p][j] = X[c + p][b + j]; Bs[i][p] = L[J0 + a + i][J0 + c + p]
        As[i][j] = (i < jc && j < jb) ? X[(c + i) + static_cast<int64_t>(b + j) * W] : 0.0;
        Bs[i][j] = (i < ja && j < jc) ? static_cast<double>(F[(v.J0 + a + i) + (v.J0 + c + j) * m]) : 0.0;
      }
      __syncthreads();
      mm64(As, Bs, o);
    }
    // acc -> As (k-major: As[p][j] = acc[p][j]); V_a -> Bs (row-major Bs[i][p] = V_a[i][p])
    __syncthreads();
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) As[wave * 16 + lane % 16][bj * 16 + M::orow(lane, r)] = o[bj][r];
    {
      const double* va = vbuf + static_cast<int64_t>(btask[v.bt + (a - b) / NB]) * NB * NB;
      for (int idx = tid; idx < NB * NB; idx += 256) {
        const int i = idx % NB, p = idx / NB;
        Bs[i][p] = va[idx];
      }
    }
    __syncthreads();
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) o[bj] = M::acc_t{0, 0, 0, 0};
    mm64(As, Bs, o);
    const int i = wave * 16 + lane % 16;
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = bj * 16 + M::orow(lane, r);
        if (i < ja && q < jb) {
          X[(a + i) + static_cast<int64_t>(b + q) * W] = -o[bj][r];
          Xt[(b + q) + static_cast<int64_t>(a + i) * W] = -o[bj][r];
        }
      }
    // the next step reads this block back from X
    __syncthreads();
  }
}

// forward: y_J = X t_J for rows [64a, 64a + 64) of the super-block (a = blockIdx.y);
// the result goes to ys (t_J is read by the other row blocks)
template <int R, int NW>
__global__ __launch_bounds__(64 * NW) void k_sfwd_inv(const FrontDesc* __restrict__ fd,
                                                                const SbTask* __restrict__ tasks,
                                                                const double* __restrict__ xinv,
                                                                const int64_t* __restrict__ sboff,
                                                                const double* __restrict__ t, double* __restrict__ ys,
                                                                int64_t tv, int nrhs) {
  constexpr int CPW = sb_maxw<NW>() / NW;
  const SbTask tk = tasks[blockIdx.x];
  const int W = tk.J1 - tk.J0, a0 = blockIdx.y * NB;
  if (a0 >= W) return;
  const FrontDesc d = fd[tk.front];
  __shared__ double tJ[SBW][R];
  __shared__ double red[NW][NB][R + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i = a0 + lane;
  const int jend = min(W, a0 + NB);          // X is lower triangular
  const int cw = (jend + NW - 1) / NW, cb = w * cw, ce = min(jend, cb + cw);
  const double* X = xinv + sboff[tk.sb];
  const double* tf = t + d.toff + tk.J0;
  double* yf = ys + d.toff + tk.J0;
  const double* xr = X + min(i, W - 1);
  double xv[CPW];
#pragma unroll
  for (int c = 0; c < CPW; ++c) xv[c] = xr[static_cast<int64_t>(min(cb + c, jend - 1)) * W];
  for (int j0 = static_cast<int>(blockIdx.z) * R; j0 < nrhs; j0 += R * static_cast<int>(gridDim.z)) {
    const int nr = min(R, nrhs - j0);
    __syncthreads();
    for (int idx = tid; idx < jend * R; idx += 64 * NW) {
      const int q = idx % jend, j = idx / jend;
      tJ[q][j] = j < nr ? tf[(j0 + j) * tv + q] : 0.0;
    }
    __syncthreads();
    double acc[R];
#pragma unroll
    for (int j = 0; j < R; ++j) acc[j] = 0.0;
#pragma unroll
    for (int c = 0; c < CPW; ++c)
      if (cb + c < ce) {
#pragma unroll
        for (int j = 0; j < R; ++j) acc[j] = fma(xv[c], tJ[cb + c][j], acc[j]);
      }
#pragma unroll
    for (int j = 0; j < R; ++j) red[w][lane][j] = acc[j];
    __syncthreads();
    if (tid < NB && i < W)
      for (int j = 0; j < nr; ++j) yf[(j0 + j) * tv + i] = wave_tree_sum<NW>(red, lane, j);
  }
}

// backward: z_J = t_J - (partial sums in row-tile order) -> ys; 64 columns per
// block (blockIdx.y), the four waves sum contiguous quarters of the row tiles
template <int R, int NW>
__global__ __launch_bounds__(64 * NW) void k_sbwd_comb(const FrontDesc* __restrict__ fd,
                                                                 const SbTask* __restrict__ tasks,
                                                                 const double* __restrict__ partial,
                                                                 const double* __restrict__ t,
                                                                 double* __restrict__ ys, int64_t tv, int nrhs) {
  const SbTask tk = tasks[blockIdx.x];
  const int W = tk.J1 - tk.J0, c0 = blockIdx.y * NB;
  if (c0 >= W) return;
  const FrontDesc d = fd[tk.front];
  __shared__ double red[NW][NB][R + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i = c0 + lane;
  const int cc = i / BC, cl = i % BC;
  const int qw = (tk.nrc + NW - 1) / NW, rb = w * qw, re = min(tk.nrc, rb + qw);
  const double* tf = t + d.toff + tk.J0;
  double* zf = ys + d.toff + tk.J0;
  for (int j0 = static_cast<int>(blockIdx.z) * R; j0 < nrhs; j0 += R * static_cast<int>(gridDim.z)) {
    const int nr = min(R, nrhs - j0);
    double acc[R];
#pragma unroll
    for (int j = 0; j < R; ++j) acc[j] = 0.0;
    if (i < W) {
      const double* src = partial + ((static_cast<int64_t>(tk.p0) + static_cast<int64_t>(cc) * tk.nrc) * nrhs + j0) * BC + cl;
      // the loads of sixteen row chunks are issued together
#pragma unroll 16
      for (int rc = rb; rc < re; ++rc)
#pragma unroll
        for (int j = 0; j < R; ++j)
          if (j < nr) acc[j] += src[(static_cast<int64_t>(rc) * nrhs + j) * BC];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) red[w][lane][j] = acc[j];
    __syncthreads();
    if (tid < NB && i < W)
      for (int j = 0; j < nr; ++j) zf[(j0 + j) * tv + i] = tf[(j0 + j) * tv + i] - wave_tree_sum<NW>(red, lane, j);
  }
}

// backward: x_J = X^T z_J for rows [64a, 64a + 64) (a = blockIdx.y) from the
// stored transpose: lane = output row, the four waves split the columns
template <int R, int NW>
__global__ __launch_bounds__(64 * NW) void k_sbwd_inv(const FrontDesc* __restrict__ fd,
                                                                const SbTask* __restrict__ tasks,
                                                                const double* __restrict__ xinvt,
                                                                const int64_t* __restrict__ sboff,
                                                                const double* __restrict__ ys, double* __restrict__ t,
                                                                int64_t tv, int nrhs) {
  constexpr int CPW = sb_maxw<NW>() / NW;
  const SbTask tk = tasks[blockIdx.x];
  const int W = tk.J1 - tk.J0, a0 = blockIdx.y * NB;
  if (a0 >= W) return;
  const FrontDesc d = fd[tk.front];
  __shared__ double zs[SBW][R];
  __shared__ double red[NW][NB][R + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i = a0 + lane;
  const int span = W - a0;                      // X^T is upper triangular: columns q >= a0
  const int cw = (span + NW - 1) / NW, cb = a0 + w * cw, ce = min(W, cb + cw);
  const double* Xt = xinvt + sboff[tk.sb];
  const double* zf = ys + d.toff + tk.J0;
  double* tf = t + d.toff + tk.J0;
  const double* xr = Xt + min(i, W - 1);
  double xv[CPW];
#pragma unroll
  for (int c = 0; c < CPW; ++c) xv[c] = xr[static_cast<int64_t>(min(cb + c, W - 1)) * W];
  for (int j0 = static_cast<int>(blockIdx.z) * R; j0 < nrhs; j0 += R * static_cast<int>(gridDim.z)) {
    const int nr = min(R, nrhs - j0);
    __syncthreads();
    for (int idx = tid; idx < span * R; idx += 64 * NW) {
      const int q = a0 + idx % span, j = idx / span;
      zs[q][j] = j < nr ? zf[(j0 + j) * tv + q] : 0.0;
    }
    __syncthreads();
    double acc[R];
#pragma unroll
    for (int j = 0; j < R; ++j) acc[j] = 0.0;
#pragma unroll
    for (int c = 0; c < CPW; ++c)
      if (cb + c < ce) {
#pragma unroll
        for (int j = 0; j < R; ++j) acc[j] = fma(xv[c], zs[cb + c][j], acc[j]);
      }
#pragma unroll
    for (int j = 0; j < R; ++j) red[w][lane][j] = acc[j];
    __syncthreads();
    if (tid < NB && i < W)
      for (int j = 0; j < nr; ++j) tf[(j0 + j) * tv + i] = wave_tree_sum<NW>(red, lane, j);
  }
}

// ------------------------------------------------ fused solves of small fronts
//
// Levels whose fronts all have at most SMM rows (SMM_MRHS with more than
// four right-hand sides) and SMK pivots (the lower levels of 3D trees, most
// levels of 2D trees, every level of a batch of small plates) are solved by
// one wave per front and one launch per level and direction, instead of five
// or six launches of four-wave workgroups.  Each step repeats the arithmetic
// of the four-wave kernel it replaces, in the same order (forward:
// k_fwd_init, k_fwd_child per slot, k_sfwd_inv, k_fwd_finish_ys, k_sfwd_rows;
// backward: k_bwd_init, k_sbwd_part, k_sbwd_comb, k_sbwd_inv, k_bwd_finish):
// a lane forms the four waves' partial sums of its row one after another and
// adds them in the waves' pairwise order, so the results are bit-identical.
// The front's vector stays in LDS between the steps; the LDS is sized per
// level (ldm rows, ldk pivots).  Such a front has one super-block
// (SMK <= SBW), at most SMM / PRT backward row tiles and one 256-column chunk.

// sum over the four waves' column ranges of a 64-row block of the four-wave
// kernels: v[w][j] = sum_{c in [cb_w, ce_w)} a(c) s[c][j], in order, then
// (v0 + v2) + (v1 + v3).  lo is the block's first column, span its width.
template <int R, typename LoadA>
__device__ __forceinline__ void four_wave_sums(LoadA a, const double* __restrict__ s, int ld, int lo, int span,
                                               double (&out)[R]) {
  const int cw = (span + 3) / 4;
  double v[4][R];
#pragma unroll
  for (int ww = 0; ww < 4; ++ww) {
    const int cb = lo + ww * cw, ce = min(lo + span, cb + cw);
#pragma unroll
    for (int j = 0; j < R; ++j) v[ww][j] = 0.0;
#pragma unroll 4
    for (int c = cb; c < ce; ++c) {
      const double av = a(c);
#pragma unroll
      for (int j = 0; j < R; ++j) v[ww][j] = fma(av, s[c * ld + j], v[ww][j]);
    }
  }
#pragma unroll
  for (int j = 0; j < R; ++j) out[j] = (v[0][j] + v[2][j]) + (v[1][j] + v[3][j]);
}

// forward, one wave per front: t = x[pivots] (0 below), += the children's
// update vectors (slot 0, then 1), y = X t[0:k], x[pivots] = y / D,
// t[k:m] -= L[k:m, 0:k] y, and t[k:m] out for the parent
template <typename T, int R>
__global__ __launch_bounds__(64) void k_small_fwd(const FrontDesc* __restrict__ fd, const SmallFront* __restrict__ sf,
                                                  const int32_t* __restrict__ cmap, const T* __restrict__ arena,
                                                  const double* __restrict__ xinv, const int64_t* __restrict__ sboff,
                                                  const double* __restrict__ vbuf, const double* __restrict__ dpiv,
                                                  double* __restrict__ x, double* __restrict__ t, int64_t n, int64_t tv,
                                                  int nrhs, int ldm) {
  static_assert(SMK <= SBW && SMK == sb_maxw<4>() && SRB == 64, "one super-block per front, 64-row update tiles");
  constexpr int LD = R + 1;
  extern __shared__ double sh[];
  double* ts = sh;              // ldm x LD: the front's vector
  double* ys = sh + ldm * LD;   // k x LD: y
  const SmallFront sm = sf[blockIdx.x];
  const FrontDesc d = fd[sm.front];
  const int lane = threadIdx.x;
  const int m = d.m, k = d.k, W = d.k;
  // X = L^-1 of the front's pivots: for at most NB pivots the pivot block's
  // inverse V itself (which is what the super-block inverse holds then),
  // column stride NB
  const double* X = k <= 0 ? nullptr : k <= NB ? vbuf + static_cast<int64_t>(sm.vtask) * NB * NB : xinv + sboff[sm.sb];
  const int64_t ldx = k <= NB ? NB : W;
  for (int j0 = static_cast<int>(blockIdx.y) * R; j0 < nrhs; j0 += R * static_cast<int>(gridDim.y)) {
    const int nr = min(R, nrhs - j0);
    __syncthreads();
    for (int idx = lane; idx < m * R; idx += 64) {
      const int i = idx % m, j = idx / m;
      ts[i * LD + j] = (i < k && j < nr) ? x[static_cast<int64_t>(j0 + j) * n + d.start + i] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int sl = 0; sl < 2; ++sl) {
      if (sm.child[sl] < 0) continue;
      const FrontDesc c = fd[sm.child[sl]];
      const int32_t* cm = cmap + c.cmap_off;
      for (int idx = lane; idx < c.u * R; idx += 64) {
        const int xx = idx % c.u, j = idx / c.u;
        if (j < nr) ts[cm[xx] * LD + j] += t[static_cast<int64_t>(j0 + j) * tv + c.toff + c.k + xx];
      }
      __syncthreads();
    }
    if (k > 0) {
      // y = X t[0:k] by 64-row blocks (k_sfwd_inv)
      for (int a0 = 0; a0 < W; a0 += 64) {
        const int i = a0 + lane;
        if (i < W) {
          const double* xr = X + i;
          double y[R];
          four_wave_sums<R>([&](int c) { return xr[static_cast<int64_t>(c) * ldx]; }, ts, LD, 0, min(W, a0 + 64), y);
#pragma unroll
          for (int j = 0; j < R; ++j) ys[i * LD + j] = j < nr ? y[j] : 0.0;
        }
      }
      __syncthreads();
      // x[pivots] = y / D (k_fwd_finish_ys)
      for (int idx = lane; idx < k * R; idx += 64) {
        const int i = idx % k, j = idx / k;
        if (j < nr) x[static_cast<int64_t>(j0 + j) * n + d.start + i] = ys[i * LD + j] / dpiv[d.start + i];
      }
      // t[k:m] -= L[k:m, 0:k] y by 64-row tiles (k_sfwd_rows)
      for (int r0 = k; r0 < m; r0 += SRB) {
        const int r = r0 + lane;
        if (r < m) {
          const T* F = arena + d.off + r;
          double v[R];
          four_wave_sums<R>([&](int c) { return static_cast<double>(F[static_cast<int64_t>(c) * m]); }, ys, LD, 0, W,
                            v);
          for (int j = 0; j < nr; ++j) ts[r * LD + j] -= v[j];
        }
      }
      __syncthreads();
    }
    // the update vector for the parent
    const int u = m - k;
    for (int idx = lane; idx < u * R; idx += 64) {
      const int i = k + idx % u, j = idx / u;
      if (j < nr) t[static_cast<int64_t>(j0 + j) * tv + d.toff + i] = ts[i * LD + j];
    }
  }
}

// backward, one wave per front: t = x[pivots] and the parent's final entries
// below, z = t[0:k] - L[k:m, 0:k]^T t[k:m] (partial sums per 64-row tile,
// combined in the order of k_sbwd_comb), t[0:k] = X^T z, x[pivots] = t[0:k],
// and the whole t out for the children
template <typename T, int R>
__global__ __launch_bounds__(64) void k_small_bwd(const FrontDesc* __restrict__ fd, const SmallFront* __restrict__ sf,
                                                  const int32_t* __restrict__ cmap, const T* __restrict__ arena,
                                                  const double* __restrict__ xinvt, const int64_t* __restrict__ sboff,
                                                  const double* __restrict__ vbuf, double* __restrict__ x,
                                                  double* __restrict__ t, int64_t n, int64_t tv, int nrhs, int ldm,
                                                  int ldk) {
  static_assert(SMK <= SBW && SMK == sb_maxw<4>() && SMK <= BC && PRT == 64,
                "one super-block and one column chunk per front");
  constexpr int LD = R + 1;
  extern __shared__ double sh[];
  double* ts = sh;                   // ldm x LD: the front's vector
  double* zs = sh + ldm * LD;        // ldk x LD: z
  double* part = zs + ldk * LD;      // (ldm / PRT rounded up) x ldk x LD: partial sums per row tile
  const SmallFront sm = sf[blockIdx.x];
  const FrontDesc d = fd[sm.front];
  const int lane = threadIdx.x;
  const int m = d.m, k = d.k, W = d.k;
  const int64_t ptoff = d.parent >= 0 ? fd[d.parent].toff : 0;
  const T* F = arena + d.off;
  const int nrc = (m - k + PRT - 1) / PRT;
  for (int j0 = static_cast<int>(blockIdx.y) * R; j0 < nrhs; j0 += R * static_cast<int>(gridDim.y)) {
    const int nr = min(R, nrhs - j0);
    __syncthreads();
    for (int idx = lane; idx < m * R; idx += 64) {
      const int i = idx % m, j = idx / m;
      double v = 0.0;
      if (j < nr)
        v = i < k ? x[static_cast<int64_t>(j0 + j) * n + d.start + i]
                  : t[static_cast<int64_t>(j0 + j) * tv + ptoff + cmap[d.cmap_off + i - k]];
      ts[i * LD + j] = v;
    }
    __syncthreads();
    if (k > 0) {
      // partial sums of 64-row tiles (k_sbwd_part): 16 lanes per column,
      // rows rl + 16 s of the tile, summed across the DPP row; four columns
      // per step
      const int rl = lane & 15, cg = lane >> 4;
      for (int rc = 0; rc < nrc; ++rc) {
        const int r0 = k + rc * PRT, rend = min(r0 + PRT, m);
        for (int c4 = 0; c4 < k; c4 += 4) {
          const int cl = c4 + cg;
          const T* col = F + static_cast<int64_t>(min(cl, k - 1)) * m;
          double lv[4], tr[4][R];
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int row = r0 + rl + 16 * s;
            lv[s] = static_cast<double>(col[min(row, rend - 1)]);
#pragma unroll
            for (int j = 0; j < R; ++j) tr[s][j] = (row < rend && j < nr) ? ts[row * LD + j] : 0.0;
          }
          double acc[R];
#pragma unroll
          for (int j = 0; j < R; ++j) {
            acc[j] = 0.0;
#pragma unroll
            for (int s = 0; s < 4; ++s) acc[j] = fma(lv[s], tr[s][j], acc[j]);
            acc[j] = row16_sum(acc[j]);
          }
          if (rl == 0 && cl < k)
#pragma unroll
            for (int j = 0; j < R; ++j) part[(rc * ldk + cl) * LD + j] = acc[j];
        }
      }
      __syncthreads();
      // z = t[0:k] - the partial sums (k_sbwd_comb: wave w adds row tiles
      // [w qw, w qw + qw) in order, then the pairwise sum of the waves)
      const int qw = (nrc + 3) / 4;
      for (int idx = lane; idx < W * R; idx += 64) {
        const int i = idx % W, j = idx / W;
        double v[4];
#pragma unroll
        for (int ww = 0; ww < 4; ++ww) {
          const int rb = ww * qw, re = min(nrc, rb + qw);
          double a = 0.0;
          if (j < nr)
            for (int rc = rb; rc < re; ++rc) a += part[(rc * ldk + i) * LD + j];
          v[ww] = a;
        }
        zs[i * LD + j] = j < nr ? ts[i * LD + j] - ((v[0] + v[2]) + (v[1] + v[3])) : 0.0;
      }
      __syncthreads();
      // t[0:k] = X^T z by 64-row blocks (k_sbwd_inv; X^T is upper triangular)
      // row i of X^T: for at most NB pivots column i of V (stride 1), else
      // row i of the stored transpose (stride W)
      const bool small_k = k <= NB;
      const double* Xt = small_k ? vbuf + static_cast<int64_t>(sm.vtask) * NB * NB : xinvt + sboff[sm.sb];
      for (int a0 = 0; a0 < W; a0 += 64) {
        const int i = a0 + lane;
        if (i < W) {
          const double* xr = small_k ? Xt + static_cast<int64_t>(i) * NB : Xt + i;
          const int64_t st = small_k ? 1 : W;
          double v[R];
          four_wave_sums<R>([&](int c) { return xr[static_cast<int64_t>(c) * st]; }, zs, LD, a0, W - a0, v);
          for (int j = 0; j < nr; ++j) ts[i * LD + j] = v[j];
        }
      }
      __syncthreads();
      for (int idx = lane; idx < k * R; idx += 64) {
        const int i = idx % k, j = idx / k;
        if (j < nr) x[static_cast<int64_t>(j0 + j) * n + d.start + i] = ts[i * LD + j];
      }
    }
    for (int idx = lane; idx < m * R; idx += 64) {
      const int i = idx % m, j = idx / m;
      if (j < nr) t[static_cast<int64_t>(j0 + j) * tv + d.toff + i] = ts[i * LD + j];
    }
  }
}

// x[pivots] = y / D from the forward result in ys
__global__ __launch_bounds__(256) void k_fwd_finish_ys(const FrontDesc* __restrict__ fd, const int32_t* __restrict__ fronts,
                                                       const double* __restrict__ ys, const double* __restrict__ dpiv,
                                                       double* __restrict__ x, int64_t n, int64_t tv, int nrhs) {
  const FrontDesc d = fd[fronts[blockIdx.x]];
  const int j = blockIdx.y, i0 = blockIdx.z * VCH, i1 = min(d.k, i0 + VCH);
  for (int i = i0 + threadIdx.x; i < i1; i += 256) x[j * n + d.start + i] = ys[j * tv + d.toff + i] / dpiv[d.start + i];
}

// ------------------------------------------------------------- refinement

// the failing-pivot word to "none" (a kernel, not a memset node, so that the
// factorization can be captured into a graph)
__global__ void k_err_reset(unsigned long long* err) { *err = ULLONG_MAX; }

__global__ void k_gather(const int64_t* __restrict__ map, const double* __restrict__ src, double* __restrict__ dst,
                         int64_t cnt) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < cnt) dst[i] = src[map[i]];
}

// The sum over a 16-lane DPP row in the order of the butterfly
// v += shfl_xor(v, o) for o = 8, 4, 2, 1, with row rotations instead of
// permutes through LDS: after the steps for 8, ..., 2o the values repeat
// with period 2o within the row, so lane l - o (mod 16) holds what lane
// l ^ o holds, and every lane adds the same two operands in the same order.
__device__ __forceinline__ double row16_sum_xor(double v) {
  v += dpp_d<0x128>(v);  // row_ror:8
  v += dpp_d<0x124>(v);  // row_ror:4
  v += dpp_d<0x122>(v);  // row_ror:2
  v += dpp_d<0x121>(v);  // row_ror:1
  return v;
}

// r = b - A x for a CSR matrix, LPR lanes per row (16, or 4 for short rows),
// column per right-hand side; RR right-hand sides per pass, so that a row of
// A is read once per pass (every column's sum is formed in the same order
// whatever RR is)
template <int RR, int LPR>
__global__ __launch_bounds__(256) void k_residual(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                  const double* __restrict__ av, const double* __restrict__ x,
                                                  const double* __restrict__ b, double* __restrict__ r, int64_t n,
                                                  int nrhs) {
  static_assert(LPR == 16 || LPR == 4, "16 or 4 lanes per row");
  const int64_t row = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / LPR;
  const int sub = threadIdx.x % LPR;
  if (row >= n) return;
  const int64_t p0 = rp[row], p1 = rp[row + 1];
  for (int j0 = 0; j0 < nrhs; j0 += RR) {
    const int nr = min(RR, nrhs - j0);
    double s[RR];
#pragma unroll
    for (int j = 0; j < RR; ++j) s[j] = 0.0;
#pragma unroll 4
    for (int64_t p = p0 + sub; p < p1; p += LPR) {
      const double a = av[p];
      const int64_t c = ci[p];
#pragma unroll
      for (int j = 0; j < RR; ++j)
        if (j < nr) s[j] = fma(a, x[(j0 + j) * n + c], s[j]);
    }
#pragma unroll
    for (int j = 0; j < RR; ++j) {
      if constexpr (LPR == 16) {
        s[j] = row16_sum_xor(s[j]);
      } else {
        s[j] += dpp_d<0x4E>(s[j]);  // quad_perm [2,3,0,1]: lane ^ 2
        s[j] += dpp_d<0xB1>(s[j]);  // quad_perm [1,0,3,2]: lane ^ 1
      }
      if (sub == 0 && j < nr) r[(j0 + j) * n + row] = b[(j0 + j) * n + row] - s[j];
    }
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

// sums of squares of every segment [seg[s], seg[s+1]) of every column, one
// block per segment; SR columns per pass, read side by side (their loads in
// flight together), each summed in the same order
constexpr int SR_SQ = 8;
__global__ __launch_bounds__(256) void k_seg_sumsq(const double* __restrict__ v, int64_t n, int nrhs,
                                                   const int64_t* __restrict__ seg, int nseg, double* __restrict__ out) {
  __shared__ double sh[4][SR_SQ];
  for (int sg = blockIdx.x; sg < nseg; sg += gridDim.x)
    for (int j0 = 0; j0 < nrhs; j0 += SR_SQ) {
      const int nr = min(SR_SQ, nrhs - j0);
      double s[SR_SQ];
#pragma unroll
      for (int j = 0; j < SR_SQ; ++j) s[j] = 0.0;
      for (int64_t i = seg[sg] + threadIdx.x; i < seg[sg + 1]; i += 256) {
#pragma unroll
        for (int j = 0; j < SR_SQ; ++j)
          if (j < nr) {
            const double x = v[(j0 + j) * n + i];
            s[j] += x * x;
          }
      }
#pragma unroll
      for (int j = 0; j < SR_SQ; ++j)
        for (int o = 32; o > 0; o >>= 1) s[j] += __shfl_xor(s[j], o, 64);
      __syncthreads();
      if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int j = 0; j < SR_SQ; ++j) sh[threadIdx.x >> 6][j] = s[j];
      __syncthreads();
      if (threadIdx.x < nr)
        out[static_cast<int64_t>(sg) * nrhs + j0 + threadIdx.x] =
            (sh[0][threadIdx.x] + sh[1][threadIdx.x]) + (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
    }
}

// one workgroup of 64 per right-hand side j (blockIdx.x): lane t adds the
// partials t, t + 64, ... in order, then lane 0 adds the 64 lane sums in
// order (a fixed order: the same sums on every call)
__global__ __launch_bounds__(64) void k_sumsq_final(const double* __restrict__ part, int nblk, int nrhs,
                                                    double* __restrict__ out) {
  __shared__ double ls[64];
  const int j = blockIdx.x;
  if (j >= nrhs) return;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64) s += part[static_cast<int64_t>(j) * nblk + b];
  ls[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < 64; ++i) t += ls[i];
    out[j] = t;
  }
}

// x += alpha c (alpha = 1: a plain sum; -1 takes a correction back)
__global__ void k_axpy(double* __restrict__ x, const double* __restrict__ c, int64_t cnt, double alpha) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < cnt) x[i] = fma(alpha, c[i], x[i]);
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

// Work vectors of one solve: front vectors t and ys (sum of m per column),
// backward partial sums, x, b, two residual/correction vectors r and c, and
// the norm reduction buffers.
struct SolveWork {
  double *t = nullptr, *ys = nullptr, *part = nullptr, *x = nullptr, *b = nullptr, *r = nullptr, *c = nullptr,
         *red = nullptr, *nrm = nullptr;
};

// Scratch of the allocation-free solve (rustyfem_solve_device_noalloc): 64
// MiB per device, held by every device the library runs on until the process
// ends.  Host code never uses this name as an address (it is a host-side
// stand-in); scratch_ptr() resolves the device address once per device.
constexpr size_t kScratchElems = size_t{8} << 20;  // doubles (64 MiB)
constexpr int kNoAllocMaxRhs = 64;
constexpr int kMaxDevices = 64;

// Makes dev current for the lifetime of the guard, then restores the caller's device.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    HIPCHK(hipGetDevice(&prev));
    if (prev != dev) HIPCHK(hipSetDevice(dev));
  }
  ~DeviceGuard() {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

}  // namespace

static __device__ __align__(256) double scratch[kScratchElems];

namespace {

// Device address of `scratch` on device dev (which must be current): looked up
// with hipGetSymbolAddress at the first call on that device and kept.
double* scratch_ptr(int dev) {
  static std::mutex mu;
  static double* ptr[kMaxDevices] = {};
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("allocation-free solve: device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      char m[200];
      snprintf(m, sizeof m, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
      throw std::runtime_error(m);
    }
    ptr[dev] = static_cast<double*>(p);
  }
  return ptr[dev];
}

template <typename T>
class Impl final : public GpuFactor {
 public:
  Impl(Plan plan, const std::vector<int64_t>& colptr, const std::vector<int64_t>& rowidx) : P_(std::move(plan)) {
    HIPCHK(hipGetDevice(&dev_));
    HIPCHK(hipStreamCreate(&s_));
    HIPCHK(hipStreamCreate(&s2_));
    HIPCHK(hipEventCreateWithFlags(&ev_look_, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_rest_, hipEventDisableTiming));
    {
      const char* la = std::getenv("RUSTYFEM_NO_LOOKAHEAD");
      lookahead_ = !(la && *la && *la != '0');
    }
    const char* ng = std::getenv("RUSTYFEM_NO_GRAPH");
    use_graphs_ = !(ng && *ng && *ng != '0');
    const char* nf = std::getenv("RUSTYFEM_NO_FACTOR_GRAPH");
    factor_graph_ = use_graphs_ && !(nf && *nf && *nf != '0');
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
    d_panel_ = dupload(P_.panel, bytes_, s_);
    d_srow_ = dupload(P_.srow, bytes_, s_);
    d_btile_ = dupload(P_.btile, bytes_, s_);
    d_l21_ = dupload(P_.l21, bytes_, s_);
    d_l21c_ = dupload(P_.l21c, bytes_, s_);
    d_last_ = dupload(P_.bwd_last, bytes_, s_);
    d_sbt_ = dupload(P_.sbt, bytes_, s_);
    d_sbrow_ = dupload(P_.sbrow, bytes_, s_);
    d_sbpart_ = dupload(P_.sbpart, bytes_, s_);
    d_btask_ = dupload(P_.btask, bytes_, s_);
    d_sboff_ = dupload(P_.sbinv_off, bytes_, s_);
    d_xinv_ = dalloc<double>(static_cast<size_t>(std::max<int64_t>(P_.sbinv_elems, 1)), bytes_);
    d_xinvt_ = dalloc<double>(static_cast<size_t>(std::max<int64_t>(P_.sbinv_elems, 1)), bytes_);
    {
      const char* e = std::getenv("RUSTYFEM_LEGACY_SOLVE");
      legacy_solve_ = e && *e && *e != '0';
      const char* q = std::getenv("RUSTYFEM_LEGACY_PANEL");
      legacy_panel_ = q && *q && *q != '0';
      const char* c = std::getenv("RUSTYFEM_SOLVE_CHAIN");
      chain_solve_ = c && *c && *c != '0';
      const char* mm = std::getenv("RUSTYFEM_NO_MRHS_MFMA");
      mrhs_mfma_ = !(mm && *mm && *mm != '0');
      const char* rb = std::getenv("RUSTYFEM_SOLVE_RHS_BLOCK");
      if (rb && *rb) rhs_block_ = std::atoi(rb);
      // RUSTYFEM_TUNE_MARGIN: a variant must be this fraction faster to be picked
      const char* tm = std::getenv("RUSTYFEM_TUNE_MARGIN");
      if (tm && *tm) tune_margin_ = 1.0f - static_cast<float>(std::atof(tm));
      // RUSTYFEM_AUTOTUNE: unset or 1 = time every GEMM kernel per launch,
      // 0 = k_gemm everywhere, "agpr" = k_gemm_agpr everywhere, "ilv" =
      // k_gemm_ilv everywhere, "k64" = k_gemm_k64 everywhere, "q" = k_gemm_q
      // everywhere (FP32; FP64 factors take k_gemm)
      const char* a = std::getenv("RUSTYFEM_AUTOTUNE");
      const int8_t force = a && std::strcmp(a, "agpr") == 0  ? 1
                           : a && std::strcmp(a, "ilv") == 0 ? 2
                           : a && std::strcmp(a, "k64") == 0 ? 3
                           : a && std::strcmp(a, "q") == 0   ? (kGemmVariants > 4 ? 4 : 0)
                                                             : 0;
      autotune_ = !(a && *a == '0') && force == 0;
      gemm_pick_.assign(P_.gemm.size(), force);
      // In a GEMM launch where most tiles update at most 64 rows and columns
      // (small fronts), those tiles go first in its range and take the
      // one-wave 64 x 64 kernel, the others the 128 x 128 kernels; in the
      // other launches a few such tiles (the corners of large fronts, with
      // long K) stay with the rest, beside which they run
      // (RUSTYFEM_NO_GEMM64=1: every tile on the 128 x 128 kernels).  A
      // tile's result does not depend on its place in the launch.
      const char* g64 = std::getenv("RUSTYFEM_NO_GEMM64");
      gemm64_.assign(P_.gemm.size(), 0);
      if (!(g64 && *g64 && *g64 != '0')) {
        bool moved = false;
        auto fits = [&](const GemmTile& gt) { return P_.fd[gt.front].m - gt.r0 <= 64 && gt.cend - gt.c0 <= 64; };
        for (const LevelPlan& lp : P_.levels)
          for (const Phase& ph : lp.phases) {
            if ((ph.kind != PhaseKind::Gemm128 && ph.kind != PhaseKind::GemmRest) || ph.r.count <= 0) continue;
            auto b = P_.gemm.begin() + ph.r.begin, e = b + ph.r.count;
            if (2 * std::count_if(b, e, fits) < ph.r.count) continue;
            auto mid = std::stable_partition(b, e, fits);
            gemm64_[ph.r.begin] = static_cast<int32_t>(mid - b);
            moved = moved || (mid != b && mid != e);
          }
        if (moved) {
          HIPCHK(hipMemcpyAsync(d_gemm_, P_.gemm.data(), P_.gemm.size() * sizeof(GemmTile), hipMemcpyHostToDevice, s_));
          HIPCHK(hipStreamSynchronize(s_));
        }
      }
    }
    d_lf_ = dupload(P_.level_fronts, bytes_, s_);
    for (int sl = 0; sl < 2; ++sl) d_cs_[sl] = dupload(P_.child_slot_fronts[sl], bytes_, s_);
    // longest front vectors per level (grid extents of the per-front vector kernels)
    lvl_maxm_.assign(P_.levels.size(), 0);
    lvl_maxk_.assign(P_.levels.size(), 0);
    for (size_t L = 0; L < P_.levels.size(); ++L)
      for (int32_t f : P_.levels[L].fronts) {
        lvl_maxm_[L] = std::max<int64_t>(lvl_maxm_[L], P_.fd[f].m);
        lvl_maxk_[L] = std::max<int64_t>(lvl_maxk_[L], P_.fd[f].k);
      }
    // levels solved by the fused small-front kernels, those whose fronts all
    // have at most SMM rows and SMK pivots (RUSTYFEM_NO_SMALL_SOLVE=1: none,
    // every level through the super-block kernels)
    {
      const char* e = std::getenv("RUSTYFEM_NO_SMALL_SOLVE");
      const bool off = e && *e && *e != '0';
      std::vector<std::array<int32_t, 2>> ch(static_cast<size_t>(P_.nf), std::array<int32_t, 2>{-1, -1});
      for (int sl = 0; sl < 2; ++sl)
        for (int32_t c : P_.child_slot_fronts[sl]) ch[P_.fd[c].parent][sl] = c;
      std::vector<SmallFront> sf;
      std::vector<char> vonly(static_cast<size_t>(P_.nf), 0);  // solved from V alone (no super-block inverse)
      small_r_.assign(P_.levels.size(), Range{});
      small_ok_.assign(P_.levels.size(), false);
      for (size_t L = 0; L < P_.levels.size(); ++L) {
        small_ok_[L] = lvl_maxm_[L] <= SMM && lvl_maxk_[L] <= SMK;
        if (off || !small_ok_[L]) continue;
        const int64_t b = static_cast<int64_t>(sf.size());
        // a level that every solve takes through the fused kernels (any
        // number of right-hand sides, and room in LDS at 16 per pass) needs
        // no super-block inverses for its fronts of at most NB pivots
        const bool always = lvl_maxm_[L] <= SMM_MRHS && small_lds<16>(L, false) <= kSmallLdsMax &&
                            small_lds<16>(L, true) <= kSmallLdsMax;
        for (int32_t f : P_.levels[L].fronts) {
          const int32_t k = P_.fd[f].k;
          sf.push_back(SmallFront{f, k > 0 ? P_.sb_first[f] : -1, {ch[f][0], ch[f][1]},
                                  k > 0 ? P_.btask[P_.btask_off[f]] : -1});
          if (always && k > 0 && k <= NB) vonly[f] = 1;
        }
        small_r_[L] = Range{b, static_cast<int64_t>(sf.size()) - b};
        tm_.small_levels += 1;
      }
      d_small_ = dupload(sf, bytes_, s_);
      // the super-block inverses the solves read: none for the fronts of at
      // most NB pivots on the levels always fused (they read V)
      std::vector<SbInvTask> inv;
      inv.reserve(P_.sbinv.size());
      for (const SbInvTask& tk : P_.sbinv)
        if (!vonly[tk.front]) inv.push_back(tk);
      nsbinv_ = static_cast<int64_t>(inv.size());
      d_sbinv_ = dupload(inv, bytes_, s_);
    }
    for (int sl = 0; sl < 2; ++sl) {
      slot_maxu_[sl].assign(P_.levels.size(), 0);
      for (size_t L = 0; L < P_.levels.size() && L < P_.child_slot_range[sl].size(); ++L) {
        const Range& cr = P_.child_slot_range[sl][L];
        for (int64_t q = cr.begin; q < cr.begin + cr.count; ++q)
          slot_maxu_[sl][L] = std::max<int64_t>(slot_maxu_[sl][L], P_.fd[P_.child_slot_fronts[sl][q]].u);
      }
    }
    // 64 entries of padding: k_gemm_q's 16-byte loads of a quad of rows that
    // starts at a front's last row read up to three entries past its last column
    d_arena_ = dalloc<T>(static_cast<size_t>(P_.arena_elems) + 64, bytes_);
    d_dpiv_ = dalloc<double>(n_, bytes_);
    d_adiag_ = dalloc<double>(n_, bytes_);
    d_lval_ = dalloc<double>(nlower_, bytes_);
    d_vbuf_ = dalloc<double>(static_cast<size_t>(std::max<int64_t>(P_.diag.size(), 1)) * NB * NB, bytes_);
    // the pivot-block inverses are written on their jb x jb blocks only: the
    // rest stays zero from here on (once, outside any captured graph)
    HIPCHK(hipMemsetAsync(d_vbuf_, 0, static_cast<size_t>(std::max<int64_t>(P_.diag.size(), 1)) * NB * NB * sizeof(double),
                          s_));

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
    (void)hipStreamSynchronize(s2_);
    (void)hipStreamSynchronize(s_);
    drop_graphs();
    if (fgraph_) (void)hipGraphExecDestroy(fgraph_);
    void* ptrs[] = {d_fd_, d_cmap_, d_amap_, d_zero_, d_scatter_, d_ea_, d_gemm_, d_diag_, d_trsm_, d_panel_, d_srow_, d_btile_, d_l21_, d_l21c_, d_last_, d_sbt_, d_sbrow_, d_sbpart_, d_btask_, d_sbinv_, d_sboff_, d_xinv_, d_xinvt_, d_ys_, d_lf_, d_cs_[0],
                    d_cs_[1], d_arena_, d_dpiv_, d_adiag_, d_lval_, d_vbuf_, d_err_, d_dpos_, d_rp_, d_ci_, d_gmap_,
                    d_av_, d_t_, d_x_, d_part_, d_b_, d_r_, d_c_, d_red_, d_nrm_, d_bc_, d_eq_, d_iso_, d_seg_, d_segsq_, d_small_};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    for (hipEvent_t e : evs_) (void)hipEventDestroy(e);
    (void)hipEventDestroy(e0_);
    (void)hipEventDestroy(e1_);
    (void)hipEventDestroy(ev_look_);
    (void)hipEventDestroy(ev_rest_);
    (void)hipStreamDestroy(s2_);
    (void)hipStreamDestroy(s_);
  }

  // The factorization is a fixed sequence of launches on fixed buffers once
  // the GEMM kernels are tuned: from the second call on it is captured once
  // into a HIP graph (both streams, their events as graph edges) and
  // replayed.  Kernel timing (per-GEMM events) and RUSTYFEM_NO_GRAPH /
  // RUSTYFEM_NO_FACTOR_GRAPH keep it eager.
  int64_t factor(const double* vals, bool on_device) override {
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (vals) HIPCHK(hipMemcpyAsync(d_lval_, vals, nlower_ * sizeof(double), kind, s_));
    const bool graph = factor_graph_ && !ktime_ && (tuned_ || !autotune_);
    if (graph) {
      if (!fgraph_) {
        HIPCHK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        try {
          factor_launches(false);
        } catch (...) {
          hipGraph_t dummy = nullptr;
          (void)hipStreamEndCapture(s_, &dummy);
          if (dummy) (void)hipGraphDestroy(dummy);
          rest_pending_ = false;
          throw;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamEndCapture(s_, &g));
        HIPCHK(hipGraphInstantiate(&fgraph_, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        fgraph_flops_ = tm_.gemm_flops;
      }
      HIPCHK(hipEventRecord(e0_, s_));
      HIPCHK(hipGraphLaunch(fgraph_, s_));
      HIPCHK(hipEventRecord(e1_, s_));
      tm_.gemm_flops = fgraph_flops_;
    } else {
      HIPCHK(hipEventRecord(e0_, s_));
      factor_launches(true);
      HIPCHK(hipEventRecord(e1_, s_));
    }
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

  // the launches of one numeric factorization on s_ (and s2_, joined back
  // to s_ at the end); may_tune: run the GEMM autotuner when it is due
  void factor_launches(bool may_tune) {
    launch(k_err_reset, 1, 1, d_err_);
    launch(k_gather, (n_ + 255) / 256, 256, d_dpos_, d_lval_, d_adiag_, n_);
    launch(k_gather, (nfull_ + 255) / 256, 256, d_gmap_, d_lval_, d_av_, nfull_);
    if (may_tune && autotune_ && !tuned_) tune_gemm();
    nev_ = 0;
    tm_.gemm_flops = 0;
    // each level's fronts are cleared before the level: level 0 here, level
    // L + 1 on the second stream while level L works (the level's start
    // waits for that stream)
    if (lookahead_) {
      // the second stream starts after everything issued before (a solve
      // still reading the arena, the autotuning)
      HIPCHK(hipEventRecord(ev_look_, s_));
      HIPCHK(hipStreamWaitEvent(s2_, ev_look_, 0));
    }
    if (!P_.levels.empty()) {
      const Range& z = P_.levels[0].zero;
      launch(k_zero<T>, lookahead_ ? z.count : static_cast<int64_t>(P_.zero.size()), 256, d_fd_, d_zero_ + z.begin,
             d_arena_);
    }
    for (size_t L = 0; L < P_.levels.size(); ++L) {
      const LevelPlan& lp = P_.levels[L];
      join_rest();
      if (lookahead_ && L + 1 < P_.levels.size() && P_.levels[L + 1].zero.count > 0) {
        const Range& z = P_.levels[L + 1].zero;
        launch_on(s2_, k_zero<T>, z.count, 256, d_fd_, d_zero_ + z.begin, d_arena_);
        HIPCHK(hipEventRecord(ev_rest_, s2_));
        rest_pending_ = true;
      }
      // child slot 0 written into the cleared fronts, the original entries
      // added, then child slot 1: every entry is orig + child0 + child1 as in
      // the reference's order (the first sum is formed as child0 + orig, the
      // same value in IEEE arithmetic), without reading the fronts for slot 0
      launch(k_extend_add<T, false>, lp.ea[0].count, 256, d_fd_, d_ea_ + lp.ea[0].begin, d_cmap_, d_arena_);
      launch(k_scatter<T>, lp.scatter.count, 256, d_scatter_ + lp.scatter.begin, d_amap_, d_lval_, d_arena_);
      launch(k_extend_add<T, true>, lp.ea[1].count, 256, d_fd_, d_ea_ + lp.ea[1].begin, d_cmap_, d_arena_);
      for (size_t pi = 0; pi < lp.phases.size(); ++pi) {
        const Phase& ph = lp.phases[pi];
        // only the pivot-block phases may run beside the rest of an update
        if (ph.kind != PhaseKind::PanelDiag && ph.kind != PhaseKind::PanelRows && ph.kind != PhaseKind::Panel)
          join_rest();
        const bool rest_next = lookahead_ && pi + 1 < lp.phases.size() &&
                               lp.phases[pi + 1].kind == PhaseKind::GemmRest && lp.phases[pi + 1].r.count > 0;
        switch (ph.kind) {
          case PhaseKind::Gemm64:
          case PhaseKind::Gemm128:
            // the rest of this update reads the same finished panel and
            // writes other columns: it may start beside this launch
            if (rest_next) HIPCHK(hipEventRecord(ev_look_, s_));
            gemm(ph, s_);
            break;
          case PhaseKind::GemmRest:
            if (lookahead_ && ph.r.count > 0) {
              // on the second stream, after everything issued on the main
              // one before the lookahead part; the next panel phases go on
              // beside it
              if (pi == 0 || lp.phases[pi - 1].kind != PhaseKind::Gemm128 || lp.phases[pi - 1].r.count == 0)
                HIPCHK(hipEventRecord(ev_look_, s_));
              HIPCHK(hipStreamWaitEvent(s2_, ev_look_, 0));
              gemm(ph, s2_);
              HIPCHK(hipEventRecord(ev_rest_, s2_));
              rest_pending_ = true;
            } else {
              gemm(ph, s_);
            }
            break;
          case PhaseKind::Diag:
            launch(k_diag<T>, ph.r.count, 256, d_fd_, d_diag_, ph.r.begin, d_arena_, d_dpiv_, d_adiag_, d_vbuf_,
                   d_err_);
            break;
          case PhaseKind::Trsm:
            launch(k_trsm<T>, ph.r.count, 256, d_fd_, d_trsm_ + ph.r.begin, d_arena_, d_vbuf_, d_dpiv_);
            break;
          case PhaseKind::Panel:
            launch(k_panel<T>, ph.r.count, 256, d_fd_, d_panel_ + ph.r.begin, d_arena_, d_dpiv_, d_adiag_, d_vbuf_,
                   d_err_);
            break;
          case PhaseKind::PanelDiag:
            if (legacy_panel_)
              launch(k_panel<T>, ph.r.count, 256, d_fd_, d_panel_ + ph.r.begin, d_arena_, d_dpiv_, d_adiag_, d_vbuf_,
                     d_err_);
            else
              launch(k_pdiag<T>, ph.r.count, 256, d_fd_, d_panel_ + ph.r.begin, d_arena_, d_dpiv_, d_adiag_, d_vbuf_,
                     d_err_);
            break;
          case PhaseKind::PanelRows:
            if (legacy_panel_)
              launch(k_panel<T>, ph.r.count, 256, d_fd_, d_panel_ + ph.r.begin, d_arena_, d_dpiv_, d_adiag_, d_vbuf_,
                     d_err_);
            else
              launch(k_prow<T>, ph.r.count, 256, d_fd_, d_panel_ + ph.r.begin, d_arena_, d_dpiv_, d_vbuf_);
            break;
        }
      }
    }
    join_rest();
    if (lookahead_) {
      // the second stream back into the first (a capture must end joined)
      HIPCHK(hipEventRecord(ev_rest_, s2_));
      HIPCHK(hipStreamWaitEvent(s_, ev_rest_, 0));
    }
    // explicit inverses of the super-blocks for the solves
    if (!legacy_solve_ && !chain_solve_)
      launch(k_sbinv<T>, nsbinv_, 256, d_fd_, d_sbinv_, d_arena_, d_vbuf_, d_btask_,
             d_sboff_, d_xinv_, d_xinvt_);
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
    const size_t nn = static_cast<size_t>(std::max(nseg_, 1)) * nrhs;
    std::vector<double> bn(nn), rn(nn);
    run_solve(bc, nrhs, tol, max_steps, bn.data(), rn.data(), true);
    if (!on_device) HIPCHK(hipMemcpyAsync(b, d_bc_, cc * sizeof(double), hipMemcpyDeviceToHost, s_));
    HIPCHK(hipStreamSynchronize(s_));
  }

  double solve_internal(const double* d_b, double* d_x, int nrhs, double tol, int max_steps) override {
    if (nrhs <= 0 || n_ == 0) return -1.0;
    ensure_vectors(nrhs);
    const int64_t cnt = n_ * nrhs;
    HIPCHK(hipMemcpyAsync(w_.b, d_b, cnt * sizeof(double), hipMemcpyDeviceToDevice, s_));
    const size_t nn = static_cast<size_t>(std::max(nseg_, 1)) * nrhs;
    std::vector<double> bn(nn), rn(nn);
    if (max_steps > 0) {
      refine(nrhs, tol, max_steps, bn.data(), rn.data());
    } else {
      HIPCHK(hipMemcpyAsync(w_.x, w_.b, cnt * sizeof(double), hipMemcpyDeviceToDevice, s_));
      solve_device(w_.x, nrhs);
      tm_.rel_residual = -1.0;
    }
    HIPCHK(hipMemcpyAsync(d_x, w_.x, cnt * sizeof(double), hipMemcpyDeviceToDevice, s_));
    HIPCHK(hipStreamSynchronize(s_));
    return tm_.rel_residual;
  }

  void set_fallback(SolveFallback f) override { fallback_ = std::move(f); }
  const double* device_lower_values() const override { return d_lval_; }

  // The solve of solve_device() without any device allocation: every work
  // vector is carved from the file-scope scratch array of the device that
  // holds this factor, and the kernels run eagerly (no graph is
  // instantiated).  Single problems only (no segments).
  bool solve_noalloc(double* d_b, int nrhs, double tol, int max_steps) override {
    if (nrhs <= 0 || nc_ == 0) return true;
    if (nseg_ > 0) throw std::runtime_error("the allocation-free solve serves single-problem handles only");
    if (nrhs > kNoAllocMaxRhs) {
      char m[160];
      snprintf(m, sizeof m, "allocation-free solve: %d right-hand sides, at most %d", nrhs, kNoAllocMaxRhs);
      throw std::runtime_error(m);
    }
    const size_t need = noalloc_elems(nrhs);
    if (need > kScratchElems) {
      char m[200];
      snprintf(m, sizeof m, "allocation-free solve needs %zu bytes of scratch for %d right-hand sides; the maximum is %zu",
               need * sizeof(double), nrhs, kScratchElems * sizeof(double));
      throw std::runtime_error(m);
    }
    DeviceGuard guard(dev_);
    double* base = scratch_ptr(dev_);
    const SolveWork saved = w_;
    const bool saved_eager = eager_;
    w_ = carve(base, nrhs);
    eager_ = true;
    double bn[kNoAllocMaxRhs], rn[kNoAllocMaxRhs];
    try {
      run_solve(d_b, nrhs, tol, max_steps, bn, rn, false);
      HIPCHK(hipStreamSynchronize(s_));
    } catch (...) {
      w_ = saved;
      eager_ = saved_eager;
      throw;
    }
    w_ = saved;
    eager_ = saved_eager;
    return max_steps <= 0 || tm_.rel_residual <= tol;
  }

  // caller numbering in bc (device) -> internal b, solve (and refine), back
  // into bc; when the refinement misses tol and a fallback is set (and
  // allowed), the fallback solves the internal right-hand sides again
  void run_solve(double* bc, int nrhs, double tol, int max_steps, double* bn, double* rn, bool allow_fallback) {
    const int64_t cc = nc_ * nrhs;
    tm_.fell_back = 0;
    tm_.fallback_rel = -1.0;
    HIPCHK(hipEventRecord(e0_, s_));
    if (n_ > 0) {
      launch(k_perm_in, (cc + 255) / 256, 256, d_eq_, bc, w_.b, nc_, n_, nrhs);
      if (max_steps > 0) {
        refine(nrhs, tol, max_steps, bn, rn);
        if (allow_fallback && fallback_ && !(tm_.rel_residual <= tol)) {
          HIPCHK(hipStreamSynchronize(s_));
          const double rel = fallback_(w_.b, w_.x, nrhs);
          if (rel >= 0.0) {
            tm_.fell_back = 1;
            tm_.fallback_rel = rel;
          }
        }
      } else {
        HIPCHK(hipMemcpyAsync(w_.x, w_.b, n_ * nrhs * sizeof(double), hipMemcpyDeviceToDevice, s_));
        solve_device(w_.x, nrhs);
        tm_.refine_steps = 0;
        tm_.rel_residual = -1.0;
      }
    }
    launch(k_perm_out, (cc + 255) / 256, 256, d_eq_, d_iso_, w_.x, bc, nc_, n_, nrhs);
    HIPCHK(hipEventRecord(e1_, s_));
    HIPCHK(hipEventSynchronize(e1_));
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
    launch_on(s_, kernel, blocks, threads, args...);
  }

  template <typename K, typename... A>
  void launch_on(hipStream_t st, K kernel, int64_t blocks, int threads, A... args) {
    if (blocks <= 0) return;
    hipLaunchKernelGGL(kernel, dim3(grid(blocks)), dim3(threads), 0, st, args...);
    HIPCHK(hipGetLastError());
  }

  // the main stream waits for the rest of the last outer block's update
  void join_rest() {
    if (!rest_pending_) return;
    HIPCHK(hipStreamWaitEvent(s_, ev_rest_, 0));
    rest_pending_ = false;
  }

  // a per-front vector kernel over `fronts` fronts, nrhs right-hand sides and
  // entries in chunks of VCH up to maxlen (the longest front vector of the launch)
  template <typename K, typename... A>
  void launch_vec(K kernel, int64_t fronts, int nrhs, int64_t maxlen, A... args) {
    if (fronts <= 0 || nrhs <= 0 || maxlen <= 0) return;
    const unsigned nch = static_cast<unsigned>((maxlen + VCH - 1) / VCH);
    hipLaunchKernelGGL(kernel, dim3(grid(fronts), static_cast<unsigned>(nrhs), nch), dim3(256), 0, s_, args...);
    HIPCHK(hipGetLastError());
  }

  // one launch of the Schur-update GEMM over a tile range, on stream st
  void gemm(const Phase& ph, hipStream_t st) {
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
      HIPCHK(hipEventRecord(evs_[2 * nev_], st));
    }
    if (ph.kind != PhaseKind::Gemm128 && ph.kind != PhaseKind::GemmRest)
      throw std::runtime_error("GEMM phase with 64-column tiles is not supported");
    const int64_t n64 = gemm64_[r.begin];
    if (n64 > 0) launch_on(st, k_gemm64<T>, n64, 64, d_fd_, d_gemm_ + r.begin, d_arena_, d_dpiv_);
    if (n64 < r.count) launch_gemm(gemm_pick_[r.begin], Range{r.begin + n64, r.count - n64}, st);
    tm_.gemm_flops += flops;
    if (ktime_) {
      HIPCHK(hipEventRecord(evs_[2 * nev_ + 1], st));
      ++nev_;
    }
  }

  // variant 0 = k_gemm, 1 = k_gemm_agpr, 2 = k_gemm_ilv, 3 = k_gemm_k64,
  // 4 = k_gemm_q (FP32 only)
  static constexpr int kGemmVariants = std::is_same<T, float>::value ? 5 : 4;
  void launch_gemm(int variant, const Range& r, hipStream_t st) {
    if constexpr (std::is_same<T, float>::value) {
      if (variant == 4) {
        launch_on(st, k_gemm_q, r.count, 256, d_fd_, d_gemm_ + r.begin, d_arena_, d_dpiv_);
        return;
      }
    }
    if (variant == 1)
      launch_on(st, k_gemm_agpr<T>, r.count, 256, d_fd_, d_gemm_ + r.begin, d_arena_, d_dpiv_);
    else if (variant == 2)
      launch_on(st, k_gemm_ilv<T>, r.count, 256, d_fd_, d_gemm_ + r.begin, d_arena_, d_dpiv_);
    else if (variant == 3)
      launch_on(st, k_gemm_k64<T>, r.count, 256, d_fd_, d_gemm_ + r.begin, d_arena_, d_dpiv_);
    else
      launch_on(st, k_gemm<T>, r.count, 256, d_fd_, d_gemm_ + r.begin, d_arena_, d_dpiv_);
  }

  // Autotuning of the GEMM kernel: every GEMM launch of the plan runs the
  // four variants (one warm-up, then three timed launches each, the minimum
  // kept); a variant replaces the current choice (k_gemm first) only where it
  // is at least 1% faster (RUSTYFEM_TUNE_MARGIN).  All four give identical
  // results.  The launches
  // work on the arena before the first factorization assembles anything in
  // it; the factorization clears the arena and every level assembles its
  // fronts before its own GEMMs, so
  // what they leave there is never read.
  void tune_gemm() {
    std::vector<const Phase*> ph;
    for (const LevelPlan& lp : P_.levels)
      for (const Phase& p : lp.phases)
        if ((p.kind == PhaseKind::Gemm128 || p.kind == PhaseKind::GemmRest) && p.r.count > gemm64_[p.r.begin])
          ph.push_back(&p);
    constexpr int NV = kGemmVariants;
    std::vector<hipEvent_t> ev(ph.size() * 4 * NV);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipEventRecord(e0_, s_));
    for (size_t i = 0; i < ph.size(); ++i) {
      // the launch's tiles for the 128 x 128 kernels
      const int64_t n64 = gemm64_[ph[i]->r.begin];
      const Range r{ph[i]->r.begin + n64, ph[i]->r.count - n64};
      for (int v = 0; v < NV; ++v) {
        launch_gemm(v, r, s_);
        for (int k = 0; k < 4; ++k) {
          HIPCHK(hipEventRecord(ev[4 * (NV * i + v) + k], s_));
          if (k < 3) launch_gemm(v, r, s_);
        }
      }
    }
    HIPCHK(hipEventRecord(e1_, s_));
    HIPCHK(hipStreamSynchronize(s_));
    int64_t count[NV] = {};
    for (size_t i = 0; i < ph.size(); ++i) {
      float best[NV];
      for (int v = 0; v < NV; ++v) {
        best[v] = 1e30f;
        for (int k = 0; k < 3; ++k) {
          float t = 0;
          HIPCHK(hipEventElapsedTime(&t, ev[4 * (NV * i + v) + k], ev[4 * (NV * i + v) + k + 1]));
          best[v] = std::min(best[v], t);
        }
      }
      int pick = 0;
      for (int v = 1; v < NV; ++v)
        if (best[v] < tune_margin_ * best[pick]) pick = v;
      gemm_pick_[ph[i]->r.begin] = static_cast<int8_t>(pick);
      ++count[pick];
    }
    for (auto& e : ev) HIPCHK(hipEventDestroy(e));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0_, e1_));
    tm_.tune_ms = ms;
    tm_.tuned_launches = static_cast<double>(ph.size());
    tm_.agpr_launches = static_cast<double>(count[1]);
    tm_.ilv_launches = static_cast<double>(count[2] + count[3]);
    tm_.q_launches = NV > 4 ? static_cast<double>(count[NV - 1]) : 0.0;
    tuned_ = true;
  }

  void ensure_vectors(int nrhs) {
    if (nrhs <= cap_rhs_) return;
    drop_graphs();
    for (void* p : {static_cast<void*>(d_t_), static_cast<void*>(d_x_), static_cast<void*>(d_part_),
                    static_cast<void*>(d_b_), static_cast<void*>(d_r_), static_cast<void*>(d_c_),
                    static_cast<void*>(d_red_), static_cast<void*>(d_nrm_), static_cast<void*>(d_bc_)})
      if (p) HIPCHK(hipFree(p));
    cap_rhs_ = nrhs;
    d_t_ = dalloc<double>(static_cast<size_t>(P_.tvec_elems) * nrhs, bytes_);
    if (d_ys_) HIPCHK(hipFree(d_ys_));
    d_ys_ = dalloc<double>(static_cast<size_t>(P_.tvec_elems) * nrhs, bytes_);
    d_x_ = dalloc<double>(static_cast<size_t>(n_) * nrhs, bytes_);
    d_b_ = dalloc<double>(static_cast<size_t>(n_) * nrhs, bytes_);
    d_r_ = dalloc<double>(static_cast<size_t>(n_) * nrhs, bytes_);
    d_c_ = dalloc<double>(static_cast<size_t>(n_) * nrhs, bytes_);
    d_part_ = dalloc<double>(part_elems() * nrhs, bytes_);
    d_red_ = dalloc<double>(static_cast<size_t>(kRedBlocks) * nrhs, bytes_);
    d_nrm_ = dalloc<double>(nrhs, bytes_);
    d_bc_ = dalloc<double>(static_cast<size_t>(std::max<int64_t>(nc_, 1)) * nrhs, bytes_);
    w_ = SolveWork{d_t_, d_ys_, d_part_, d_x_, d_b_, d_r_, d_c_, d_red_, d_nrm_};
  }

  size_t part_elems() const {
    return static_cast<size_t>(std::max<int64_t>(std::max(P_.max_l21, P_.max_sbpart), 1)) * BC;
  }

  // the work vectors of one solve of nrhs columns, each start 256-byte aligned
  static size_t up32(size_t e) { return (e + 31) / 32 * 32; }
  size_t noalloc_elems(int nrhs) const {
    const size_t r = static_cast<size_t>(nrhs), tv = static_cast<size_t>(P_.tvec_elems), n = static_cast<size_t>(n_);
    return 2 * up32(tv * r) + up32(part_elems() * r) + 4 * up32(n * r) + up32(kRedBlocks * r) + up32(r);
  }
  SolveWork carve(double* base, int nrhs) const {
    const size_t r = static_cast<size_t>(nrhs), tv = static_cast<size_t>(P_.tvec_elems), n = static_cast<size_t>(n_);
    SolveWork w;
    double* p = base;
    auto take = [&](size_t e) {
      double* q = p;
      p += up32(e);
      return q;
    };
    w.t = take(tv * r);
    w.ys = take(tv * r);
    w.part = take(part_elems() * r);
    w.x = take(n * r);
    w.b = take(n * r);
    w.r = take(n * r);
    w.c = take(n * r);
    w.red = take(kRedBlocks * r);
    w.nrm = take(r);
    return w;
  }

  // x = A^-1 b with FP64 refinement; b in w_.b (internal numbering), result
  // in w_.x; bn and rn are host arrays of (segments x nrhs) entries.  The
  // residuals (and the corrections solved in place of them) alternate
  // between w_.r and w_.c, so the last correction is still at hand when the
  // residual it produced is measured: a correction that raised the worst
  // residual is subtracted again, and the iterate before it is the result.
  void refine(int nrhs, double tol, int max_steps, double* bn, double* rn) {
    const int64_t cnt = n_ * nrhs;
    HIPCHK(hipMemcpyAsync(w_.x, w_.b, cnt * sizeof(double), hipMemcpyDeviceToDevice, s_));
    solve_device(w_.x, nrhs);
    const int nseg = nseg_ > 0 ? nseg_ : 1;
    auto norms = [&](const double* v, double* out) {
      if (nseg_ > 0) {
        if (!d_segsq_ || segsq_cap_ < nseg * nrhs) {
          if (d_segsq_) HIPCHK(hipFree(d_segsq_));
          segsq_cap_ = nseg * nrhs;
          d_segsq_ = dalloc<double>(segsq_cap_, bytes_);
        }
        launch(k_seg_sumsq, std::min(nseg, 4096), 256, v, n_, nrhs, d_seg_, nseg, d_segsq_);
        HIPCHK(hipMemcpyAsync(out, d_segsq_, static_cast<size_t>(nseg) * nrhs * sizeof(double), hipMemcpyDeviceToHost, s_));
        HIPCHK(hipStreamSynchronize(s_));
      } else {
        sumsq(v, nrhs, out);
      }
    };
    norms(w_.b, bn);
    int steps = 0;
    double worst = 0.0, prev = HUGE_VAL;
    double* last = nullptr;  // the last correction added to x
    for (double& h : tm_.refine_history) h = -1.0;
    tm_.refine_undone = 0;
    for (;;) {
      double* r = (steps % 2 == 0) ? w_.r : w_.c;
      // four lanes per row when the rows average fewer than 32 entries (2D
      // elements), sixteen otherwise
      const bool short_rows = n_ > 0 && nfull_ < 32 * n_;
      if (nrhs == 1 && short_rows)
        launch(k_residual<1, 4>, (n_ * 4 + 255) / 256, 256, d_rp_, d_ci_, d_av_, w_.x, w_.b, r, n_, nrhs);
      else if (nrhs == 1)
        launch(k_residual<1, 16>, (n_ * 16 + 255) / 256, 256, d_rp_, d_ci_, d_av_, w_.x, w_.b, r, n_, nrhs);
      else if (short_rows)
        launch(k_residual<8, 4>, (n_ * 4 + 255) / 256, 256, d_rp_, d_ci_, d_av_, w_.x, w_.b, r, n_, nrhs);
      else
        launch(k_residual<8, 16>, (n_ * 16 + 255) / 256, 256, d_rp_, d_ci_, d_av_, w_.x, w_.b, r, n_, nrhs);
      norms(r, rn);
      worst = 0.0;
      bool done = true;
      for (int j = 0; j < nseg * nrhs; ++j) {
        const double rel = bn[j] > 0 ? std::sqrt(rn[j] / bn[j]) : 0.0;
        worst = std::max(worst, rel);
        if (!(rel <= tol)) done = false;
      }
      if (steps < 4) tm_.refine_history[steps] = worst;
      if (last && !(worst <= prev)) {
        // the last correction raised the worst residual (or made it NaN)
        launch(k_axpy, (cnt + 255) / 256, 256, w_.x, last, cnt, -1.0);
        worst = prev;
        --steps;
        tm_.refine_undone = 1;
        break;
      }
      // stop at the target, at the step limit, or when a correction no longer
      // halves the worst residual (the FP64 rounding floor of the residual)
      if (done || steps >= max_steps || !(worst < 0.5 * prev)) break;
      prev = worst;
      solve_device(r, nrhs);
      launch(k_axpy, (cnt + 255) / 256, 256, w_.x, r, cnt, 1.0);
      last = r;
      ++steps;
    }
    tm_.refine_steps = steps;
    tm_.rel_residual = worst;
  }

  void set_segments(const std::vector<int64_t>& bounds) override {
    if (d_seg_) HIPCHK(hipFree(d_seg_));
    d_seg_ = nullptr;
    nseg_ = bounds.size() >= 2 ? static_cast<int>(bounds.size()) - 1 : 0;
    if (nseg_ > 0) d_seg_ = dupload(bounds, bytes_, s_);
    HIPCHK(hipStreamSynchronize(s_));
  }

  void sumsq(const double* v, int nrhs, double* host_out) {
    launch(k_sumsq_partial, kRedBlocks, 256, v, n_, nrhs, w_.red);
    launch(k_sumsq_final, nrhs, 64, w_.red, kRedBlocks, nrhs, w_.nrm);
    HIPCHK(hipMemcpyAsync(host_out, w_.nrm, nrhs * sizeof(double), hipMemcpyDeviceToHost, s_));
    HIPCHK(hipStreamSynchronize(s_));
  }

  // Several right-hand sides go through the kernels in chunks of 4 by
  // default, each chunk its own workgroups.  Measured on the 64^3 block with
  // 16 RHS (three solves): chunks of 1 take 195 ms, of 4 102 ms, of 8 171 ms
  // and of 16 232 ms -- a chunk of 16 needs about 100 KB of LDS per
  // workgroup (one per CU), chunks of 1 read the factor 16 times.
  void solve_eager(double* dx, int nrhs) {
    if (nrhs == 1 || rhs_block_ == 1)
      solve_levels<1>(dx, nrhs);
    else if (nrhs <= 4 || rhs_block_ != 16)
      solve_levels<4>(dx, nrhs);
    else
      solve_levels<16>(dx, nrhs);
  }

  // The solve is a fixed sequence of kernel launches on fixed buffers: it is
  // captured once per (vector, nrhs) into a HIP graph and replayed.
  void solve_device(double* dx, int nrhs) {
    if (!use_graphs_ || eager_) {
      solve_eager(dx, nrhs);
      return;
    }
    for (const GraphEntry& e : graphs_)
      if (e.x == dx && e.nrhs == nrhs) {
        HIPCHK(hipGraphLaunch(e.exec, s_));
        return;
      }
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    try {
      solve_eager(dx, nrhs);
    } catch (...) {
      hipGraph_t dummy = nullptr;
      (void)hipStreamEndCapture(s_, &dummy);
      if (dummy) (void)hipGraphDestroy(dummy);
      throw;
    }
    HIPCHK(hipStreamEndCapture(s_, &g));
    hipGraphExec_t ge = nullptr;
    HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    graphs_.push_back(GraphEntry{dx, nrhs, ge});
    HIPCHK(hipGraphLaunch(ge, s_));
  }

  void drop_graphs() {
    for (const GraphEntry& e : graphs_) (void)hipGraphExecDestroy(e.exec);
    graphs_.clear();
  }

  template <int R>
  void solve_levels(double* dx, int nrhs) {
    if (!legacy_solve_) {
      solve_levels_sb<R>(dx, nrhs);
      return;
    }
    const int64_t tv = P_.tvec_elems;
    for (int64_t L = 0; L < static_cast<int64_t>(P_.levels.size()); ++L) {
      const LevelPlan& lp = P_.levels[L];
      const int32_t* fr = d_lf_ + lp.fronts_r.begin;
      launch_vec(k_fwd_init, lp.fronts_r.count, nrhs, lvl_maxm_[L], d_fd_, fr, dx, w_.t, n_, tv, nrhs);
      for (int sl = 0; sl < 2; ++sl) {
        const Range& cr = P_.child_slot_range[sl][L];
        launch_vec(k_fwd_child, cr.count, nrhs, slot_maxu_[sl][L], d_fd_, d_cs_[sl] + cr.begin, d_cmap_, w_.t, tv,
                   nrhs);
      }
      if (!lp.solve.empty())
        launch(k_fwd_first<R>, lp.solve[0].diag.count, 256, d_fd_, d_diag_, lp.solve[0].diag.begin, d_vbuf_, w_.t,
               tv, nrhs);
      for (const SolveStep& sp : lp.solve)
        launch(k_fwd_step<T, R>, sp.rows.count, 256, d_fd_, d_srow_ + sp.rows.begin, d_arena_, d_vbuf_, w_.t, tv,
               nrhs);
      launch_vec(k_fwd_finish, lp.fronts_r.count, nrhs, lvl_maxk_[L], d_fd_, fr, w_.t, d_dpiv_, dx, n_, tv, nrhs);
    }
    for (int64_t L = static_cast<int64_t>(P_.levels.size()) - 1; L >= 0; --L) {
      const LevelPlan& lp = P_.levels[L];
      const BwdPlan& bp = lp.bwd;
      const int32_t* fr = d_lf_ + lp.fronts_r.begin;
      launch_vec(k_bwd_init, lp.fronts_r.count, nrhs, lvl_maxm_[L], d_fd_, fr, d_cmap_, dx, w_.t, n_, tv, nrhs);
      launch(k_l21_partial<T, R>, bp.l21.count, 256, d_fd_, d_l21_ + bp.l21.begin, d_arena_, w_.t, tv, nrhs, w_.part);
      launch(k_l21_combine, bp.comb.count, 256, d_fd_, d_l21c_ + bp.comb.begin, w_.part, w_.t, tv, nrhs);
      launch(k_bwd_last<R>, bp.last.count, 256, d_fd_, d_diag_, d_last_ + bp.last.begin, d_vbuf_, w_.t, tv, nrhs);
      for (const Range& st : bp.steps)
        launch(k_bwd_step<T, R>, st.count, 256, d_fd_, d_btile_ + st.begin, d_arena_, d_vbuf_, w_.t, tv, nrhs);
      launch_vec(k_bwd_finish, lp.fronts_r.count, nrhs, lvl_maxk_[L], d_fd_, fr, w_.t, dx, n_, tv, nrhs);
    }
  }

  // Chunks of right-hand sides per launch: a launch with enough workgroups
  // for the GPU takes every chunk in each workgroup (the factor is read once),
  // a small one spreads the chunks over workgroups (more parallelism at the
  // top of the tree, the factor read once per chunk).
  static constexpr int64_t kChunkAll = 512;
  // the MFMA products for more than four right-hand sides, unless
  // RUSTYFEM_NO_MRHS_MFMA=1, at every level but those the fused small-front
  // kernels can take (which keep the VALU order whether fused or not)
  bool use_mrhs(int64_t L, int nrhs) const { return mrhs_mfma_ && nrhs > 4 && !fused_fits(L, nrhs); }
  // the level's fronts fit the fused kernels for this many right-hand sides:
  // up to SMM rows with at most four, SMM_MRHS with more (beyond that the
  // MFMA products of the super-block path are faster)
  bool fused_fits(int64_t L, int nrhs) const { return small_ok_[L] && (nrhs <= 4 || lvl_maxm_[L] <= SMM_MRHS); }
  static unsigned mr_chunks(int nrhs) { return static_cast<unsigned>((nrhs + MR - 1) / MR); }
  // dynamic LDS of the fused small-front kernels at level L: the front's
  // vector (its level's longest), y or z (its most pivots) and, backward, the
  // partial sums of the row tiles; a level that would need more than
  // kSmallLdsMax bytes (16 right-hand sides per pass only) takes the
  // super-block kernels
  static constexpr size_t kSmallLdsMax = 64 * 1024;
  template <int R>
  size_t small_lds(int64_t L, bool bwd) const {
    const int64_t lm = lvl_maxm_[L], lk = lvl_maxk_[L];
    const int64_t rows = lm + lk + (bwd ? (lm + PRT - 1) / PRT * lk : 0);
    return static_cast<size_t>(rows) * (R + 1) * sizeof(double);
  }
  static unsigned rhs_chunks(int64_t blocks, unsigned nch) { return blocks >= kChunkAll ? 1u : nch; }

  // the super-block steps of one level, NW waves per workgroup
  template <int R, int NW>
  void sb_fwd_steps(const LevelPlan& lp, int nrhs, int64_t tv, unsigned nch, bool mfma) {
    const int nbk = SBW / NB;
    for (const SbFwdStep& st : lp.sfwd) {
      if (st.chain.count > 0) {
        hipLaunchKernelGGL((k_sfwd_inv<R, NW>),
                           dim3(static_cast<unsigned>(st.chain.count), nbk, rhs_chunks(st.chain.count * nbk, nch)),
                           dim3(64 * NW), 0, s_, d_fd_, d_sbt_ + st.chain.begin, d_xinv_, d_sboff_, w_.t, w_.ys, tv,
                           nrhs);
        HIPCHK(hipGetLastError());
      }
      if (st.rows.count > 0 && mfma) {
        hipLaunchKernelGGL((k_sfwd_rows_m<T>), dim3(grid(st.rows.count), rhs_chunks(st.rows.count, mr_chunks(nrhs))),
                           dim3(256), 0, s_, d_fd_, d_sbrow_ + st.rows.begin, d_arena_, w_.ys, w_.t, tv, nrhs);
        HIPCHK(hipGetLastError());
      } else if (st.rows.count > 0) {
        hipLaunchKernelGGL((k_sfwd_rows<T, R, NW>), dim3(grid(st.rows.count), rhs_chunks(st.rows.count, nch)),
                           dim3(64 * NW), 0, s_, d_fd_, d_sbrow_ + st.rows.begin, d_arena_, w_.ys, w_.t, tv, nrhs);
        HIPCHK(hipGetLastError());
      }
    }
  }

  template <int R, int NW>
  void sb_bwd_steps(const LevelPlan& lp, int nrhs, int64_t tv, unsigned nch, bool mfma) {
    const int nbk = SBW / NB;
    for (const SbBwdStep& st : lp.sbwd) {
      if (st.part.count > 0 && mfma) {
        hipLaunchKernelGGL((k_sbwd_part_m<T>),
                           dim3(grid(st.part.count), rhs_chunks(st.part.count * 4, mr_chunks(nrhs)), BC / 64), dim3(256),
                           0, s_, d_fd_, d_sbpart_ + st.part.begin, d_arena_, w_.t, tv, nrhs, w_.part);
        HIPCHK(hipGetLastError());
      } else if (st.part.count > 0) {
        hipLaunchKernelGGL((k_sbwd_part<T, R>), dim3(grid(st.part.count), rhs_chunks(st.part.count, nch)), dim3(256), 0,
                           s_, d_fd_, d_sbpart_ + st.part.begin, d_arena_, w_.t, tv, nrhs, w_.part);
        HIPCHK(hipGetLastError());
      }
      if (st.chain.count > 0) {
        hipLaunchKernelGGL((k_sbwd_comb<R, NW>), dim3(static_cast<unsigned>(st.chain.count), nbk, nch), dim3(64 * NW), 0,
                           s_, d_fd_, d_sbt_ + st.chain.begin, w_.part, w_.t, w_.ys, tv, nrhs);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL((k_sbwd_inv<R, NW>),
                           dim3(static_cast<unsigned>(st.chain.count), nbk, rhs_chunks(st.chain.count * nbk, nch)),
                           dim3(64 * NW), 0, s_, d_fd_, d_sbt_ + st.chain.begin, d_xinvt_, d_sboff_, w_.ys, w_.t, tv,
                           nrhs);
        HIPCHK(hipGetLastError());
      }
    }
  }

  template <int R>
  void solve_levels_inv(double* dx, int nrhs) {
    const int64_t tv = P_.tvec_elems;
    const unsigned nch = static_cast<unsigned>((nrhs + R - 1) / R);  // right-hand-side chunks, one per workgroup
    for (int64_t L = 0; L < static_cast<int64_t>(P_.levels.size()); ++L) {
      const LevelPlan& lp = P_.levels[L];
      const int32_t* fr = d_lf_ + lp.fronts_r.begin;
      if (small_r_[L].count > 0 && fused_fits(L, nrhs) && small_lds<R>(L, false) <= kSmallLdsMax) {
        const Range& sr = small_r_[L];
        hipLaunchKernelGGL((k_small_fwd<T, R>), dim3(grid(sr.count), rhs_chunks(sr.count, nch)), dim3(64),
                           small_lds<R>(L, false), s_, d_fd_, d_small_ + sr.begin, d_cmap_, d_arena_, d_xinv_,
                           d_sboff_, d_vbuf_, d_dpiv_, dx, w_.t, n_, tv, nrhs, static_cast<int>(lvl_maxm_[L]));
        HIPCHK(hipGetLastError());
        continue;
      }
      launch_vec(k_fwd_init, lp.fronts_r.count, nrhs, lvl_maxm_[L], d_fd_, fr, dx, w_.t, n_, tv, nrhs);
      for (int sl = 0; sl < 2; ++sl) {
        const Range& cr = P_.child_slot_range[sl][L];
        launch_vec(k_fwd_child, cr.count, nrhs, slot_maxu_[sl][L], d_fd_, d_cs_[sl] + cr.begin, d_cmap_, w_.t, tv,
                   nrhs);
      }
      if (lvl_maxk_[L] <= 128)
        sb_fwd_steps<R, 4>(lp, nrhs, tv, nch, use_mrhs(L, nrhs));
      else
        sb_fwd_steps<R, sb_waves<R>()>(lp, nrhs, tv, nch, use_mrhs(L, nrhs));
      launch_vec(k_fwd_finish_ys, lp.fronts_r.count, nrhs, lvl_maxk_[L], d_fd_, fr, w_.ys, d_dpiv_, dx, n_, tv, nrhs);
    }
    for (int64_t L = static_cast<int64_t>(P_.levels.size()) - 1; L >= 0; --L) {
      const LevelPlan& lp = P_.levels[L];
      const int32_t* fr = d_lf_ + lp.fronts_r.begin;
      if (small_r_[L].count > 0 && fused_fits(L, nrhs) && small_lds<R>(L, true) <= kSmallLdsMax) {
        const Range& sr = small_r_[L];
        hipLaunchKernelGGL((k_small_bwd<T, R>), dim3(grid(sr.count), rhs_chunks(sr.count, nch)), dim3(64),
                           small_lds<R>(L, true), s_, d_fd_, d_small_ + sr.begin, d_cmap_, d_arena_, d_xinvt_,
                           d_sboff_, d_vbuf_, dx, w_.t, n_, tv, nrhs, static_cast<int>(lvl_maxm_[L]),
                           static_cast<int>(lvl_maxk_[L]));
        HIPCHK(hipGetLastError());
        continue;
      }
      launch_vec(k_bwd_init, lp.fronts_r.count, nrhs, lvl_maxm_[L], d_fd_, fr, d_cmap_, dx, w_.t, n_, tv, nrhs);
      if (lvl_maxk_[L] <= 128)
        sb_bwd_steps<R, 4>(lp, nrhs, tv, nch, use_mrhs(L, nrhs));
      else
        sb_bwd_steps<R, sb_waves<R>()>(lp, nrhs, tv, nch, use_mrhs(L, nrhs));
      launch_vec(k_bwd_finish, lp.fronts_r.count, nrhs, lvl_maxk_[L], d_fd_, fr, w_.t, dx, n_, tv, nrhs);
    }
  }

  template <int R>
  void solve_levels_sb(double* dx, int nrhs) {
    if (!chain_solve_) {
      solve_levels_inv<R>(dx, nrhs);
      return;
    }
    const int64_t tv = P_.tvec_elems;
    for (int64_t L = 0; L < static_cast<int64_t>(P_.levels.size()); ++L) {
      const LevelPlan& lp = P_.levels[L];
      const int32_t* fr = d_lf_ + lp.fronts_r.begin;
      launch_vec(k_fwd_init, lp.fronts_r.count, nrhs, lvl_maxm_[L], d_fd_, fr, dx, w_.t, n_, tv, nrhs);
      for (int sl = 0; sl < 2; ++sl) {
        const Range& cr = P_.child_slot_range[sl][L];
        launch_vec(k_fwd_child, cr.count, nrhs, slot_maxu_[sl][L], d_fd_, d_cs_[sl] + cr.begin, d_cmap_, w_.t, tv,
                   nrhs);
      }
      for (const SbFwdStep& st : lp.sfwd) {
        launch(k_sfwd_chain<T, R>, st.chain.count, 256, d_fd_, d_sbt_ + st.chain.begin, d_arena_, d_vbuf_, d_btask_,
               w_.t, tv, nrhs);
        launch(k_sfwd_rows<T, R, sb_waves<R>()>, st.rows.count, 64 * sb_waves<R>(), d_fd_, d_sbrow_ + st.rows.begin,
               d_arena_, w_.t, w_.t, tv, nrhs);
      }
      launch_vec(k_fwd_finish, lp.fronts_r.count, nrhs, lvl_maxk_[L], d_fd_, fr, w_.t, d_dpiv_, dx, n_, tv, nrhs);
    }
    for (int64_t L = static_cast<int64_t>(P_.levels.size()) - 1; L >= 0; --L) {
      const LevelPlan& lp = P_.levels[L];
      const int32_t* fr = d_lf_ + lp.fronts_r.begin;
      launch_vec(k_bwd_init, lp.fronts_r.count, nrhs, lvl_maxm_[L], d_fd_, fr, d_cmap_, dx, w_.t, n_, tv, nrhs);
      for (const SbBwdStep& st : lp.sbwd) {
        launch(k_sbwd_part<T, R>, st.part.count, 256, d_fd_, d_sbpart_ + st.part.begin, d_arena_, w_.t, tv, nrhs,
               w_.part);
        launch(k_sbwd_chain<T, R>, st.chain.count, 256, d_fd_, d_sbt_ + st.chain.begin, d_arena_, d_vbuf_, d_btask_,
               w_.part, w_.t, tv, nrhs);
      }
      launch_vec(k_bwd_finish, lp.fronts_r.count, nrhs, lvl_maxk_[L], d_fd_, fr, w_.t, dx, n_, tv, nrhs);
    }
  }

  static constexpr int kRedBlocks = 1024;
  Plan P_;
  hipStream_t s_ = nullptr;
  // the rest of an outer block's update (PhaseKind::GemmRest) runs on s2_
  // beside the next outer block's panel on s_; ev_look_ orders it after the
  // work before it, ev_rest_ orders the next non-panel phase after it
  hipStream_t s2_ = nullptr;
  hipEvent_t ev_look_ = nullptr, ev_rest_ = nullptr;
  bool lookahead_ = true, rest_pending_ = false;
  float tune_margin_ = 0.99f;  // a GEMM variant replaces the current one when its time is below this factor
  hipEvent_t e0_ = nullptr, e1_ = nullptr;
  int64_t n_ = 0, nlower_ = 0, nfull_ = 0;
  size_t bytes_ = 0;
  Timings tm_;
  struct GraphEntry {
    double* x;
    int nrhs;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs_;
  bool use_graphs_ = true;
  bool factor_graph_ = true;           // capture and replay the refactorization
  hipGraphExec_t fgraph_ = nullptr;
  double fgraph_flops_ = 0;
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
  PanelTile* d_panel_ = nullptr;
  SolveTile* d_srow_ = nullptr;
  BwdTile* d_btile_ = nullptr;
  L21Tile* d_l21_ = nullptr;
  L21Comb* d_l21c_ = nullptr;
  SbTask* d_sbt_ = nullptr;
  SbRowTile* d_sbrow_ = nullptr;
  SbPartTile* d_sbpart_ = nullptr;
  int32_t* d_btask_ = nullptr;
  bool legacy_solve_ = false;
  int64_t* d_seg_ = nullptr;
  double* d_segsq_ = nullptr;
  int nseg_ = 0, segsq_cap_ = 0;
  bool legacy_panel_ = false;
  bool autotune_ = true, tuned_ = false;
  std::vector<int8_t> gemm_pick_;  // per GEMM launch (by its first tile): the launch_gemm variant
  std::vector<int32_t> gemm64_;    // per GEMM launch (by its first tile): its leading tiles for the 64 x 64 kernel
  bool chain_solve_ = false;
  std::vector<int64_t> lvl_maxm_, lvl_maxk_, slot_maxu_[2];
  SmallFront* d_small_ = nullptr;
  int64_t nsbinv_ = 0;  // super-block inverse tasks launched (d_sbinv_)
  std::vector<bool> small_ok_;  // per level: its fronts fit the fused small-front kernels
  bool mrhs_mfma_ = true;
  std::vector<Range> small_r_;  // per level: its fronts in d_small_ (count 0: not fused)
  int rhs_block_ = 4;  // right-hand sides per pass of the solve kernels when nrhs > 1 (1, 4 or 16)
  SbInvTask* d_sbinv_ = nullptr;
  int64_t* d_sboff_ = nullptr;
  double* d_xinv_ = nullptr;
  double* d_xinvt_ = nullptr;
  double* d_ys_ = nullptr;
  int32_t* d_last_ = nullptr;
  int32_t* d_lf_ = nullptr;
  int32_t* d_cs_[2] = {nullptr, nullptr};
  T* d_arena_ = nullptr;
  double *d_dpiv_ = nullptr, *d_adiag_ = nullptr, *d_lval_ = nullptr, *d_vbuf_ = nullptr;

  unsigned long long* d_err_ = nullptr;
  int64_t* d_dpos_ = nullptr;
  int64_t* d_rp_ = nullptr;
  int32_t* d_ci_ = nullptr;
  int64_t* d_gmap_ = nullptr;
  double* d_av_ = nullptr;
  double *d_t_ = nullptr, *d_x_ = nullptr, *d_part_ = nullptr, *d_b_ = nullptr, *d_r_ = nullptr, *d_c_ = nullptr,
         *d_red_ = nullptr, *d_nrm_ = nullptr;
  // the work vectors the solve kernels use: the buffers above, or a carve of
  // the scratch array during solve_noalloc()
  SolveWork w_;
  bool eager_ = false;  // solve_device() without graphs (set by solve_noalloc)
  SolveFallback fallback_;  // solves again when the refinement misses its target
  int dev_ = 0;         // the device that holds this factor
  double* d_bc_ = nullptr;
  int64_t* d_eq_ = nullptr;
  double* d_iso_ = nullptr;
  int64_t nc_ = 0;
};

}  // namespace

void copy_to_host(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

void copy_to_device(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

std::unique_ptr<GpuFactor> make_gpu_factor(Plan plan, const std::vector<int64_t>& colptr,
                                           const std::vector<int64_t>& rowidx, Precision prec) {
  if (prec == Precision::MIXED) return std::make_unique<Impl<float>>(std::move(plan), colptr, rowidx);
  return std::make_unique<Impl<double>>(std::move(plan), colptr, rowidx);
}

}  // namespace rf
