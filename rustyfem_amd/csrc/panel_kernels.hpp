// Inner pivot step of the multifrontal LDL^T on gfx950: one 64-column pivot
// block B = [j0, j0 + jb) of each front, inside an outer block [J0, J0 + 256).
#pragma once
#include <hip/hip_runtime.h>

#include "mf.hpp"
#include "mfma.hpp"

namespace rf {

// LDL^T of the jb x jb pivot block (jb <= 64) and the inverse of its unit
// lower factor, blocked by 16 columns, by all four waves of the workgroup.
//   in:  A[q][i] = A_BB[i][q] (column q, row i), both triangles, zero outside
//        jb x jb
//   out: A's lower triangle holds L below the diagonal and D on it (the
//        identity beyond jb); dout[i] = D_i (0 beyond jb); V[q][i] =
//        L^-1[i][q], zero outside the lower triangle of jb x jb; the pivot
//        test (src/sparse.rs:528-531) against orig = the original diagonal of
//        lane i of wave 0, the first failing equation eq0 + i wins in *err (no
//        test when err is null)
//   T:   scratch of 4 x 256 doubles
// Block column b: wave 0 eliminates its 16 columns for every row (lane i
// owns row i, columns 16b..16b+15 in registers; step j broadcasts column j
// through LDS, one write per lane and uniform-address reads, as one step of
// a right-looking elimination); then the waves apply the block's update
// -L[:, blk] D[blk] L[:, blk]^T to the trailing lower 16 x 16 tiles with FP64
// MFMA (K = 16).  The inverse: wave 1 inverts the four diagonal 16 x 16
// blocks (lane 16a + q solves L_aa x = e_q), then the blocks below them
// follow by distance from the diagonal, V_ab = -V_aa sum_{c=b}^{a-1} L_ac
// V_cb, one tile per wave on MFMA.  Loops stay short, so the code of the
// kernels that call this stays small.
__device__ __forceinline__ void block_ldlt_blk(double (*A)[NB + 1], double (*V)[NB + 2], double* T, double* cbuf,
                                               double* dout, int jb, double orig, int64_t eq0,
                                               unsigned long long* err) {
  using M = Mfma<double>;
  constexpr int BS = 16, NBK = NB / BS;
  const int tid = threadIdx.x, lane = tid & 63, l16 = lane & 15, g = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // identity beyond jb, so that every step below is uniform; the 16-column
  // blocks past jb (nbe on) stay identity and are skipped
  const int nbe = __builtin_amdgcn_readfirstlane((jb + BS - 1) / BS);
  if (tid < NB && tid >= jb) A[tid][tid] = 1.0;
  for (int idx = tid; idx < NB * NB; idx += 256) V[idx / NB][idx % NB] = 0.0;
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NBK; ++b) {
    if (b >= nbe) continue;
    const int c0 = BS * b;
    if (wave == 0) {
      double r[BS];
#pragma unroll
      for (int q = 0; q < BS; ++q) r[q] = A[c0 + q][lane];
#pragma unroll
      for (int jj = 0; jj < BS; ++jj) {
        cbuf[lane] = r[jj];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double w[BS];
#pragma unroll
        for (int q = jj; q < BS; ++q) w[q] = cbuf[c0 + q];
        __builtin_amdgcn_sched_barrier(0);
        const double l = r[jj] / w[jj];
#pragma unroll
        for (int q = jj + 1; q < BS; ++q) r[q] = fma(-l, w[q], r[q]);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane > c0 + jj) r[jj] = l;
      }
      if (lane >= c0) {
#pragma unroll
        for (int q = 0; q < BS; ++q) A[c0 + q][lane] = r[q];
      }
    }
    __syncthreads();
    // trailing lower tiles (ti, tc), b < tc <= ti < nbe, one per wave
    const int nt = nbe - 1 - b;
    for (int tt = wave; tt < nt * (nt + 1) / 2; tt += 4) {
      int ti = b + 1, rem = tt;
      while (rem >= ti - b) {
        rem -= ti - b;
        ++ti;
      }
      const int tc = b + 1 + rem;
      // output (row 16 ti + g + 4 r, column 16 tc + l16) in register r
      M::acc_t acc;
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = A[BS * tc + l16][BS * ti + g + 4 * r];
#pragma unroll
      for (int s4 = 0; s4 < BS / 4; ++s4) {
        const int k = c0 + 4 * s4 + g;
        acc = M::op(-A[k][BS * ti + l16] * A[k][k], A[k][BS * tc + l16], acc);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) A[BS * tc + l16][BS * ti + g + 4 * r] = acc[r];
    }
    __syncthreads();
  }
  // D and the pivot test
  if (wave == 0) {
    const double dl = A[lane][lane];
    if (err != nullptr && lane < jb && (!(orig > 0.0) || !(dl > 1e-12 * orig)))
      atomicMin(err, static_cast<unsigned long long>(eq0 + lane));
    dout[lane] = lane < jb ? dl : 0.0;
  } else if (wave == 1) {
    // column q of L_aa^-1 (unit lower), a = g
    const int r0 = BS * g;
    double x[BS];
#pragma unroll
    for (int i = 0; i < BS; ++i) x[i] = i == l16 ? 1.0 : 0.0;
#pragma unroll
    for (int i = 1; i < BS; ++i)
#pragma unroll
      for (int p = 0; p < i; ++p) x[i] = fma(-A[r0 + p][r0 + i], x[p], x[i]);
#pragma unroll
    for (int i = 0; i < BS; ++i) V[r0 + l16][r0 + i] = x[i];
  }
  __syncthreads();
  for (int dist = 1; dist < nbe; ++dist) {
    const int a = dist + wave;
    if (a < nbe) {
      const int b = a - dist;
      // T = sum_{c=b}^{a-1} L_ac V_cb; register r holds T[g + 4 r][l16]
      M::acc_t t = M::acc_t{0, 0, 0, 0};
      for (int c = b; c < a; ++c)
#pragma unroll
        for (int s4 = 0; s4 < BS / 4; ++s4) {
          const int k = BS * c + 4 * s4 + g;
          t = M::op(A[k][BS * a + l16], V[BS * b + l16][k], t);
        }
      double* Tw = T + wave * BS * BS;  // Tw[j * 16 + i] = T[i][j]
#pragma unroll
      for (int r = 0; r < 4; ++r) Tw[l16 * BS + g + 4 * r] = t[r];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // V_ab = -V_aa T
      M::acc_t o = M::acc_t{0, 0, 0, 0};
#pragma unroll
      for (int s4 = 0; s4 < BS / 4; ++s4) {
        const int k = 4 * s4 + g;
        o = M::op(V[BS * a + k][BS * a + l16], Tw[l16 * BS + k], o);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) V[BS * b + l16][BS * a + g + 4 * r] = -o[r];
    }
    __syncthreads();
  }
  // zero outside jb x jb (the identity padding and the signs of its zeros)
  for (int idx = tid; idx < NB * NB; idx += 256) {
    const int q = idx / NB, i = idx % NB;
    if (q >= jb || i >= jb || i < q) V[q][i] = 0.0;
  }
}

// Columns of the left-looking update staged in LDS per pass: the update of a
// pivot block reads at most NBO - NB = 192 earlier columns of its outer
// block; one pass in FP32 (50 KB of LDS), two in FP64.
template <typename T>
struct PanelPass {
  static constexpr int PK = 96 * 8 / static_cast<int>(sizeof(T));
};

// The pass's global loads are issued together, before the LDS of the previous
// pass is released, so that one memory latency is exposed per pass.  Rows
// beyond jb (or beyond the front) are clamped to a valid row: they only reach
// update entries that are discarded.  The MFMA order over k is that of one
// k-step of 4 at a time (lane group g = lane / 16 takes k = 4 t + g at step
// t), the order of the fused panel kernel, so both give identical factors.
//
// k_pdiag (one workgroup per front): C_BB = Y D Y^T with Y = L[B, J0:j0],
// A_BB = F_BB - C_BB, then block_ldlt_blk; writes D and V = L_BB^-1 (the
// arena's pivot block stays as assembled).
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_pdiag(
    const FrontDesc* __restrict__ fd, const PanelTile* __restrict__ tiles,
                                               const T* __restrict__ arena, double* __restrict__ dpiv,
                                               const double* __restrict__ adiag, double* __restrict__ vbuf,
                                               unsigned long long* __restrict__ err) {
  using M = Mfma<double>;
  constexpr int PK = PanelPass<T>::PK;
  constexpr int NS = PK / 4;
  constexpr int YBYTES = PK * (NB + 2) * static_cast<int>(sizeof(T)) + PK * 8;
  constexpr int ABYTES = NB * (NB + 1) * 8;  // Ad, factored in place
  constexpr int TOFF = (ABYTES + 15) / 16 * 16;  // the inverse's scratch after Ad
  constexpr int VOFF = TOFF + 4 * 256 * 8;         // then Vs
  constexpr int LBYTES = VOFF + NB * (NB + 2) * 8;
  // the update's staging and the factorization's buffers share one region
  // (75 KB in all), so that two workgroups fit in a CU's LDS
  __shared__ __attribute__((aligned(16))) unsigned char smem[YBYTES > LBYTES ? YBYTES : LBYTES];
  __shared__ double cbuf[NB], dd[NB];
  double(*Vs)[NB + 2] = reinterpret_cast<double(*)[NB + 2]>(smem + VOFF);  // Vs[q][i] = L^-1[i][q]
  T(*Yk)[NB + 2] = reinterpret_cast<T(*)[NB + 2]>(smem);                            // Yk[k][i] = L[j0 + i, p0 + k]
  double* Ds = reinterpret_cast<double*>(smem + PK * (NB + 2) * sizeof(T));          // D[p0 + k]
  double(*Ad)[NB + 1] = reinterpret_cast<double(*)[NB + 1]>(smem);                  // after the update
  const PanelTile t = tiles[blockIdx.x];
  const FrontDesc d = fd[t.front];
  const int jb = t.jb, tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const T* F = arena + d.off;
  const int64_t m = d.m;
  const int i = wave * 16 + (lane & 15);
  // the original diagonal of lane i's pivot (for the test after the
  // elimination) and the block's own entries (lower triangle, mirrored)
  const double orig = lane < jb ? adiag[static_cast<int64_t>(d.start) + t.j0 + lane] : 1.0;
  // (read before the update in FP32; in FP64 after it, where the 32 registers
  // they take would push the update past two waves per SIMD)
  constexpr bool kEarly = sizeof(T) == 4;
  T fb[4][4];
  auto load_fb = [&]() {
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = bj * 16 + M::orow(lane, r);
        const int hi = min(max(i, q), jb - 1), lo = min(min(i, q), jb - 1);
        fb[bj][r] = F[t.j0 + hi + static_cast<int64_t>(t.j0 + lo) * m];
      }
  };
  if constexpr (kEarly) load_fb();
  M::acc_t cb[4];
#pragma unroll
  for (int bj = 0; bj < 4; ++bj) cb[bj] = M::acc_t{0, 0, 0, 0};
  const int li = tid & 63, kq = tid >> 6;
  const T* Fy = F + t.j0 + min(li, jb - 1);
  for (int p0 = t.J0; p0 < t.j0; p0 += PK) {
    const int kn = min(PK, t.j0 - p0);
    T ry[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (4 * s < kn) ry[s] = Fy[static_cast<int64_t>(p0 + kq + 4 * s) * m];
    const double rd = tid < kn ? dpiv[d.start + p0 + tid] : 0.0;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (4 * s < kn) Yk[kq + 4 * s][li] = ry[s];
    if (tid < kn) Ds[tid] = rd;
    __syncthreads();
    // four k-steps per group; the next group's operands are read before the
    // current group's products
    T ya[2][4][4], yb[2][4];
    double yd[2][4];
    auto rd4 = [&](int buf, int k4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kr = k4 + 4 * u + lane / 16;
        yd[buf][u] = Ds[kr];
        yb[buf][u] = Yk[kr][i];
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) ya[buf][u][bj] = Yk[kr][bj * 16 + (lane & 15)];
      }
    };
    rd4(0, 0);
#pragma unroll
    for (int gq = 0; gq < PK / 16; ++gq) {
      if (16 * gq < kn) {
        if (16 * gq + 16 < kn) rd4((gq + 1) & 1, 16 * gq + 16);
        __builtin_amdgcn_sched_barrier(0);
        const int cur = gq & 1;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double b0 = static_cast<double>(yb[cur][u]);
#pragma unroll
          for (int bj = 0; bj < 4; ++bj)
            cb[bj] = M::op(static_cast<double>(ya[cur][u][bj]) * yd[cur][u], b0, cb[bj]);
        }
      }
    }
  }
  if constexpr (!kEarly) load_fb();
  __syncthreads();
#pragma unroll
  for (int bj = 0; bj < 4; ++bj)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = bj * 16 + M::orow(lane, r);
      Ad[q][i] = (i < jb && q < jb) ? static_cast<double>(fb[bj][r]) - cb[bj][r] : 0.0;
    }
  __syncthreads();
  block_ldlt_blk(Ad, Vs, reinterpret_cast<double*>(smem + TOFF), cbuf, dd, jb, orig,
                 static_cast<int64_t>(d.start) + t.j0, err);
  __syncthreads();
  // only the jb x jb block: the buffer is zero outside it from its
  // allocation on (and every pivot step of the plan writes the same jb)
  double* vb = vbuf + static_cast<int64_t>(t.task) * NB * NB;
  for (int idx = tid; idx < NB * NB; idx += 256)
    if (idx / NB < jb && idx % NB < jb) vb[idx] = Vs[idx / NB][idx % NB];
  if (tid < jb) dpiv[d.start + t.j0 + tid] = dd[tid];
}

// k_prow (one workgroup per 64-row tile below the block): C_R = X D Y^T with
// X = L[r0:r0+64, J0:j0] (in registers, in the MFMA operand layout), then
// L_R = (F_R - C_R) L_BB^-T D^-1 on MFMA with W built from V and D.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_prow(
    const FrontDesc* __restrict__ fd, const PanelTile* __restrict__ tiles, T* __restrict__ arena,
    const double* __restrict__ dpiv, const double* __restrict__ vbuf) {
  using M = Mfma<double>;
  constexpr int PK = PanelPass<T>::PK;
  constexpr int NS = PK / 4;
  constexpr int YBYTES = PK * (NB + 2) * static_cast<int>(sizeof(T)) + PK * 8;
  constexpr int WBYTES = NB * (NB + 1) * 8;
  __shared__ __attribute__((aligned(16))) unsigned char smem[YBYTES > WBYTES ? YBYTES : WBYTES];
  __shared__ double dinv[NB];
  T(*Yk)[NB + 2] = reinterpret_cast<T(*)[NB + 2]>(smem);
  double* Ds = reinterpret_cast<double*>(smem + PK * (NB + 2) * sizeof(T));
  double(*Wd)[NB + 1] = reinterpret_cast<double(*)[NB + 1]>(smem);  // W[p][q] = L_BB^-1[q][p], after the update
  const PanelTile t = tiles[blockIdx.x];
  const FrontDesc d = fd[t.front];
  const int jb = t.jb, tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  T* F = arena + d.off;
  const int64_t m = d.m;
  // the pivot block's inverse and D (written by the previous launch) first
  const double* vb = vbuf + static_cast<int64_t>(t.task) * NB * NB;
  double wv[NB * NB / 256];
#pragma unroll
  for (int s = 0; s < NB * NB / 256; ++s) {
    const int idx = tid + 256 * s;  // zero outside the jb x jb block
    wv[s] = (idx / NB < jb && idx % NB < jb) ? vb[idx] : 0.0;
  }
  const double dv = tid < jb ? dpiv[d.start + t.j0 + tid] : 1.0;
  const int i = wave * 16 + (lane & 15);
  const int row = t.r0 + i;
  const int rowc = min(row, d.m - 1);
  // A_R entries of this lane
  T fa[4][4];
#pragma unroll
  for (int bj = 0; bj < 4; ++bj)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = min(bj * 16 + M::orow(lane, r), jb - 1);
      fa[bj][r] = F[rowc + static_cast<int64_t>(t.j0 + q) * m];
    }
  M::acc_t cr[4];
#pragma unroll
  for (int bj = 0; bj < 4; ++bj) cr[bj] = M::acc_t{0, 0, 0, 0};
  const int li = tid & 63, kq = tid >> 6, g = lane / 16;
  const T* Fy = F + t.j0 + min(li, jb - 1);
  const T* Fx = F + rowc;
  for (int p0 = t.J0; p0 < t.j0; p0 += PK) {
    const int kn = min(PK, t.j0 - p0);
    T ry[NS], rx[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (4 * s < kn) {
        ry[s] = Fy[static_cast<int64_t>(p0 + kq + 4 * s) * m];
        rx[s] = Fx[static_cast<int64_t>(p0 + 4 * s + g) * m];
      }
    const double rd = tid < kn ? dpiv[d.start + p0 + tid] : 0.0;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (4 * s < kn) Yk[kq + 4 * s][li] = ry[s];
    if (tid < kn) Ds[tid] = rd;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (4 * s < kn) {
        const int kr = 4 * s + g;
        const double dk = Ds[kr];
        const double bx = static_cast<double>(rx[s]);
#pragma unroll
        for (int bj = 0; bj < 4; ++bj)
          cr[bj] = M::op(static_cast<double>(Yk[kr][bj * 16 + (lane & 15)]) * dk, bx, cr[bj]);
      }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NB * NB / 256; ++s) {
    const int idx = tid + 256 * s, q = idx % NB, p = idx / NB;
    Wd[p][q] = q < jb ? wv[s] : 0.0;
  }
  if (tid < NB) dinv[tid] = tid < jb ? 1.0 / dv : 0.0;
  // A_R = F_R - C_R stays in the registers that hold C_R: at TRSM step s the
  // lane's B operand is column 4s + lane/16 of its row, which is register
  // (s / 4, s % 4) of cr
  double ar[16];
#pragma unroll
  for (int bj = 0; bj < 4; ++bj)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = bj * 16 + M::orow(lane, r);
      ar[bj * 4 + r] = (row < d.m && q < jb) ? static_cast<double>(fa[bj][r]) - cr[bj][r] : 0.0;
    }
  __syncthreads();
  M::acc_t o[4];
#pragma unroll
  for (int bj = 0; bj < 4; ++bj) o[bj] = M::acc_t{0, 0, 0, 0};
#pragma unroll
  for (int s4 = 0; s4 < 16; ++s4) {
    const int kr = 4 * s4 + lane / 16;
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) o[bj] = M::op(Wd[kr][bj * 16 + (lane & 15)], ar[s4], o[bj]);
  }
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

}  // namespace rf
