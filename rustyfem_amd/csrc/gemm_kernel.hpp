// Schur-update GEMM of the multifrontal factorization on MFMA (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include "mf.hpp"
#include "mfma.hpp"

namespace rf {

// Workgroup bid of a launch runs on XCD bid % 8 (round-robin dispatch).  The
// tile of bid is chosen so that every XCD works through a contiguous run of
// the tile list (tiles of one front column block share their B panel, which
// then stays in that XCD's L2).  A bijection on [0, nb).
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
  const int x = bid & 7, s = bid >> 3, q = nb >> 3, r = nb & 7;
  return x < r ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}

// C[r0:r0+128, c0:min(c0+128, cend)) -= (L[r0:, k0:k1] D[k0:k1]) L[c0:, k0:k1]^T
// on one 128 x 128 tile of a column-major front.
//
// Four waves, 2 x 2, each own a 64 x 64 block (4 x 4 MFMA 16 x 16 outputs).
// The MFMA row index runs over C's columns, so that lanes 0..15 of an output
// register hold 16 consecutive rows of the front (coalesced C access).
//
// K tiles of KT columns (32 in FP32, 16 in FP64) are staged in LDS with k
// contiguous per row: a lane's MFMA steps read their A, B and D operands with
// 16-byte LDS reads (four steps per read in FP32, two in FP64; lane group
// g = lane / 16 uses k = 16 h + 4 g + s at step s of chunk h, for both
// operands).  Row
// pitch KT + 16 bytes keeps the 16-lane reads on distinct banks.
//
// Loads carry no per-element masks, so that the next K tile stays in flight
// behind the current tile's products: rows beyond the front (A) or beyond
// cend (B) are clamped to the last valid row, whose finite lower-triangle
// values reach only C entries that are not stored; columns beyond k1 are
// clamped to k1 - 1 and their D is 0.  Column addresses are wave-uniform.
// The C tile is read 16 entries at a time before it is written, so that the
// reads of a group overlap.
template <typename T>
__device__ __forceinline__ void gemm_tile(const FrontDesc* __restrict__ fd, const GemmTile* __restrict__ tiles,
                                          T* __restrict__ arena, const double* __restrict__ dpiv) {
  using M = Mfma<T>;
  using V4 = typename M::acc_t;
  using L16 = typename M::lds16_t;
  constexpr int E = 16 / static_cast<int>(sizeof(T));  // elements per 16-byte LDS read
  constexpr int KT = M::KT;
  constexpr int KP = KT + 16 / static_cast<int>(sizeof(T));
  constexpr int NQ = KT / 8;
  const GemmTile t = tiles[blockIdx.x];
  const FrontDesc d = fd[t.front];
  __shared__ __attribute__((aligned(16))) T As[GT_M][KP];
  __shared__ __attribute__((aligned(16))) T Bs[GT_M][KP];
  __shared__ __attribute__((aligned(16))) T Ds[KT];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1, kg = wave >> 1;
  const int li = tid & 127;
  T* F = arena + d.off;
  const int64_t m = d.m;
  const int arow = min(t.r0 + li, d.m - 1);
  const int brow = min(t.c0 + li, t.cend - 1);
  V4 acc[4][4];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = V4{0, 0, 0, 0};
  T ra[4 * NQ], rb[4 * NQ];
  double rd = 0.0;
  // thread (li, kg) loads row li of both tiles at k = 4 (kg + 2 q) + j
  auto gload = [&](int kb) {
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = min(kb + 4 * (kg + 2 * q) + j, t.k1 - 1);
        const T* col = F + static_cast<int64_t>(c) * m;
        ra[4 * q + j] = col[arow];
        rb[4 * q + j] = col[brow];
      }
    if (tid < KT) rd = dpiv[d.start + min(kb + tid, t.k1 - 1)];
  };
  auto store = [&](int kb) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int k4 = 4 * (kg + 2 * q);
      lds_put4(&As[li][k4], ra[4 * q], ra[4 * q + 1], ra[4 * q + 2], ra[4 * q + 3]);
      lds_put4(&Bs[li][k4], rb[4 * q], rb[4 * q + 1], rb[4 * q + 2], rb[4 * q + 3]);
    }
    if (tid < KT) Ds[tid] = kb + tid < t.k1 ? static_cast<T>(rd) : T(0);
  };
  const int g = lane >> 4, l16 = lane & 15;
  gload(t.k0);
  for (int kb = t.k0; kb < t.k1; kb += KT) {
    __syncthreads();
    store(kb);
    __syncthreads();
    if (kb + KT < t.k1) gload(kb + KT);
#pragma unroll
    for (int h = 0; h < KT / 16; ++h)
#pragma unroll
      for (int p = 0; p < 4 / E; ++p) {
        const int kk = 16 * h + 4 * g + E * p;
        L16 av[4], bv[4];
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) av[bi] = *reinterpret_cast<const L16*>(&As[wm * 64 + bi * 16 + l16][kk]);
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) bv[bj] = *reinterpret_cast<const L16*>(&Bs[wn * 64 + bj * 16 + l16][kk]);
        const L16 dv = *reinterpret_cast<const L16*>(&Ds[kk]);
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) av[bi] *= dv;
#pragma unroll
        for (int s = 0; s < E; ++s)
#pragma unroll
          for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = M::op(bv[bj][s], av[bi][s], acc[bi][bj]);
      }
  }
#pragma unroll
  for (int bi = 0; bi < 4; ++bi) {
    const int row = t.r0 + wm * 64 + bi * 16 + l16;
    const int rowc = min(row, d.m - 1);
    T cv[4][4];
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int colc = min(t.c0 + wn * 64 + bj * 16 + M::orow(lane, r), t.cend - 1);
        cv[bj][r] = F[rowc + static_cast<int64_t>(colc) * m];
      }
    if (row >= d.m) continue;
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = t.c0 + wn * 64 + bj * 16 + M::orow(lane, r);
        if (col < t.cend) F[row + static_cast<int64_t>(col) * m] = cv[bj][r] - acc[bi][bj][r];
      }
  }
}

// The 64 x 64 upper-left quarter of gemm_tile's tile, with one wave: for
// update blocks of at most 64 rows and columns (the small fronts of the lower
// levels), the other three waves of gemm_tile would only compute entries it
// does not store.  The wave loads and stages the same K tiles (its lane is
// the row, all KT columns per lane) and issues the MFMA sequence of
// gemm_tile's wave 0 (wm = wn = 0), so the results are identical.
template <typename T>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_gemm64(
    const FrontDesc* __restrict__ fd, const GemmTile* __restrict__ tiles,
                                               T* __restrict__ arena, const double* __restrict__ dpiv) {
  using M = Mfma<T>;
  using V4 = typename M::acc_t;
  using L16 = typename M::lds16_t;
  constexpr int E = 16 / static_cast<int>(sizeof(T));
  constexpr int KT = M::KT;
  constexpr int KP = KT + 16 / static_cast<int>(sizeof(T));
  const GemmTile t = tiles[blockIdx.x];
  const FrontDesc d = fd[t.front];
  __shared__ __attribute__((aligned(16))) T As[64][KP];
  __shared__ __attribute__((aligned(16))) T Bs[64][KP];
  __shared__ __attribute__((aligned(16))) T Ds[KT];
  const int lane = threadIdx.x;
  T* F = arena + d.off;
  const int64_t m = d.m;
  const int arow = min(t.r0 + lane, d.m - 1);
  const int brow = min(t.c0 + lane, t.cend - 1);
  V4 acc[4][4];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = V4{0, 0, 0, 0};
  T ra[KT], rb[KT];
  double rd = 0.0;
  auto gload = [&](int kb) {
#pragma unroll
    for (int c4 = 0; c4 < KT; ++c4) {
      const T* col = F + static_cast<int64_t>(min(kb + c4, t.k1 - 1)) * m;
      ra[c4] = col[arow];
      rb[c4] = col[brow];
    }
    if (lane < KT) rd = dpiv[d.start + min(kb + lane, t.k1 - 1)];
  };
  auto store = [&](int kb) {
#pragma unroll
    for (int k4 = 0; k4 < KT; k4 += 4) {
      lds_put4(&As[lane][k4], ra[k4], ra[k4 + 1], ra[k4 + 2], ra[k4 + 3]);
      lds_put4(&Bs[lane][k4], rb[k4], rb[k4 + 1], rb[k4 + 2], rb[k4 + 3]);
    }
    if (lane < KT) Ds[lane] = kb + lane < t.k1 ? static_cast<T>(rd) : T(0);
  };
  const int g = lane >> 4, l16 = lane & 15;
  gload(t.k0);
  for (int kb = t.k0; kb < t.k1; kb += KT) {
    __syncthreads();
    store(kb);
    __syncthreads();
    if (kb + KT < t.k1) gload(kb + KT);
#pragma unroll
    for (int h = 0; h < KT / 16; ++h)
#pragma unroll
      for (int p = 0; p < 4 / E; ++p) {
        const int kk = 16 * h + 4 * g + E * p;
        L16 av[4], bv[4];
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) av[bi] = *reinterpret_cast<const L16*>(&As[bi * 16 + l16][kk]);
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) bv[bj] = *reinterpret_cast<const L16*>(&Bs[bj * 16 + l16][kk]);
        const L16 dv = *reinterpret_cast<const L16*>(&Ds[kk]);
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) av[bi] *= dv;
#pragma unroll
        for (int s = 0; s < E; ++s)
#pragma unroll
          for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = M::op(bv[bj][s], av[bi][s], acc[bi][bj]);
      }
  }
#pragma unroll
  for (int bi = 0; bi < 4; ++bi) {
    const int row = t.r0 + bi * 16 + l16;
    const int rowc = min(row, d.m - 1);
    T cv[4][4];
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int colc = min(t.c0 + bj * 16 + M::orow(lane, r), t.cend - 1);
        cv[bj][r] = F[rowc + static_cast<int64_t>(colc) * m];
      }
    if (row >= d.m) continue;
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = t.c0 + bj * 16 + M::orow(lane, r);
        if (col < t.cend) F[row + static_cast<int64_t>(col) * m] = cv[bj][r] - acc[bi][bj][r];
      }
  }
}

// gemm_tile with K tiles twice as wide (64 columns in FP32, 32 in FP64: half
// the barriers per product and twice the time for the next tile's loads to
// land) and cheaper loads spread among the products.  A column's address is
// the K tile's first column (one 64-bit base per tile) plus a 32-bit offset,
// one scalar multiply-add per column instead of a 64-bit product, and the
// lane's row is a 32-bit vector offset on it.  The next K tile's loads sit in
// the same basic block as the current tile's MFMAs (the last K tile is peeled
// off), and a scheduling pattern interleaves them with the first products.
// Every accumulator receives the same MFMA sequence as in gemm_tile (k in
// increasing order, four k per 16-byte LDS read in FP32, two in FP64): both
// give identical results.
template <typename T, bool PATTERN>
__device__ __forceinline__ void gemm_tile_ilv(const FrontDesc* __restrict__ fd, const GemmTile* __restrict__ tiles,
                                              T* __restrict__ arena, const double* __restrict__ dpiv) {
  using M = Mfma<T>;
  using V4 = typename M::acc_t;
  using L16 = typename M::lds16_t;
  constexpr int E = 16 / static_cast<int>(sizeof(T));
  constexpr int KT = 2 * M::KT;
  constexpr int KP = KT + 16 / static_cast<int>(sizeof(T));
  constexpr int NQ = KT / 8;
  const GemmTile t = tiles[blockIdx.x];
  const FrontDesc d = fd[t.front];
  __shared__ __attribute__((aligned(16))) T As[GT_M][KP];
  __shared__ __attribute__((aligned(16))) T Bs[GT_M][KP];
  __shared__ __attribute__((aligned(16))) T Ds[KT];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1, kg = wave >> 1;
  const int li = tid & 127;
  T* F = arena + d.off;
  const int64_t m = d.m;
  const int arow = min(t.r0 + li, d.m - 1);
  const int brow = min(t.c0 + li, t.cend - 1);
  // a front holds m^2 entries, so KT columns of it span far less than 4 GiB
  const uint32_t mb = static_cast<uint32_t>(m * static_cast<int64_t>(sizeof(T)));
  const uint32_t aoff = static_cast<uint32_t>(arow) * static_cast<uint32_t>(sizeof(T));
  const uint32_t boff = static_cast<uint32_t>(brow) * static_cast<uint32_t>(sizeof(T));
  const char* Fb = reinterpret_cast<const char*>(F);
  V4 acc[4][4];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = V4{0, 0, 0, 0};
  T ra[4 * NQ], rb[4 * NQ];
  double rd = 0.0;
  // thread (li, kg) loads row li of both tiles at k = 4 (kg + 2 q) + j; every
  // thread loads a D entry (the store keeps those of threads < KT)
  auto gload = [&](int kb) {
    const char* base = Fb + static_cast<int64_t>(kb) * m * static_cast<int64_t>(sizeof(T));
    const int lim = t.k1 - 1 - kb;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = min(4 * (kg + 2 * q) + j, lim);
        const char* col = base + static_cast<uint32_t>(c) * mb;
        ra[4 * q + j] = *reinterpret_cast<const T*>(col + aoff);
        rb[4 * q + j] = *reinterpret_cast<const T*>(col + boff);
      }
    rd = dpiv[d.start + min(kb + (tid & (KT - 1)), t.k1 - 1)];
  };
  auto store = [&](int kb) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int k4 = 4 * (kg + 2 * q);
      lds_put4(&As[li][k4], ra[4 * q], ra[4 * q + 1], ra[4 * q + 2], ra[4 * q + 3]);
      lds_put4(&Bs[li][k4], rb[4 * q], rb[4 * q + 1], rb[4 * q + 2], rb[4 * q + 3]);
    }
    if (tid < KT) Ds[tid] = kb + tid < t.k1 ? static_cast<T>(rd) : T(0);
  };
  const int g = lane >> 4, l16 = lane & 15;
  auto products = [&]() {
#pragma unroll
    for (int h = 0; h < KT / 16; ++h)
#pragma unroll
      for (int p = 0; p < 4 / E; ++p) {
        const int kk = 16 * h + 4 * g + E * p;
        L16 av[4], bv[4];
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) av[bi] = *reinterpret_cast<const L16*>(&As[wm * 64 + bi * 16 + l16][kk]);
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) bv[bj] = *reinterpret_cast<const L16*>(&Bs[wn * 64 + bj * 16 + l16][kk]);
        const L16 dv = *reinterpret_cast<const L16*>(&Ds[kk]);
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) av[bi] *= dv;
#pragma unroll
        for (int s = 0; s < E; ++s)
#pragma unroll
          for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = M::op(bv[bj][s], av[bi][s], acc[bi][bj]);
      }
  };
  gload(t.k0);
  int kb = t.k0;
  for (; kb + KT < t.k1; kb += KT) {
    __syncthreads();
    store(kb);
    __syncthreads();
    gload(kb + KT);
    products();
    // one load (and the two scalar instructions of its address) after each
    // of the first 8 NQ products
    if constexpr (PATTERN) {
#pragma unroll
      for (int i = 0; i < 8 * NQ; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x004, 2, 0);
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
      }
    }
  }
  __syncthreads();
  store(kb);
  __syncthreads();
  products();
#pragma unroll
  for (int bi = 0; bi < 4; ++bi) {
    const int row = t.r0 + wm * 64 + bi * 16 + l16;
    const int rowc = min(row, d.m - 1);
    T cv[4][4];
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int colc = min(t.c0 + wn * 64 + bj * 16 + M::orow(lane, r), t.cend - 1);
        cv[bj][r] = F[rowc + static_cast<int64_t>(colc) * m];
      }
    if (row >= d.m) continue;
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = t.c0 + wn * 64 + bj * 16 + M::orow(lane, r);
        if (col < t.cend) F[row + static_cast<int64_t>(col) * m] = cv[bj][r] - acc[bi][bj][r];
      }
  }
}

// Four floats at a 4-byte aligned address: one 16-byte access (gfx950 runs
// in unaligned access mode; the front leading dimension m is arbitrary).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

// FP32 Schur-update tile with 16-byte accesses: the same 128 x 128 tile and
// the same MFMA sequence per accumulator as gemm_tile<float> (k = 16 h + 4 g
// + s at step s of chunk h for lane group g, k increasing), so the results
// are identical; what differs is the data movement.
//  - Global loads: thread (rq, kq) loads rows 4 rq .. 4 rq + 3 of the columns
//    4 kq .. 4 kq + 3 of both panels, four 16-byte loads per panel (a wave
//    reads two whole 512-byte column segments per load).  A quad of rows
//    that starts at or beyond the last valid row starts at that row instead;
//    it then reads up to three entries past the column (the next column's
//    unused upper part, or the arena's padding), which reach only C rows
//    that are not stored.  Columns beyond k1 are clamped to k1 - 1 with D = 0.
//  - LDS: [k quad][row][k % 4], two buffers, one barrier per K tile: the
//    next tile's loads are in flight during the current tile's products and
//    land in the other buffer.  D multiplies the A panel when it is stored.
//    A lane's MFMA operands for four steps are one 16-byte LDS read.
//  - C: the A panel is the MFMA's first operand, so output register r of
//    lane l is row 4 (l / 16) + r, column l % 16 of the 16 x 16 block: every
//    lane reads and writes four consecutive rows of one column (16 bytes);
//    a quad that crosses the front's last row is stored entry by entry.
//  - Tile order: xcd_tile.
__device__ __forceinline__ void gemm_tile_q(const FrontDesc* __restrict__ fd, const GemmTile* __restrict__ tiles,
                                            float* __restrict__ arena, const double* __restrict__ dpiv) {
  using M = Mfma<float>;
  using V4 = M::acc_t;
  constexpr int KT = M::KT;      // 32 k per tile
  constexpr int NKQ = KT / 4;    // k quads per tile
  __shared__ __attribute__((aligned(16))) float S[2][2][NKQ][GT_M][4];  // [buffer][A, B][k quad][row][k % 4]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int rq = tid & 31, kq = tid >> 5;
  const int g = lane >> 4, l16 = lane & 15;
  const GemmTile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const FrontDesc d = fd[t.front];
  float* F = arena + d.off;
  const int64_t m = d.m;
  const int arow = min(t.r0 + 4 * rq, d.m - 1);
  const int brow = min(t.c0 + 4 * rq, t.cend - 1);
  // the K tile in flight: its 16-byte row quads of both panels and the D of
  // its columns (loaded unconditionally, D of the clamped column, so that no
  // branch waits for the loads before the products)
  f4u ra[4], rb[4];
  double rd[4];
  int kload = 0;
  auto gload = [&](int kb) {
    kload = kb;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = min(kb + 4 * kq + j, t.k1 - 1);
      const float* col = F + static_cast<int64_t>(c) * m;
      ra[j] = *reinterpret_cast<const f4u*>(col + arow);
      rb[j] = *reinterpret_cast<const f4u*>(col + brow);
      rd[j] = dpiv[d.start + c];
    }
  };
  // A thread writes its four rows in a rotated order, row 4 rq + (rr + o) % 4
  // at write rr, so that the 16-byte writes of 8 (and of 16) consecutive
  // lanes land on distinct banks (rows 4 rq + rr alone repeat every 4 lanes)
  const int rot = ((rq >> 1) + (rq >> 3)) & 3;
  auto pick = [](const f4u& v, int i) { return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3]; };
  auto store = [&](int buf) {
    float dk[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) dk[j] = kload + 4 * kq + j < t.k1 ? static_cast<float>(rd[j]) : 0.0f;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int e = (rr + rot) & 3;
      *reinterpret_cast<v4f*>(&S[buf][0][kq][4 * rq + e][0]) =
          v4f{pick(ra[0], e) * dk[0], pick(ra[1], e) * dk[1], pick(ra[2], e) * dk[2], pick(ra[3], e) * dk[3]};
      *reinterpret_cast<v4f*>(&S[buf][1][kq][4 * rq + e][0]) =
          v4f{pick(rb[0], e), pick(rb[1], e), pick(rb[2], e), pick(rb[3], e)};
    }
  };
  V4 acc[4][4];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = V4{0, 0, 0, 0};
  auto products = [&](int buf) {
#pragma unroll
    for (int h = 0; h < KT / 16; ++h) {
      const int q = 4 * h + g;
      v4f av[4], bv[4];
#pragma unroll
      for (int bi = 0; bi < 4; ++bi) av[bi] = *reinterpret_cast<const v4f*>(&S[buf][0][q][wm * 64 + bi * 16 + l16][0]);
#pragma unroll
      for (int bj = 0; bj < 4; ++bj) bv[bj] = *reinterpret_cast<const v4f*>(&S[buf][1][q][wn * 64 + bj * 16 + l16][0]);
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
          for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = M::op(av[bi][s], bv[bj][s], acc[bi][bj]);
    }
  };
  // the C tile's entries of this lane, read before the last K tile's
  // products so that their latency hides behind them
  f4u cv[4][4];
  auto cload = [&]() {
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
      const int rowc = min(t.r0 + wm * 64 + bi * 16 + 4 * g, d.m - 1);
#pragma unroll
      for (int bj = 0; bj < 4; ++bj) {
        const int colc = min(t.c0 + wn * 64 + bj * 16 + l16, t.cend - 1);
        cv[bi][bj] = *reinterpret_cast<const f4u*>(F + rowc + static_cast<int64_t>(colc) * m);
      }
    }
  };
  gload(t.k0);
  store(0);
  __syncthreads();
  // every K tile but the last, then the last with the C reads in flight
  int buf = 0;
  for (int kb = t.k0; kb + KT < t.k1; kb += KT) {
    gload(kb + KT);
    products(buf);
    store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  cload();
  __builtin_amdgcn_sched_barrier(0);
  products(buf);
#pragma unroll
  for (int bi = 0; bi < 4; ++bi) {
    const int row = t.r0 + wm * 64 + bi * 16 + 4 * g;
    if (row >= d.m) continue;
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) {
      const int col = t.c0 + wn * 64 + bj * 16 + l16;
      if (col >= t.cend) continue;
      const f4u o = cv[bi][bj] - f4u{acc[bi][bj][0], acc[bi][bj][1], acc[bi][bj][2], acc[bi][bj][3]};
      float* p = F + row + static_cast<int64_t>(col) * m;
      if (row + 3 < d.m) {
        *reinterpret_cast<f4u*>(p) = o;
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (row + r < d.m) p[r] = o[r];
      }
    }
  }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_gemm_q(
    const FrontDesc* __restrict__ fd, const GemmTile* __restrict__ tiles, float* __restrict__ arena,
    const double* __restrict__ dpiv) {
  gemm_tile_q(fd, tiles, arena, dpiv);
}

// The same tile under two register budgets: k_gemm asks for two waves per
// SIMD (accumulators in VGPRs; FP32 reaches three), k_gemm_agpr leaves the
// accumulators in AGPRs (two waves per SIMD in FP32, one in FP64).  Which is
// faster depends on the launch (tile count, K, front sizes); the solver times
// both on each launch of its plan once and keeps the faster.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_gemm(
    const FrontDesc* __restrict__ fd, const GemmTile* __restrict__ tiles, T* __restrict__ arena,
    const double* __restrict__ dpiv) {
  gemm_tile<T>(fd, tiles, arena, dpiv);
}

template <typename T>
__global__ __launch_bounds__(256) void k_gemm_agpr(const FrontDesc* __restrict__ fd, const GemmTile* __restrict__ tiles,
                                                   T* __restrict__ arena, const double* __restrict__ dpiv) {
  gemm_tile<T>(fd, tiles, arena, dpiv);
}

// gemm_tile_ilv, two waves per SIMD, with the load interleaving pattern
// (k_gemm_ilv) and without it, the loads where the compiler puts them
// (k_gemm_k64)
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_gemm_ilv(const FrontDesc* __restrict__ fd, const GemmTile* __restrict__ tiles,
                                                  T* __restrict__ arena, const double* __restrict__ dpiv) {
  gemm_tile_ilv<T, true>(fd, tiles, arena, dpiv);
}

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_gemm_k64(
    const FrontDesc* __restrict__ fd, const GemmTile* __restrict__ tiles, T* __restrict__ arena,
    const double* __restrict__ dpiv) {
  gemm_tile_ilv<T, false>(fd, tiles, arena, dpiv);
}

}  // namespace rf
