// Multifrontal LDL^T on one MI355X: plan (host, integer work) and device driver.
//
// Layout in HBM: every front f owns a dense m x m column-major block (leading
// dimension m = k + u) in one arena; its first k columns become the factor
// (L11 strictly lower, L21 below), its trailing u x u lower block is the update
// matrix its parent extend-adds.  D lives in a separate FP64 array indexed by
// equation.  Fronts are processed level by level (a front's level is one more
// than its highest child), so every launch covers all fronts of a level.
//
// Blocking of the k pivot columns of a front: outer blocks of NBO columns,
// each factored as inner blocks of NB columns (left-looking inside the outer
// block), followed by a right-looking update of the remaining pivot columns
// with K = NBO; the update matrix receives one Schur update with K = k at the
// end.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "symbolic.hpp"

namespace rf {

constexpr int NB = 64;        // inner pivot block (diagonal factor, TRSM, solve)
constexpr int NBO = 256;      // outer pivot block (right-looking updates)
constexpr int TRS = 64;       // rows per TRSM tile
constexpr int SR = 1024;      // rows per solve tile (4 per thread)
constexpr int BC = 256;       // columns per backward tile (one per thread)
constexpr int GT_M = 128;     // GEMM tile rows
constexpr int ZC = 64;        // columns per arena-clearing tile
constexpr int EA_T = 64;      // extend-add tile edge (square tiles)
constexpr int EA_TX = 256;    // rows of a tall extend-add tile (EA_T * EA_T / EA_TX columns)
constexpr int EA_TALL = 1024; // contribution blocks with at least this many rows take tall tiles
constexpr int SBW = 512;      // super-block width of the triangular solves
constexpr int SRB = 64;       // rows per tile of the forward super-block update
constexpr int PRT = 64;       // rows per tile of the backward super-block partial sums

struct FrontDesc {
  int64_t off;     // element offset of the front in the arena
  int64_t toff;    // element offset of the front's vector in the solve buffers
  int32_t m, k, u, start;
  int32_t parent;
  int32_t cmap_off;  // offset of this front's child->parent row map (length u), -1 for a root
};

struct ZeroTile { int32_t front, c0; };
struct ScatterTile { int64_t p0, p1; };
// rows [x0, x0 + rows) and columns [y0, y0 + EA_T * EA_T / rows) of a child's
// contribution block; tall = 1: rows = EA_TX, else rows = EA_T
struct EaTile { int32_t child, x0, y0, tall; };
constexpr int ea_rows(int tall) { return tall ? EA_TX : EA_T; }
constexpr int ea_cols(int tall) { return EA_T * EA_T / ea_rows(tall); }
// C[r0:r0+128, c0:min(c0+TN, cend)) -= L[r0:, k0:k1] D[k0:k1] L[c0:, k0:k1]^T
struct GemmTile { int32_t front, r0, c0, cend, k0, k1; };
// one NB-wide pivot block of one front; its index is the slot of the block's
// FP64 inverse in the persistent V buffer (used by the TRSM and the solves)
struct DiagTask { int32_t front, j0, jb, tile0, ntiles, pad; };
struct TrsmTile { int32_t front, j0, jb, r0, task, pad; };
// one inner step of one front for the rows [r0, r0 + TRS): left-looking update
// from the outer block's columns [J0, j0), LDL^T of the diagonal block and
// TRSM; the tile with r0 == j0 is the diagonal block itself and writes D and
// the block inverse (slot `task`)
struct PanelTile { int32_t front, J0, j0, jb, r0, task; };
// solve row tile below pivot block `task`; `next` is the front's following
// block (-1 if none), `local`/`ntiles` its position among the block's tiles
struct SolveTile { int32_t front, j0, jb, r0, task, next, local, ntiles; };
// backward step: columns [c0, c1) of front `front` receive -L[B, c0:c1]^T z_B
// for the solved block B = [j0, j0 + jb); when c1 == j0 the tile then solves
// the previous block (task `prev`, first column c1 - jbp)
struct BwdTile { int32_t front, j0, jb, c0, c1, prev, jbp, pad; };
// rows [r0, r0 + SR) x columns [c0, c0 + BC) of a front's L21 (partial slot = tile index)
struct L21Tile { int32_t front, r0, c0, pad; };
// sum of the L21 partials of columns [c0, c0 + BC) of a front: tiles [t0, t0 + nt) of the level
struct L21Comb { int32_t front, c0, t0, nt; };

// one super-block [J0, J1) of one front for the triangular solves: its 64-wide
// pivot blocks are solved in sequence by one workgroup; bt is the index of the
// first block's entry in Plan::btask.  Backward: the block's partial sums are
// the np tiles from slot p0 of the step, nrc row tiles per 256-column chunk.
struct SbTask { int32_t front, J0, J1, bt, p0, np, nrc, sb; };
// explicit inverse of a super-block's unit-lower factor: block column b (64
// wide) of X = L_JJ^-1, W x W column-major at Plan::sbinv_off[sb]
struct SbInvTask { int32_t front, J0, W, b, bt, sb; };
// forward: rows [r0, r0 + SRB) of the front receive -L[r, J0:J1] y[J0:J1]
struct SbRowTile { int32_t front, J0, J1, r0; };
// backward: partial sums over rows [r0, r0 + PRT) of L[r, c0:c1]^T t[r]
struct SbPartTile { int32_t front, r0, c0, c1; };
// a front of a level solved by the fused small-front kernels: its super-block
// (-1 when it has no pivots), its children by slot (-1 when absent) and the
// pivot-block task of its first pivots (whose inverse V is the super-block
// inverse when it has at most NB pivots)
struct SmallFront { int32_t front, sb, child[2], vtask; };
constexpr int SMM = 384;       // rows of the largest front the fused solves take (at most 4 RHS)
constexpr int SMM_MRHS = 192;  // with more than 4 RHS (the MFMA products take the larger fronts)
constexpr int SMK = 128;  // and its pivots (the four-wave super-block kernels' widest)

struct Range { int64_t begin = 0, count = 0; };

// Panel: the fused inner step (every tile factors the pivot block itself);
// PanelDiag / PanelRows: the same step as two launches (pivot blocks, then
// the row tiles below them)
// GemmRest: the part of an outer block's right-looking update that the next
// outer block's panel does not read (pivot columns from J1 + NBO on); it may
// run beside the panel phases that follow it, up to the next non-panel phase.
enum class PhaseKind : int32_t { Gemm64, Gemm128, Diag, Trsm, Panel, PanelDiag, PanelRows, GemmRest };

struct Phase {
  PhaseKind kind;
  Range r;
  double flops = 0;  // algorithmic flops (GEMM phases)
};

struct SolveStep {
  Range diag;   // DiagTask of one inner block for every front of the level that has it
  Range rows;   // solve row tiles (SolveTile, SR rows each, at least one per block) below those blocks
};

struct BwdPlan {
  Range l21;     // L21Tile partials of s = t[0:k] - L21^T t[k:m]
  Range comb;    // L21Comb
  Range last;    // DiagTask indices (into bwd_last) of each front's last block
  std::vector<Range> steps;  // BwdTile ranges, step s handles block nb_f - 1 - s of every front
};

struct SbFwdStep { Range chain, rows; };
struct SbBwdStep { Range part, chain; };

struct LevelPlan {
  std::vector<int32_t> fronts;
  Range scatter;   // ScatterTile
  Range ea[2];     // EaTile per child slot
  std::vector<Phase> phases;       // factorization, in launch order
  std::vector<SolveStep> solve;    // inner blocks in pivot order (forward)
  BwdPlan bwd;                     // backward
  std::vector<SbFwdStep> sfwd;     // super-block forward steps
  std::vector<SbBwdStep> sbwd;     // super-block backward steps (last super-block first)
  Range fronts_r;  // range in the level-ordered front list (solve kernels)
  Range zero;      // ZeroTile: the lower triangles of the level's fronts
};

struct Plan {
  int64_t n = 0, nf = 0;
  std::vector<FrontDesc> fd;
  std::vector<int32_t> level_of;
  std::vector<LevelPlan> levels;
  std::vector<int32_t> level_fronts;  // fronts grouped by level (solve kernels)
  std::vector<int32_t> cmap;          // child row -> parent local index
  std::vector<int64_t> amap;          // Lower entry -> arena offset
  std::vector<ZeroTile> zero;          // lower triangles of all fronts, ZC columns per tile, level by level
  std::vector<ScatterTile> scatter;
  std::vector<EaTile> ea;
  std::vector<GemmTile> gemm;
  std::vector<DiagTask> diag;
  std::vector<TrsmTile> trsm;   // TRS-row tiles of the factorization
  std::vector<PanelTile> panel; // fused inner-step tiles of the factorization
  std::vector<SolveTile> srow;  // SR-row tiles of the forward solve
  std::vector<BwdTile> btile;   // column tiles of the backward solve
  std::vector<L21Tile> l21;
  std::vector<L21Comb> l21c;
  std::vector<int32_t> bwd_last;  // DiagTask index of each front's last block, per level
  std::vector<SbTask> sbt;        // super-block tasks (forward, then backward, per level)
  std::vector<SbRowTile> sbrow;
  std::vector<SbPartTile> sbpart;
  std::vector<int32_t> btask;     // DiagTask of pivot block b of front f at btask[btask_off[f] + b]
  std::vector<int32_t> btask_off;
  int64_t max_sbpart = 0;         // most partial tiles in one backward step
  std::vector<SbInvTask> sbinv;   // all block columns of all super-block inverses
  std::vector<int64_t> sbinv_off; // element offset of super-block sb's inverse
  std::vector<int32_t> sb_first;  // index of front f's first super-block
  int64_t sbinv_elems = 0;
  std::vector<int32_t> child_slot_fronts[2];  // fronts that are child slot s (per level ranges below)
  std::vector<Range> child_slot_range[2];     // per level
  int64_t arena_elems = 0;
  int64_t tvec_elems = 0;  // sum of m
  int64_t max_l21 = 0;       // most L21 partial tiles in one level
  double flops = 0.0;        // sum k^3/3 + u k^2 + u^2 k
};

// Builds the plan from the symbolic analysis and the Lower pattern (internal numbering).
Plan build_plan(const Symbolic& sym, const std::vector<int64_t>& colptr, const std::vector<int64_t>& rowidx);

// Static hazard check: "" if no tile of a launch writes what another tile of
// the same launch reads or writes, else a description of the first conflict.
// plant = 1 adds an in-place write of each pivot block by its diagonal tile
// (a known race) so that a test can see the check fire.
std::string check_plan(const Plan& P, int plant = 0);

}  // namespace rf
