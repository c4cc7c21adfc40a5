// Level schedule, arena layout, assembly maps and batched task lists for the
// multifrontal factorization and solves (integer work, built once per pattern).
#include <algorithm>
#include <stdexcept>

#include "mf.hpp"

namespace rf {

namespace {
Range take(int64_t begin, int64_t end) { return Range{begin, end - begin}; }

int32_t local_index(const Front& p, int64_t eq) {
  if (eq >= p.start && eq < p.start + p.k) return static_cast<int32_t>(eq - p.start);
  auto it = std::lower_bound(p.rows.begin(), p.rows.end(), eq);
  if (it == p.rows.end() || *it != eq) throw std::runtime_error("internal: row missing in the parent front");
  return static_cast<int32_t>(p.k + (it - p.rows.begin()));
}
}  // namespace

Plan build_plan(const Symbolic& sym, const std::vector<int64_t>& colptr, const std::vector<int64_t>& rowidx) {
  Plan P;
  const auto& T = sym.fronts;
  P.n = sym.n;
  P.nf = static_cast<int64_t>(T.size());
  P.fd.resize(P.nf);
  P.level_of.assign(P.nf, 0);
  int64_t off = 0, toff = 0;
  int64_t ncmap = 0;
  for (int64_t f = 0; f < P.nf; ++f) {
    const Front& F = T[f];
    FrontDesc& d = P.fd[f];
    const int64_t u = static_cast<int64_t>(F.rows.size());
    const int64_t m = F.k + u;
    if (m > INT32_MAX / 2) throw std::runtime_error("front too large");
    d.off = off;
    d.toff = toff;
    d.m = static_cast<int32_t>(m);
    d.k = static_cast<int32_t>(F.k);
    d.u = static_cast<int32_t>(u);
    d.start = static_cast<int32_t>(F.start);
    d.parent = static_cast<int32_t>(F.parent);
    d.cmap_off = F.parent >= 0 ? static_cast<int32_t>(ncmap) : -1;
    if (F.parent >= 0) ncmap += u;
    off += m * m;
    toff += m;
    int32_t lev = 0;
    for (int64_t c : F.children) lev = std::max(lev, P.level_of[c] + 1);
    P.level_of[f] = lev;
    const double k = static_cast<double>(F.k), uu = static_cast<double>(u);
    P.flops += k * k * k / 3.0 + uu * k * k + uu * uu * k;
  }
  P.arena_elems = off;
  P.tvec_elems = toff;
  // child row -> parent local index
  P.cmap.resize(ncmap);
  for (int64_t f = 0; f < P.nf; ++f) {
    if (T[f].parent < 0) continue;
    const Front& p = T[T[f].parent];
    int32_t* out = P.cmap.data() + P.fd[f].cmap_off;
    for (size_t x = 0; x < T[f].rows.size(); ++x) out[x] = local_index(p, T[f].rows[x]);
  }
  // Lower entry -> arena offset (every Lower column belongs to exactly one front)
  P.amap.resize(rowidx.size());
  for (int64_t f = 0; f < P.nf; ++f) {
    const Front& F = T[f];
    const FrontDesc& d = P.fd[f];
    for (int64_t c = F.start; c < F.start + F.k; ++c)
      for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p) {
        const int64_t lr = local_index(F, rowidx[p]);
        P.amap[p] = d.off + lr + (c - F.start) * static_cast<int64_t>(d.m);
      }
  }
  int32_t nlev = 0;
  for (int32_t l : P.level_of) nlev = std::max(nlev, l + 1);
  P.levels.resize(nlev);
  for (int64_t f = 0; f < P.nf; ++f) P.levels[P.level_of[f]].fronts.push_back(static_cast<int32_t>(f));
  for (int s = 0; s < 2; ++s) P.child_slot_range[s].resize(nlev);
  for (int32_t L = 0; L < nlev; ++L) {
    LevelPlan& lp = P.levels[L];
    int64_t b = static_cast<int64_t>(P.level_fronts.size());
    P.level_fronts.insert(P.level_fronts.end(), lp.fronts.begin(), lp.fronts.end());
    lp.fronts_r = take(b, static_cast<int64_t>(P.level_fronts.size()));
    // zero tiles
    b = static_cast<int64_t>(P.zero.size());
    for (int32_t f : lp.fronts)
      for (int32_t c0 = 0; c0 < P.fd[f].m; c0 += ZC) P.zero.push_back({f, c0});
    lp.zero = take(b, static_cast<int64_t>(P.zero.size()));
    // scatter tiles over the Lower entries of the level's pivot columns
    b = static_cast<int64_t>(P.scatter.size());
    for (int32_t f : lp.fronts) {
      const int64_t p0 = colptr[T[f].start], p1 = colptr[T[f].start + T[f].k];
      for (int64_t p = p0; p < p1; p += 8192) P.scatter.push_back({p, std::min(p + 8192, p1)});
    }
    lp.scatter = take(b, static_cast<int64_t>(P.scatter.size()));
    // extend-add tiles, child slot 0 then slot 1 (keeps the reference's summation order)
    for (int s = 0; s < 2; ++s) {
      b = static_cast<int64_t>(P.ea.size());
      int64_t cb = static_cast<int64_t>(P.child_slot_fronts[s].size());
      for (int32_t f : lp.fronts) {
        if (static_cast<int>(T[f].children.size()) <= s) continue;
        const int32_t c = static_cast<int32_t>(T[f].children[s]);
        P.child_slot_fronts[s].push_back(c);
        const int32_t uc = P.fd[c].u;
        for (int32_t x0 = 0; x0 < uc; x0 += EA_T)
          for (int32_t y0 = 0; y0 <= x0; y0 += EA_T) P.ea.push_back({c, x0, y0, 0});
      }
      lp.ea[s] = take(b, static_cast<int64_t>(P.ea.size()));
      P.child_slot_range[s][L] = take(cb, static_cast<int64_t>(P.child_slot_fronts[s].size()));
    }
    // pivot columns: outer blocks of NBO, inner blocks of NB
    int32_t maxk = 0;
    for (int32_t f : lp.fronts) maxk = std::max(maxk, P.fd[f].k);
    auto phase = [&](PhaseKind kind, int64_t b0, int64_t b1, double flops) {
      if (b1 > b0) lp.phases.push_back(Phase{kind, take(b0, b1), flops});
    };
    for (int32_t J0 = 0; J0 < maxk; J0 += NBO) {
      for (int32_t j0 = J0; j0 < std::min(J0 + NBO, maxk); j0 += NB) {
        // (a) left-looking update of the inner block from the outer block's earlier columns
        int64_t b = static_cast<int64_t>(P.gemm.size());
        double fl = 0;
        if (j0 > J0)
          for (int32_t f : lp.fronts) {
            const FrontDesc& d = P.fd[f];
            if (d.k <= j0) continue;
            const int32_t j1 = std::min(j0 + NB, d.k);
            for (int32_t r0 = j0; r0 < d.m; r0 += GT_M) P.gemm.push_back({f, r0, j0, j1, J0, j0});
            fl += 2.0 * (d.m - j0) * static_cast<double>(j1 - j0) * (j0 - J0);
          }
        phase(PhaseKind::Gemm64, b, static_cast<int64_t>(P.gemm.size()), fl);
        // (b) diagonal blocks, (c) TRSM row tiles; the solve's row tiles alongside
        const int64_t db = static_cast<int64_t>(P.diag.size());
        const int64_t tb = static_cast<int64_t>(P.trsm.size());
        const int64_t sb = static_cast<int64_t>(P.srow.size());
        int32_t slot = 0;
        for (int32_t f : lp.fronts) {
          const FrontDesc& d = P.fd[f];
          if (d.k <= j0) continue;
          const int32_t jb = std::min(NB, d.k - j0);
          const int32_t j1 = j0 + jb;
          for (int32_t r0 = j1; r0 < d.m; r0 += TRS) P.trsm.push_back({f, j0, jb, r0, slot, 0});
          const int32_t t0 = static_cast<int32_t>(P.srow.size() - sb);
          int32_t local = 0;
          for (int32_t r0 = j1; r0 < d.m; r0 += SR) P.srow.push_back({f, j0, jb, r0, slot, local++});
          P.diag.push_back({f, j0, jb, slot, t0, local, 0, 0});
          ++slot;
        }
        phase(PhaseKind::Diag, db, static_cast<int64_t>(P.diag.size()), 0);
        phase(PhaseKind::Trsm, tb, static_cast<int64_t>(P.trsm.size()), 0);
        lp.solve.push_back(SolveStep{take(db, static_cast<int64_t>(P.diag.size())),
                                     take(sb, static_cast<int64_t>(P.srow.size()))});
        P.max_slots = std::max(P.max_slots, slot);
        P.max_partials = std::max<int64_t>(P.max_partials, static_cast<int64_t>(P.srow.size()) - sb);
      }
      // right-looking update of the remaining pivot columns [J1, k), K = [J0, J1)
      const int32_t J1 = J0 + NBO;
      int64_t b = static_cast<int64_t>(P.gemm.size());
      double fl = 0;
      for (int32_t f : lp.fronts) {
        const FrontDesc& d = P.fd[f];
        if (d.k <= J1) continue;
        for (int32_t c0 = J1; c0 < d.k; c0 += 128)
          for (int32_t r0 = J1; r0 < d.m; r0 += GT_M)
            if (c0 <= r0 + GT_M - 1) P.gemm.push_back({f, r0, c0, d.k, J0, J1});
        const double w = d.k - J1, below = d.m - d.k;
        fl += 2.0 * NBO * (w * (w + 1.0) / 2.0 + below * w);
      }
      phase(PhaseKind::Gemm128, b, static_cast<int64_t>(P.gemm.size()), fl);
    }
    // Schur update of the update matrices: lower tiles of rows/cols [k, m), K = [0, k)
    b = static_cast<int64_t>(P.gemm.size());
    double fl = 0;
    for (int32_t f : lp.fronts) {
      const FrontDesc& d = P.fd[f];
      if (d.u == 0 || d.k == 0) continue;
      fl += static_cast<double>(d.u) * (d.u + 1.0) * d.k;
      for (int32_t c0 = d.k; c0 < d.m; c0 += 128)
        for (int32_t r0 = d.k; r0 < d.m; r0 += GT_M)
          if (c0 <= r0 + GT_M - 1) P.gemm.push_back({f, r0, c0, d.m, 0, d.k});
    }
    phase(PhaseKind::Gemm128, b, static_cast<int64_t>(P.gemm.size()), fl);
  }
  return P;
}

}  // namespace rf
