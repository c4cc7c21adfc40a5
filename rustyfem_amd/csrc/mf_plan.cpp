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
  P.btask_off.resize(P.nf);
  {
    int32_t o = 0;
    for (int64_t f = 0; f < P.nf; ++f) {
      P.btask_off[f] = o;
      o += (P.fd[f].k + NB - 1) / NB;
    }
    P.btask.assign(o, -1);
  }
  // super-block inverses: front f's super-blocks are sb_first[f] .. + nsb
  P.sb_first.resize(P.nf);
  {
    int32_t sb = 0;
    int64_t off = 0;
    for (int64_t f = 0; f < P.nf; ++f) {
      P.sb_first[f] = sb;
      const int32_t k = P.fd[f].k;
      for (int32_t J0 = 0; J0 < k; J0 += SBW, ++sb) {
        const int32_t W = std::min(SBW, k - J0);
        P.sbinv_off.push_back(off);
        off += static_cast<int64_t>(W) * W;
        for (int32_t b = 0; b < W; b += NB)
          P.sbinv.push_back({static_cast<int32_t>(f), J0, W, b, P.btask_off[f] + (J0 + b) / NB, sb});
      }
    }
    P.sbinv_elems = off;
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
    // clearing tiles of the level's fronts
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
        // large blocks in tall tiles (their column segments 256 rows long),
        // the rest in square ones; every entry y <= x < u in exactly one tile
        const int32_t uc = P.fd[c].u;
        const int32_t tall = uc >= EA_TALL ? 1 : 0, tx = ea_rows(tall), ty = ea_cols(tall);
        for (int32_t x0 = 0; x0 < uc; x0 += tx)
          for (int32_t y0 = 0; y0 < std::min(x0 + tx, uc); y0 += ty) P.ea.push_back({c, x0, y0, tall});
      }
      lp.ea[s] = take(b, static_cast<int64_t>(P.ea.size()));
      P.child_slot_range[s][L] = take(cb, static_cast<int64_t>(P.child_slot_fronts[s].size()));
    }
    // pivot columns: outer blocks of NBO, inner blocks of NB
    std::vector<std::pair<int64_t, int32_t>> last_tiles(P.nf, {-1, 0});  // solve tiles of a front's latest block
    int32_t maxk = 0;
    for (int32_t f : lp.fronts) maxk = std::max(maxk, P.fd[f].k);
    auto phase = [&](PhaseKind kind, int64_t b0, int64_t b1, double flops) {
      if (b1 > b0) lp.phases.push_back(Phase{kind, take(b0, b1), flops});
    };
    for (int32_t J0 = 0; J0 < maxk; J0 += NBO) {
      for (int32_t j0 = J0; j0 < std::min(J0 + NBO, maxk); j0 += NB) {
        const int64_t db = static_cast<int64_t>(P.diag.size());
        const int64_t pb = static_cast<int64_t>(P.panel.size());
        const int64_t sb = static_cast<int64_t>(P.srow.size());
        // pivot-block tiles of every front first, then the row tiles
        for (int32_t f : lp.fronts) {
          const FrontDesc& d = P.fd[f];
          if (d.k <= j0) continue;
          P.panel.push_back({f, J0, j0, std::min(NB, d.k - j0), j0, static_cast<int32_t>(P.diag.size()) +
                             static_cast<int32_t>(P.panel.size() - pb)});
        }
        const int64_t prb = static_cast<int64_t>(P.panel.size());
        for (int32_t f : lp.fronts) {
          const FrontDesc& d = P.fd[f];
          if (d.k <= j0) continue;
          const int32_t jb = std::min(NB, d.k - j0);
          const int32_t j1 = j0 + jb;
          const int32_t task = static_cast<int32_t>(P.diag.size());
          for (int32_t r0 = j1; r0 < d.m; r0 += TRS) P.panel.push_back({f, J0, j0, jb, r0, task});
          const int32_t t0 = static_cast<int32_t>(P.srow.size() - sb);
          int32_t local = 0;
          // at least one (possibly empty) tile, so that the backward step of the block runs
          for (int32_t r0 = j1; r0 < d.m || local == 0; r0 += SR) {
            P.srow.push_back({f, j0, jb, std::min(r0, d.m), task, -1, local++, 0});
            if (r0 >= d.m) break;
          }
          for (int32_t q = 0; q < local; ++q) P.srow[sb + t0 + q].ntiles = local;
          if (last_tiles[f].first >= 0)  // the previous block's tiles compute this block's solve
            for (int32_t q = 0; q < last_tiles[f].second; ++q) P.srow[last_tiles[f].first + q].next = task;
          P.btask[P.btask_off[f] + j0 / NB] = task;
          P.diag.push_back({f, j0, jb, t0, local, 0});
          last_tiles[f] = {sb + t0, local};
        }
        phase(PhaseKind::PanelDiag, pb, prb, 0);
        phase(PhaseKind::PanelRows, prb, static_cast<int64_t>(P.panel.size()), 0);
        lp.solve.push_back(SolveStep{take(db, static_cast<int64_t>(P.diag.size())),
                                     take(sb, static_cast<int64_t>(P.srow.size()))});
      }
      // right-looking update of the remaining pivot columns [J1, k), K = [J0, J1):
      // first the columns [J1, J1 + NBO) of the next outer block (its panel
      // waits for them), then the rest, which may run beside that panel
      const int32_t J1 = J0 + NBO;
      // flops of the update of columns [a, b) (rows c..m of column c)
      auto colflops = [](double a, double b, double m) { return 2.0 * NBO * ((b - a) * m - (a + b - 1.0) * (b - a) / 2.0); };
      for (int part = 0; part < 2; ++part) {
        const int64_t b = static_cast<int64_t>(P.gemm.size());
        double fl = 0;
        for (int32_t f : lp.fronts) {
          const FrontDesc& d = P.fd[f];
          if (d.k <= J1) continue;
          const int32_t cb = part == 0 ? J1 : std::min(J1 + NBO, d.k), ce = part == 0 ? std::min(J1 + NBO, d.k) : d.k;
          for (int32_t c0 = cb; c0 < ce; c0 += 128)
            for (int32_t r0 = J1; r0 < d.m; r0 += GT_M)
              if (c0 <= r0 + GT_M - 1) P.gemm.push_back({f, r0, c0, d.k, J0, J1});
          fl += colflops(cb, ce, d.m);
        }
        phase(part == 0 ? PhaseKind::Gemm128 : PhaseKind::GemmRest, b, static_cast<int64_t>(P.gemm.size()), fl);
      }
    }
    // backward solve of the level
    {
      BwdPlan& bp = lp.bwd;
      const int64_t lb = static_cast<int64_t>(P.l21.size());
      const int64_t cb = static_cast<int64_t>(P.l21c.size());
      for (int32_t f : lp.fronts) {
        const FrontDesc& d = P.fd[f];
        if (d.k == 0 || d.u == 0) continue;
        for (int32_t c0 = 0; c0 < d.k; c0 += BC) {
          const int32_t t0 = static_cast<int32_t>(P.l21.size() - lb);
          int32_t nt = 0;
          for (int32_t r0 = d.k; r0 < d.m; r0 += SR, ++nt) P.l21.push_back({f, r0, c0, 0});
          P.l21c.push_back({f, c0, t0, nt});
        }
      }
      bp.l21 = take(lb, static_cast<int64_t>(P.l21.size()));
      bp.comb = take(cb, static_cast<int64_t>(P.l21c.size()));
      P.max_l21 = std::max(P.max_l21, bp.l21.count);
      // every front's blocks in reverse, aligned at the last block
      std::vector<std::vector<int32_t>> tasks(lp.fronts.size());
      for (int64_t q = lp.solve.empty() ? 0 : lp.solve.front().diag.begin;
           !lp.solve.empty() && q < lp.solve.back().diag.begin + lp.solve.back().diag.count; ++q) {
        const int32_t f = P.diag[q].front;
        const size_t li = std::lower_bound(lp.fronts.begin(), lp.fronts.end(), f) - lp.fronts.begin();
        tasks[li].push_back(static_cast<int32_t>(q));
      }
      const int64_t kb = static_cast<int64_t>(P.bwd_last.size());
      size_t nsteps = 0;
      for (size_t li = 0; li < tasks.size(); ++li)
        if (!tasks[li].empty()) {
          P.bwd_last.push_back(tasks[li].back());
          nsteps = std::max(nsteps, tasks[li].size());
        }
      bp.last = take(kb, static_cast<int64_t>(P.bwd_last.size()));
      for (size_t st = 0; st + 1 < nsteps; ++st) {
        const int64_t tb0 = static_cast<int64_t>(P.btile.size());
        for (size_t li = 0; li < tasks.size(); ++li) {
          const auto& tl = tasks[li];
          if (tl.size() < st + 2) continue;
          const DiagTask& B = P.diag[tl[tl.size() - 1 - st]];
          const DiagTask& Bp = P.diag[tl[tl.size() - 2 - st]];
          for (int32_t c1 = B.j0; c1 > 0; c1 -= BC) {
            const int32_t c0 = std::max(0, c1 - BC);
            const bool tail = c1 == B.j0;
            P.btile.push_back({B.front, B.j0, B.jb, c0, c1, tail ? tl[tl.size() - 2 - st] : -1, tail ? Bp.jb : 0, 0});
          }
        }
        bp.steps.push_back(take(tb0, static_cast<int64_t>(P.btile.size())));
      }
    }
    // super-block solves of the level
    {
      int32_t maxnsb = 0;
      for (int32_t f : lp.fronts) maxnsb = std::max(maxnsb, (P.fd[f].k + SBW - 1) / SBW);
      for (int32_t st = 0; st < maxnsb; ++st) {
        const int64_t cb = static_cast<int64_t>(P.sbt.size()), rb = static_cast<int64_t>(P.sbrow.size());
        for (int32_t f : lp.fronts) {
          const FrontDesc& d = P.fd[f];
          const int32_t J0 = st * SBW;
          if (J0 >= d.k) continue;
          const int32_t J1 = std::min(J0 + SBW, d.k);
          P.sbt.push_back({f, J0, J1, P.btask_off[f] + J0 / NB, 0, 0, 0, P.sb_first[f] + st});
          for (int32_t r0 = J1; r0 < d.m; r0 += SRB) P.sbrow.push_back({f, J0, J1, r0});
        }
        lp.sfwd.push_back({take(cb, static_cast<int64_t>(P.sbt.size())), take(rb, static_cast<int64_t>(P.sbrow.size()))});
      }
      for (int32_t q = 0; q < maxnsb; ++q) {
        const int64_t pb = static_cast<int64_t>(P.sbpart.size());
        std::vector<SbTask> chain;
        for (int32_t f : lp.fronts) {
          const FrontDesc& d = P.fd[f];
          const int32_t nsb = (d.k + SBW - 1) / SBW;
          if (nsb <= q) continue;
          const int32_t J0 = (nsb - 1 - q) * SBW, J1 = std::min(J0 + SBW, d.k);
          const int32_t nrc = (d.m - J1 + PRT - 1) / PRT;
          const int32_t p0 = static_cast<int32_t>(P.sbpart.size() - pb);
          for (int32_t c0 = J0; c0 < J1; c0 += BC)
            for (int32_t rc = 0; rc < nrc; ++rc)
              P.sbpart.push_back({f, J1 + rc * PRT, c0, std::min(c0 + BC, J1)});
          const int32_t np = static_cast<int32_t>(P.sbpart.size() - pb) - p0;
          chain.push_back({f, J0, J1, P.btask_off[f] + J0 / NB, p0, np, nrc, P.sb_first[f] + nsb - 1 - q});
        }
        const int64_t cb = static_cast<int64_t>(P.sbt.size());
        P.sbt.insert(P.sbt.end(), chain.begin(), chain.end());
        const Range pr = take(pb, static_cast<int64_t>(P.sbpart.size()));
        P.max_sbpart = std::max(P.max_sbpart, pr.count);
        lp.sbwd.push_back({pr, take(cb, static_cast<int64_t>(P.sbt.size()))});
      }
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
