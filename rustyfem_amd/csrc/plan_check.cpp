// Static hazard check of a factorization plan: inside one launch, no tile may
// write an arena element that another tile of the same launch reads or
// writes; and the rest of an outer block's update (a GemmRest launch, which
// runs on a second stream) may not write what the launches running beside it
// (the lookahead part of the same update, the next panel launches) read or
// write, nor read what they write.  Footprints are
// rectangles (front, rows, columns) of the column-major fronts; the
// extend-add footprints are sets of parent entries reached through the
// child->parent row map.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "mf.hpp"

namespace rf {

namespace {

struct Rect {
  int32_t front;
  int32_t r0, r1, c0, c1;  // half-open
  bool write;
  int32_t tile;
};

bool overlap(const Rect& a, const Rect& b) {
  return a.front == b.front && a.r0 < b.r1 && b.r0 < a.r1 && a.c0 < b.c1 && b.c0 < a.c1;
}

std::string describe(const char* what, const Rect& a, const Rect& b) {
  char buf[256];
  snprintf(buf, sizeof buf, "%s: tile %d %s front %d rows [%d,%d) cols [%d,%d) vs tile %d %s rows [%d,%d) cols [%d,%d)",
           what, a.tile, a.write ? "writes" : "reads", a.front, a.r0, a.r1, a.c0, a.c1, b.tile,
           b.write ? "writes" : "reads", b.r0, b.r1, b.c0, b.c1);
  return buf;
}

// pairwise check of the rectangles of one launch (grouped by front)
std::string check_rects(std::vector<Rect>& v, const char* what) {
  std::sort(v.begin(), v.end(), [](const Rect& a, const Rect& b) { return a.front < b.front; });
  size_t i = 0;
  while (i < v.size()) {
    size_t j = i;
    while (j < v.size() && v[j].front == v[i].front) ++j;
    for (size_t a = i; a < j; ++a) {
      if (!v[a].write) continue;
      for (size_t b = i; b < j; ++b)
        if (b != a && v[b].tile != v[a].tile && overlap(v[a], v[b])) return describe(what, v[a], v[b]);
    }
    i = j;
  }
  return "";
}

// rectangles of two launches that run at the same time
std::string check_cross(const std::vector<Rect>& a, const std::vector<Rect>& b, const char* what) {
  for (const Rect& x : a)
    for (const Rect& y : b)
      if ((x.write || y.write) && overlap(x, y)) return describe(what, x, y);
  return "";
}

bool is_panel(PhaseKind k) { return k == PhaseKind::Panel || k == PhaseKind::PanelDiag || k == PhaseKind::PanelRows; }

}  // namespace

// the super-block solve schedule: every index in bounds, every super-block of
// every front solved once in order, row tiles covering the rows below exactly
std::string check_solve_plan(const Plan& P) {
  char buf[256];
  for (int64_t f = 0; f < P.nf; ++f) {
    const FrontDesc& d = P.fd[f];
    for (int32_t b = 0; b < (d.k + NB - 1) / NB; ++b) {
      const int64_t q = P.btask_off[f] + b;
      if (q >= static_cast<int64_t>(P.btask.size()) || P.btask[q] < 0 ||
          P.btask[q] >= static_cast<int32_t>(P.diag.size()) || P.diag[P.btask[q]].front != f ||
          P.diag[P.btask[q]].j0 != b * NB) {
        snprintf(buf, sizeof buf, "solve plan: block %d of front %lld has no pivot-block inverse", b, static_cast<long long>(f));
        return buf;
      }
    }
  }
  for (const SbInvTask& v : P.sbinv) {
    const FrontDesc& d = P.fd[v.front];
    if (v.sb < 0 || v.sb >= static_cast<int32_t>(P.sbinv_off.size()) || v.sb != P.sb_first[v.front] + v.J0 / SBW ||
        v.J0 + v.W > d.k || v.W > SBW || v.b >= v.W || v.b % NB != 0 ||
        v.bt != P.btask_off[v.front] + (v.J0 + v.b) / NB ||
        P.sbinv_off[v.sb] + static_cast<int64_t>(v.W) * v.W > P.sbinv_elems) {
      snprintf(buf, sizeof buf, "solve plan: super-block inverse task of front %d [%d,+%d) column %d", v.front, v.J0, v.W, v.b);
      return buf;
    }
  }
  for (size_t L = 0; L < P.levels.size(); ++L) {
    const LevelPlan& lp = P.levels[L];
    std::vector<int32_t> next(P.nf, 0);  // forward: next expected J0 per front
    for (const SbFwdStep& st : lp.sfwd) {
      for (int64_t q = st.chain.begin; q < st.chain.begin + st.chain.count; ++q) {
        const SbTask& t = P.sbt[q];
        const FrontDesc& d = P.fd[t.front];
        if (t.J0 != next[t.front] || t.J1 <= t.J0 || t.J1 > d.k || t.J1 - t.J0 > SBW ||
            t.bt != P.btask_off[t.front] + t.J0 / NB || t.sb != P.sb_first[t.front] + t.J0 / SBW) {
          snprintf(buf, sizeof buf, "solve plan: level %zu forward super-block [%d,%d) of front %d out of order or bounds", L,
                   t.J0, t.J1, t.front);
          return buf;
        }
        next[t.front] = t.J1;
        int32_t rexp = t.J1;
        for (int64_t r = st.rows.begin; r < st.rows.begin + st.rows.count; ++r) {
          const SbRowTile& rt = P.sbrow[r];
          if (rt.front != t.front) continue;
          if (rt.J0 != t.J0 || rt.J1 != t.J1 || rt.r0 != rexp || rt.r0 >= d.m) {
            snprintf(buf, sizeof buf, "solve plan: level %zu forward row tile r0=%d of front %d", L, rt.r0, rt.front);
            return buf;
          }
          rexp += SRB;
        }
        if (rexp < d.m) {
          snprintf(buf, sizeof buf, "solve plan: level %zu forward rows of front %d not covered", L, t.front);
          return buf;
        }
      }
    }
    for (int32_t f : lp.fronts)
      if (next[f] != P.fd[f].k) {
        snprintf(buf, sizeof buf, "solve plan: level %zu forward super-blocks of front %d stop at %d of %d", L, f, next[f],
                 P.fd[f].k);
        return buf;
      }
    std::vector<int32_t> prev(P.nf);  // backward: expected J1 per front
    for (int32_t f : lp.fronts) prev[f] = P.fd[f].k;
    for (const SbBwdStep& st : lp.sbwd) {
      for (int64_t q = st.chain.begin; q < st.chain.begin + st.chain.count; ++q) {
        const SbTask& t = P.sbt[q];
        const FrontDesc& d = P.fd[t.front];
        if (t.J1 != prev[t.front] || t.J0 < 0 || t.J0 >= t.J1 || t.J1 - t.J0 > SBW || t.J0 % SBW != 0 ||
            t.bt != P.btask_off[t.front] + t.J0 / NB || t.sb != P.sb_first[t.front] + t.J0 / SBW) {
          snprintf(buf, sizeof buf, "solve plan: level %zu backward super-block [%d,%d) of front %d", L, t.J0, t.J1, t.front);
          return buf;
        }
        prev[t.front] = t.J0;
        const int32_t ncc = (t.J1 - t.J0 + BC - 1) / BC, nrc = (d.m - t.J1 + PRT - 1) / PRT;
        if (t.nrc != nrc || t.np != ncc * nrc || t.p0 < 0 || t.p0 + t.np > st.part.count) {
          snprintf(buf, sizeof buf, "solve plan: level %zu backward partial range of front %d", L, t.front);
          return buf;
        }
        for (int32_t cc = 0; cc < ncc; ++cc)
          for (int32_t rc = 0; rc < nrc; ++rc) {
            const SbPartTile& pt = P.sbpart[st.part.begin + t.p0 + cc * nrc + rc];
            if (pt.front != t.front || pt.r0 != t.J1 + rc * PRT || pt.r0 >= d.m || pt.c0 != t.J0 + cc * BC ||
                pt.c1 != std::min(pt.c0 + BC, t.J1)) {
              snprintf(buf, sizeof buf, "solve plan: level %zu partial tile (%d,%d) of front %d", L, cc, rc, t.front);
              return buf;
            }
          }
      }
    }
    for (int32_t f : lp.fronts)
      if (prev[f] != 0) {
        snprintf(buf, sizeof buf, "solve plan: level %zu backward super-blocks of front %d stop at %d", L, f, prev[f]);
        return buf;
      }
  }
  return "";
}

std::string check_plan(const Plan& P, int plant) {
  {
    const std::string e = check_solve_plan(P);
    if (!e.empty()) return e;
  }
  for (size_t L = 0; L < P.levels.size(); ++L) {
    const LevelPlan& lp = P.levels[L];
    // a level's clearing launch runs beside the previous level's launches:
    // it may touch only the level's own fronts, with the tile count of each
    {
      std::vector<int32_t> cols(P.nf, 0);
      for (int64_t t = lp.zero.begin; t < lp.zero.begin + lp.zero.count; ++t) {
        const ZeroTile& z = P.zero[t];
        if (z.front < 0 || z.front >= P.nf || P.level_of[z.front] != static_cast<int32_t>(L) || z.c0 % ZC != 0)
          return "level " + std::to_string(L) + " clearing tile " + std::to_string(t) + " outside the level";
        cols[z.front] += 1;
      }
      for (int32_t f : lp.fronts)
        if (cols[f] != (P.fd[f].m + ZC - 1) / ZC) return "level " + std::to_string(L) + " front not cleared exactly once";
    }
    // extend-add: each launch writes parent entries (map[x], map[y]); they must be distinct
    for (int s = 0; s < 2; ++s) {
      std::vector<std::pair<int64_t, int32_t>> w;  // (arena offset, tile)
      std::vector<int32_t> kids;                   // children of the launch, for the coverage count
      for (int64_t t = lp.ea[s].begin; t < lp.ea[s].begin + lp.ea[s].count; ++t) {
        const EaTile& e = P.ea[t];
        kids.push_back(e.child);
        // planted: the launch's first tall tile is left out (a gap in its block)
        if (plant == 3 && e.tall && (t == lp.ea[s].begin || !P.ea[t - 1].tall)) continue;
        const FrontDesc& c = P.fd[e.child];
        const FrontDesc& p = P.fd[c.parent];
        const int32_t* cm = P.cmap.data() + c.cmap_off;
        if (e.x0 < 0 || e.y0 < 0 || e.x0 >= c.u || e.y0 >= c.u || (e.tall != 0 && e.tall != 1) ||
            e.x0 % ea_rows(e.tall) != 0 || e.y0 % ea_cols(e.tall) != 0)
          return "extend-add level " + std::to_string(L) + " tile " + std::to_string(t) + " outside its block";
        for (int x = e.x0; x < std::min(e.x0 + ea_rows(e.tall), c.u); ++x)
          for (int y = e.y0; y < std::min(e.y0 + ea_cols(e.tall), x + 1); ++y)
            w.push_back({p.off + cm[x] + static_cast<int64_t>(cm[y]) * p.m, static_cast<int32_t>(t)});
      }
      // every entry y <= x < u of each child's block is written
      std::sort(kids.begin(), kids.end());
      kids.erase(std::unique(kids.begin(), kids.end()), kids.end());
      int64_t want = 0;
      for (int32_t c : kids) want += static_cast<int64_t>(P.fd[c].u) * (P.fd[c].u + 1) / 2;
      if (static_cast<int64_t>(w.size()) != want)
        return "extend-add level " + std::to_string(L) + " slot " + std::to_string(s) + " does not cover its blocks";
      std::sort(w.begin(), w.end());
      for (size_t i = 1; i < w.size(); ++i)
        if (w[i].first == w[i - 1].first) {
          char buf[160];
          snprintf(buf, sizeof buf, "extend-add level %zu slot %d: tiles %d and %d write arena entry %lld", L, s,
                   w[i - 1].second, w[i].second, static_cast<long long>(w[i].first));
          return buf;
        }
    }
    // footprints of one launch
    auto rects_of = [&](const Phase& ph) {
      std::vector<Rect> v;
      if (ph.kind == PhaseKind::PanelDiag || ph.kind == PhaseKind::PanelRows) {
        for (int64_t t = ph.r.begin; t < ph.r.begin + ph.r.count; ++t) {
          const PanelTile& pt = P.panel[t];
          const int32_t j1 = pt.j0 + pt.jb, f = pt.front, m = P.fd[f].m, id = static_cast<int32_t>(t);
          const int32_t rend = std::min(pt.r0 + TRS, m);
          v.push_back({f, pt.j0, j1, pt.J0, pt.j0, false, id});  // Y
          if (ph.kind == PhaseKind::PanelDiag) {
            v.push_back({f, pt.j0, j1, pt.j0, j1, false, id});   // the pivot block (only read)
          } else {
            // planted: the first row tile of a front writes the updated rows of
            // the pivot block back over Y, which the other row tiles read
            if (plant == 1 && (t == ph.r.begin || P.panel[t - 1].front != f))
              v.push_back({f, pt.j0, j1, pt.J0, pt.j0, true, id});
            v.push_back({f, pt.r0, rend, pt.J0, pt.j0, false, id});  // X
            v.push_back({f, pt.r0, rend, pt.j0, j1, true, id});      // A_R -> L_R
          }
        }
      } else if (ph.kind == PhaseKind::Panel) {
        for (int64_t t = ph.r.begin; t < ph.r.begin + ph.r.count; ++t) {
          const PanelTile& pt = P.panel[t];
          const int32_t j1 = pt.j0 + pt.jb, f = pt.front, m = P.fd[f].m, id = static_cast<int32_t>(t);
          const int32_t rend = std::min(pt.r0 + TRS, m);
          v.push_back({f, pt.j0, j1, pt.J0, pt.j0, false, id});   // Y: rows of the block, earlier columns
          v.push_back({f, pt.j0, j1, pt.j0, j1, false, id});      // the pivot block as assembled
          if (plant == 1 && pt.r0 == pt.j0) v.push_back({f, pt.j0, j1, pt.j0, j1, true, id});  // planted: L_BB in place
          if (pt.r0 != pt.j0) {
            v.push_back({f, pt.r0, rend, pt.J0, pt.j0, false, id});  // X: the tile's rows, earlier columns
            v.push_back({f, pt.r0, rend, pt.j0, j1, false, id});     // A_R
            v.push_back({f, pt.r0, rend, pt.j0, j1, true, id});      // L_R
          }
        }
      } else if (ph.kind == PhaseKind::Gemm64 || ph.kind == PhaseKind::Gemm128 || ph.kind == PhaseKind::GemmRest) {
        const int tn = ph.kind == PhaseKind::Gemm64 ? 64 : 128;
        for (int64_t t = ph.r.begin; t < ph.r.begin + ph.r.count; ++t) {
          const GemmTile& g = P.gemm[t];
          const int32_t m = P.fd[g.front].m, id = static_cast<int32_t>(t);
          const int32_t rend = std::min(g.r0 + GT_M, m), cend = std::min(g.c0 + tn, g.cend);
          v.push_back({g.front, g.r0, rend, g.k0, g.k1, false, id});
          v.push_back({g.front, g.c0, cend, g.k0, g.k1, false, id});
          v.push_back({g.front, g.r0, rend, g.c0, cend, true, id});
          // planted: the first tile of a rest update also writes the first
          // column of the next outer block, which the panel beside it reads
          if (plant == 2 && ph.kind == PhaseKind::GemmRest && t == ph.r.begin)
            v.push_back({g.front, g.k1, m, g.k1, g.k1 + 1, true, id});
        }
      }
      return v;
    };
    for (size_t i = 0; i < lp.phases.size(); ++i) {
      const Phase& ph = lp.phases[i];
      std::vector<Rect> v = rects_of(ph);
      if (v.empty()) continue;
      const bool panel = is_panel(ph.kind);
      const std::string e = check_rects(v, panel ? "panel" : "gemm");
      if (!e.empty()) return "level " + std::to_string(L) + " " + e;
      if (ph.kind != PhaseKind::GemmRest) continue;
      const std::vector<Rect> rest = rects_of(ph);
      // beside the lookahead part of the same update, which precedes it
      if (i > 0 && lp.phases[i - 1].kind == PhaseKind::Gemm128) {
        const std::string c = check_cross(rest, rects_of(lp.phases[i - 1]), "lookahead");
        if (!c.empty()) return "level " + std::to_string(L) + " " + c;
      }
      for (size_t j = i + 1; j < lp.phases.size() && is_panel(lp.phases[j].kind); ++j) {
        const std::string c = check_cross(rest, rects_of(lp.phases[j]), "lookahead");
        if (!c.empty()) return "level " + std::to_string(L) + " " + c;
      }
    }
  }
  return "";
}

}  // namespace rf
