// Host setup of the smoothed-aggregation AMG hierarchy (restates
// src/solver/iterative/amg.rs and the CSR kernels of csr.rs).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <stdexcept>
#include <thread>

#include "amg.hpp"

namespace rf {

namespace {

int host_threads() {
  const unsigned t = std::thread::hardware_concurrency();
  return t ? static_cast<int>(std::min(t, 64u)) : 1;
}

}  // namespace

Csr csr_from_lower(int64_t n, const int64_t* colptr, const int64_t* rowidx, const double* val) {
  Csr m;
  m.nrows = m.ncols = n;
  m.ptr.assign(n + 1, 0);
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p) {
      const int64_t r = rowidx[p];
      ++m.ptr[r + 1];
      if (r != c) ++m.ptr[c + 1];
    }
  for (int64_t i = 0; i < n; ++i) m.ptr[i + 1] += m.ptr[i];
  m.idx.resize(m.ptr[n]);
  m.val.resize(m.ptr[n]);
  std::vector<int64_t> pos(m.ptr.begin(), m.ptr.end() - 1);
  // lower entries first: a column-wise sweep fills every row in ascending column order
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p) {
      const int64_t r = rowidx[p];
      m.idx[pos[r]] = static_cast<int32_t>(c);
      m.val[pos[r]] = val[p];
      ++pos[r];
    }
  // then the mirrored strictly-lower entries, again ascending
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p) {
      const int64_t r = rowidx[p];
      if (r == c) continue;
      m.idx[pos[c]] = static_cast<int32_t>(r);
      m.val[pos[c]] = val[p];
      ++pos[c];
    }
  return m;
}

std::vector<double> csr_diag(const Csr& a) {
  std::vector<double> d(a.nrows, 0.0);
  for (int64_t i = 0; i < a.nrows; ++i) {
    const auto b = a.idx.begin() + a.ptr[i], e = a.idx.begin() + a.ptr[i + 1];
    const auto it = std::lower_bound(b, e, static_cast<int32_t>(i));
    if (it != e && *it == i) d[i] = a.val[it - a.idx.begin()];
  }
  return d;
}

Csr csr_transpose(const Csr& a) {
  Csr t;
  t.nrows = a.ncols;
  t.ncols = a.nrows;
  t.ptr.assign(a.ncols + 1, 0);
  for (int32_t c : a.idx) ++t.ptr[c + 1];
  for (int64_t i = 0; i < a.ncols; ++i) t.ptr[i + 1] += t.ptr[i];
  t.idx.resize(a.nnz());
  t.val.resize(a.nnz());
  std::vector<int64_t> pos(t.ptr.begin(), t.ptr.end() - 1);
  for (int64_t r = 0; r < a.nrows; ++r)
    for (int64_t p = a.ptr[r]; p < a.ptr[r + 1]; ++p) {
      const int32_t c = a.idx[p];
      t.idx[pos[c]] = static_cast<int32_t>(r);
      t.val[pos[c]] = a.val[p];
      ++pos[c];
    }
  return t;
}

Csr csr_mul(const Csr& a, const Csr& b) {
  if (a.ncols != b.nrows) throw std::runtime_error("internal: csr_mul shape mismatch");
  const int64_t n = a.nrows;
  const int t = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(host_threads(), n / 2000)));
  const int64_t chunk = std::max<int64_t>(1, (n + t - 1) / t);
  struct Part {
    std::vector<int64_t> cnt;
    std::vector<int32_t> idx;
    std::vector<double> val;
  };
  std::vector<Part> parts(t);
  auto work = [&](int c) {
    const int64_t r0 = std::min<int64_t>(n, c * chunk), r1 = std::min<int64_t>(n, r0 + chunk);
    Part& P = parts[c];
    std::vector<int64_t> mark(b.ncols, -1);
    std::vector<double> acc(b.ncols, 0.0);
    std::vector<int32_t> cols;
    for (int64_t i = r0; i < r1; ++i) {
      cols.clear();
      for (int64_t p = a.ptr[i]; p < a.ptr[i + 1]; ++p) {
        const int32_t k = a.idx[p];
        const double av = a.val[p];
        for (int64_t q = b.ptr[k]; q < b.ptr[k + 1]; ++q) {
          const int32_t cc = b.idx[q];
          if (mark[cc] != i) {
            mark[cc] = i;
            acc[cc] = 0.0;
            cols.push_back(cc);
          }
          acc[cc] += av * b.val[q];
        }
      }
      std::sort(cols.begin(), cols.end());
      for (int32_t cc : cols) {
        P.idx.push_back(cc);
        P.val.push_back(acc[cc]);
      }
      P.cnt.push_back(static_cast<int64_t>(cols.size()));
    }
  };
  if (t == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int c = 0; c < t; ++c) th.emplace_back(work, c);
    for (auto& x : th) x.join();
  }
  Csr m;
  m.nrows = n;
  m.ncols = b.ncols;
  m.ptr.assign(1, 0);
  m.ptr.reserve(n + 1);
  size_t tot = 0;
  for (const Part& P : parts) tot += P.idx.size();
  m.idx.reserve(tot);
  m.val.reserve(tot);
  for (Part& P : parts) {
    for (int64_t c : P.cnt) m.ptr.push_back(m.ptr.back() + c);
    m.idx.insert(m.idx.end(), P.idx.begin(), P.idx.end());
    m.val.insert(m.val.end(), P.val.begin(), P.val.end());
  }
  return m;
}

std::vector<double> rigid_body_modes(int64_t nn, const double* coords, const int64_t* ndof, int& nb) {
  bool planar = true;
  for (int64_t v = 0; v < nn && planar; ++v)
    if (coords[3 * v + 2] != 0.0) planar = false;
  nb = planar ? 3 : 6;
  int64_t n = 0;
  for (int64_t v = 0; v < nn; ++v) n += ndof[v];
  double c0[3] = {0, 0, 0};
  const double cnt = static_cast<double>(nn ? nn : 1);
  for (int64_t v = 0; v < nn; ++v)
    for (int a = 0; a < 3; ++a) c0[a] += coords[3 * v + a] / cnt;
  std::vector<double> b(static_cast<size_t>(n) * nb, 0.0);
  int64_t row = 0;
  for (int64_t v = 0; v < nn; ++v) {
    const int64_t d = ndof[v];
    const double x[3] = {coords[3 * v] - c0[0], coords[3 * v + 1] - c0[1], coords[3 * v + 2] - c0[2]};
    for (int64_t l = 0; l < d; ++l) {
      double* br = b.data() + (row + l) * nb;
      if (planar) {
        if (l == 0) {
          br[0] = 1.0;
          br[2] = -x[1];
        } else if (l == 1) {
          br[1] = 1.0;
          br[2] = x[0];
        } else if (d == 3 && l == 2) {
          br[2] = 1.0;
        } else {
          br[l % nb] = 1.0;
        }
      } else {
        if (l == 0) {
          br[0] = 1.0;
          br[4] = x[2];
          br[5] = -x[1];
        } else if (l == 1) {
          br[1] = 1.0;
          br[3] = -x[2];
          br[5] = x[0];
        } else if (l == 2) {
          br[2] = 1.0;
          br[3] = x[1];
          br[4] = -x[0];
        } else if (l >= 3 && l <= 5 && d >= 6) {
          br[l] = 1.0;
        } else {
          br[l % nb] = 1.0;
        }
      }
    }
    row += d;
  }
  return b;
}

std::vector<int64_t> aggregate_passes(int64_t nb, const std::vector<int64_t>& sptr, const std::vector<int32_t>& sidx,
                                      int64_t& nagg) {
  std::vector<int64_t> agg(nb, -1);
  int64_t n = 0;
  // pass 1: blocks whose strong neighbours are all free start an aggregate
  for (int64_t b = 0; b < nb; ++b) {
    if (agg[b] != -1) continue;
    bool all = true;
    for (int64_t q = sptr[b]; q < sptr[b + 1]; ++q)
      if (agg[sidx[q]] != -1) {
        all = false;
        break;
      }
    if (!all) continue;
    agg[b] = n;
    for (int64_t q = sptr[b]; q < sptr[b + 1]; ++q) agg[sidx[q]] = n;
    ++n;
  }
  // pass 2: join the aggregate of the first strong neighbour (as of pass 1)
  const std::vector<int64_t> snap = agg;
  for (int64_t b = 0; b < nb; ++b) {
    if (agg[b] != -1) continue;
    for (int64_t q = sptr[b]; q < sptr[b + 1]; ++q)
      if (snap[sidx[q]] != -1) {
        agg[b] = snap[sidx[q]];
        break;
      }
  }
  // pass 3: the rest form aggregates with their free strong neighbours
  for (int64_t b = 0; b < nb; ++b) {
    if (agg[b] != -1) continue;
    agg[b] = n;
    for (int64_t q = sptr[b]; q < sptr[b + 1]; ++q)
      if (agg[sidx[q]] == -1) agg[sidx[q]] = n;
    ++n;
  }
  nagg = n;
  return agg;
}

std::vector<int64_t> amg_aggregate(const Csr& a, const std::vector<int64_t>& bptr, double theta, int64_t& nagg) {
  const int64_t nb = static_cast<int64_t>(bptr.size()) - 1;
  std::vector<int64_t> blk(bptr[nb]);
  for (int64_t b = 0; b < nb; ++b)
    for (int64_t i = bptr[b]; i < bptr[b + 1]; ++i) blk[i] = b;
  // squared Frobenius norms of the diagonal blocks and of the coupling blocks
  std::vector<double> dnorm(nb, 0.0);
  for (int64_t i = 0; i < a.nrows; ++i)
    for (int64_t p = a.ptr[i]; p < a.ptr[i + 1]; ++p)
      if (blk[a.idx[p]] == blk[i]) dnorm[blk[i]] += a.val[p] * a.val[p];
  // strong neighbours of every block, in the order of first appearance
  std::vector<int64_t> sptr(nb + 1, 0);
  std::vector<int32_t> sidx;
  {
    std::vector<double> acc(nb, 0.0);
    std::vector<int64_t> mark(nb, -1), list;
    for (int64_t b = 0; b < nb; ++b) {
      list.clear();
      for (int64_t i = bptr[b]; i < bptr[b + 1]; ++i)
        for (int64_t p = a.ptr[i]; p < a.ptr[i + 1]; ++p) {
          const int64_t c = blk[a.idx[p]];
          if (c == b) continue;
          if (mark[c] != b) {
            mark[c] = b;
            acc[c] = 0.0;
            list.push_back(c);
          }
          acc[c] += a.val[p] * a.val[p];
        }
      for (int64_t c : list)
        if (acc[c] > theta * theta * std::sqrt(dnorm[b] * dnorm[c])) sidx.push_back(static_cast<int32_t>(c));
      sptr[b + 1] = static_cast<int64_t>(sidx.size());
    }
  }
  return aggregate_passes(nb, sptr, sidx, nagg);
}

void coarse_ldlt(const Csr& a, std::vector<double>& l, std::vector<double>& d) {
  // dense LDL^T of the coarsest matrix (amg.rs:319-340)
  const int64_t n = a.nrows;
  l.assign(static_cast<size_t>(n) * n, 0.0);
  d.assign(n, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = a.ptr[i]; p < a.ptr[i + 1]; ++p) l[i * n + a.idx[p]] = a.val[p];
  double dmax = 0.0;
  for (int64_t i = 0; i < n; ++i) dmax = std::fmax(dmax, std::fabs(l[i * n + i]));
  for (int64_t j = 0; j < n; ++j) {
    double dj = l[j * n + j];
    for (int64_t k = 0; k < j; ++k) dj -= l[j * n + k] * l[j * n + k] * d[k];
    if (!(dj > 1e-13 * dmax)) throw AmgBreakdown("AMG setup: the coarse LDL^T broke down");
    d[j] = dj;
    for (int64_t i = j + 1; i < n; ++i) {
      double s = l[i * n + j];
      for (int64_t k = 0; k < j; ++k) s -= l[i * n + k] * l[j * n + k] * d[k];
      l[i * n + j] = s / dj;
    }
  }
  for (int64_t i = 0; i < n; ++i) {
    l[i * n + i] = 1.0;
    for (int64_t j = i + 1; j < n; ++j) l[i * n + j] = 0.0;
  }
}

namespace {

// tentative prolongator (src/solver/iterative/amg.rs:184-257): per aggregate, a
// modified Gram-Schmidt QR (two passes) of the near-null space rows; columns
// with norm <= 1e-10 x the largest column norm are dropped
Csr tentative(const std::vector<int64_t>& bptr, const std::vector<int64_t>& agg, int64_t nagg, const std::vector<double>& b,
              int nb, std::vector<double>& bc, std::vector<int64_t>& cptr) {
  const int64_t nblk = static_cast<int64_t>(bptr.size()) - 1, n = bptr[nblk];
  std::vector<std::vector<int64_t>> members(nagg);
  for (int64_t k = 0; k < nblk; ++k) members[agg[k]].push_back(k);
  struct Ent {
    int64_t r, c;
    double v;
  };
  std::vector<Ent> ent;
  bc.clear();
  cptr.assign(1, 0);
  int64_t nc = 0;
  std::vector<int64_t> dofs;
  std::vector<double> q, r, kept;
  for (int64_t g = 0; g < nagg; ++g) {
    dofs.clear();
    for (int64_t k : members[g])
      for (int64_t i = bptr[k]; i < bptr[k + 1]; ++i) dofs.push_back(i);
    const int64_t nd = static_cast<int64_t>(dofs.size());
    q.assign(static_cast<size_t>(nb) * nd, 0.0);
    for (int c = 0; c < nb; ++c)
      for (int64_t i = 0; i < nd; ++i) q[c * nd + i] = b[dofs[i] * nb + c];
    const int64_t rr = std::min<int64_t>(nb, nd);
    r.assign(static_cast<size_t>(rr) * nb, 0.0);
    kept.assign(static_cast<size_t>(nb) * nd + nd, 0.0);
    int64_t nk = 0;
    double scale = 0.0;
    for (int c = 0; c < nb; ++c) {
      double s = 0.0;
      for (int64_t i = 0; i < nd; ++i) s += q[c * nd + i] * q[c * nd + i];
      scale = std::fmax(scale, s);
    }
    scale = std::sqrt(scale);
    for (int c = 0; c < nb; ++c) {
      double* v = q.data() + c * nd;
      for (int pass = 0; pass < 2; ++pass)
        for (int64_t k = 0; k < nk; ++k) {
          const double* qk = kept.data() + k * nd;
          double s = 0.0;
          for (int64_t i = 0; i < nd; ++i) s += qk[i] * v[i];
          if (pass == 0) r[k * nb + c] = s;
          else r[k * nb + c] += s;
          for (int64_t i = 0; i < nd; ++i) v[i] -= s * qk[i];
        }
      double norm = 0.0;
      for (int64_t i = 0; i < nd; ++i) norm += v[i] * v[i];
      norm = std::sqrt(norm);
      if (nk < nd && norm > 1e-10 * std::fmax(scale, std::numeric_limits<double>::min())) {
        r[nk * nb + c] = norm;
        for (int64_t i = 0; i < nd; ++i) kept[nk * nd + i] = v[i] / norm;
        ++nk;
      }
    }
    if (nk == 0) {
      const double s = 1.0 / std::sqrt(static_cast<double>(nd));
      for (int64_t i = 0; i < nd; ++i) kept[i] = s;
      nk = 1;
    }
    for (int64_t k = 0; k < nk; ++k)
      for (int64_t li = 0; li < nd; ++li) {
        const double x = kept[k * nd + li];
        if (x != 0.0) ent.push_back({dofs[li], nc + k, x});
      }
    for (int64_t k = 0; k < nk; ++k)
      for (int c = 0; c < nb; ++c) bc.push_back(k < rr ? r[k * nb + c] : 0.0);
    nc += nk;
    cptr.push_back(nc);
  }
  std::sort(ent.begin(), ent.end(), [](const Ent& x, const Ent& y) { return x.r != y.r ? x.r < y.r : x.c < y.c; });
  Csr p;
  p.nrows = n;
  p.ncols = nc;
  p.ptr.assign(n + 1, 0);
  for (const Ent& e : ent) ++p.ptr[e.r + 1];
  for (int64_t i = 0; i < n; ++i) p.ptr[i + 1] += p.ptr[i];
  p.idx.resize(ent.size());
  p.val.resize(ent.size());
  for (size_t e = 0; e < ent.size(); ++e) {
    p.idx[e] = static_cast<int32_t>(ent[e].c);
    p.val[e] = ent[e].v;
  }
  return p;
}

// P = Ptilde - (omega D^-1) A Ptilde, merged row by row, zeros dropped (amg.rs:413-445)
Csr sub_scaled(const Csr& pt, const Csr& ap, const std::vector<double>& diag, double omega) {
  Csr m;
  m.nrows = pt.nrows;
  m.ncols = pt.ncols;
  m.ptr.assign(1, 0);
  m.ptr.reserve(pt.nrows + 1);
  for (int64_t i = 0; i < pt.nrows; ++i) {
    const double s = omega / diag[i];
    int64_t p = pt.ptr[i], pe = pt.ptr[i + 1], q = ap.ptr[i], qe = ap.ptr[i + 1];
    while (p < pe || q < qe) {
      const int64_t cp = p < pe ? pt.idx[p] : std::numeric_limits<int64_t>::max();
      const int64_t cq = q < qe ? ap.idx[q] : std::numeric_limits<int64_t>::max();
      int64_t c;
      double v;
      if (cp == cq) {
        v = pt.val[p] - s * ap.val[q];
        ++p;
        ++q;
        c = cp;
      } else if (cp < cq) {
        v = pt.val[p];
        ++p;
        c = cp;
      } else {
        v = -s * ap.val[q];
        ++q;
        c = cq;
      }
      if (v != 0.0) {
        m.idx.push_back(static_cast<int32_t>(c));
        m.val.push_back(v);
      }
    }
    m.ptr.push_back(static_cast<int64_t>(m.idx.size()));
  }
  return m;
}

}  // namespace

AmgHost amg_setup(Csr a, std::vector<int64_t> bptr, std::vector<double> b, int nb,
                  const std::function<double(const Csr&, const std::vector<double>&)>& spectral_radius, bool keep) {
  AmgHost H;
  H.sizes.push_back(a.nrows);
  const double fine_nnz = static_cast<double>(a.nnz());
  double total = fine_nnz;
  double theta = 0.08;
  // RUSTYFEM_AMG_TIMING=1: the seconds of each setup phase per level, on stderr
  const char* tenv = std::getenv("RUSTYFEM_AMG_TIMING");
  const bool timing = tenv && *tenv && *tenv != '0';
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto secs = [](std::chrono::steady_clock::time_point a0, std::chrono::steady_clock::time_point a1) {
    return std::chrono::duration<double>(a1 - a0).count();
  };
  while (a.nrows > 1500 && H.levels.size() < 12) {
    const auto t0 = now();
    std::vector<double> diag = csr_diag(a);
    for (double d : diag)
      if (!(d > 0.0)) throw AmgBreakdown("AMG setup: a diagonal entry is not positive");
    int64_t nagg = 0;
    const std::vector<int64_t> agg = amg_aggregate(a, bptr, theta, nagg);
    const auto t1 = now();
    std::vector<double> bc;
    std::vector<int64_t> cptr;
    Csr pt = tentative(bptr, agg, nagg, b, nb, bc, cptr);
    const auto t2 = now();
    if (static_cast<double>(pt.ncols) > 0.85 * static_cast<double>(a.nrows)) break;
    const double rho = spectral_radius(a, diag);
    const double omega = 4.0 / 3.0 / rho;
    const auto t3 = now();
    Csr ap = csr_mul(a, pt);
    Csr p = sub_scaled(pt, ap, diag, omega);
    ap = Csr();
    const auto t4 = now();
    AmgHostLevel L;
    if (keep) {
      L.agg = agg;
      L.Pt = std::move(pt);
      L.B = b;
      L.nbm = nb;
    }
    pt = Csr();
    Csr r = csr_transpose(p);
    const auto t5 = now();
    Csr apm = csr_mul(a, p);
    const auto t6 = now();
    Csr ac = csr_mul(r, apm);
    apm = Csr();
    const auto t7 = now();
    if (timing)
      std::fprintf(stderr,
                   "[amg] level %zu: n %lld nnz %lld | aggregate %.3f s, tentative %.3f, spectral radius %.3f, "
                   "A P~ and smoothing %.3f, transpose %.3f, A P %.3f, R (A P) %.3f\n",
                   H.levels.size(), static_cast<long long>(a.nrows), static_cast<long long>(a.nnz()), secs(t0, t1),
                   secs(t1, t2), secs(t2, t3), secs(t3, t4), secs(t4, t5), secs(t5, t6), secs(t6, t7));
    total += static_cast<double>(ac.nnz());
    H.sizes.push_back(ac.nrows);
    L.nblk = static_cast<int64_t>(bptr.size()) - 1;
    L.bptr = std::move(bptr);
    L.A = std::move(a);
    L.P = std::move(p);
    L.R = std::move(r);
    L.diag = std::move(diag);
    L.rho = rho;
    L.omega = omega;
    H.levels.push_back(std::move(L));
    a = std::move(ac);
    bptr = std::move(cptr);
    b = std::move(bc);
    theta *= 0.5;
  }
  H.coarse_bptr = bptr;
  if (keep) H.coarse_B = b;
  std::vector<double> l, d;
  coarse_ldlt(a, l, d);
  H.coarse = std::move(a);
  H.coarse_l = std::move(l);
  H.coarse_d = std::move(d);
  H.complexity = total / fine_nnz;
  return H;
}

std::vector<double> coarse_inverse(const std::vector<double>& l, const std::vector<double>& d, int64_t n) {
  // X = L^-1 (unit lower), column by column
  std::vector<double> x(static_cast<size_t>(n) * n, 0.0);
  const int t = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(host_threads(), n / 64)));
  auto par = [&](auto&& body) {
    std::vector<std::thread> th;
    for (int c = 0; c < t; ++c)
      th.emplace_back([&, c] {
        for (int64_t i = c; i < n; i += t) body(i);
      });
    for (auto& h : th) h.join();
  };
  par([&](int64_t c) {
    x[c * n + c] = 1.0;
    for (int64_t i = c + 1; i < n; ++i) {
      double s = 0.0;
      for (int64_t k = c; k < i; ++k) s += l[i * n + k] * x[k * n + c];
      x[i * n + c] = -s;
    }
  });
  // M = X^T D^-1 X
  std::vector<double> m(static_cast<size_t>(n) * n, 0.0);
  par([&](int64_t i) {
    for (int64_t j = 0; j < n; ++j) {
      double s = 0.0;
      for (int64_t k = std::max(i, j); k < n; ++k) s += x[k * n + i] * x[k * n + j] / d[k];
      m[i * n + j] = s;
    }
  });
  return m;
}

}  // namespace rf
