// The C ABI of ffi/src/lib.rs, served by the MI355X multifrontal solver.
//
// build() restates ffi/src/lib.rs:100-229 (validation, uncoupled DOFs,
// node compaction, adjacency, nested dissection with leaf 16, equation map,
// entry filter per mode); the numeric factorization and the solves run on
// the GPU (mf_gpu.hip, indef_gpu.hip).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rustyfem.h"
#include "amg.hpp"
#include "assembly.hpp"
#include "eigen.hpp"
#include "gpu.hpp"
#include "indef.hpp"
#include "mf.hpp"
#include "mpnn_train.hpp"
#include "chebyshev.hpp"
#include "pcg.hpp"
#include "symbolic.hpp"

namespace {

enum class Mode { Pd, Indef, Lu };

struct Options {
  rf::Precision prec = rf::Precision::FP64;
  int64_t leaf = 16;  // ffi/src/lib.rs:17
  int max_refine = 0;
  double refine_tol = 1e-10;
};

Options resolve(const rustyfem_options* o) {
  Options r;
  if (!o) return r;
  r.prec = o->precision == 1 ? rf::Precision::MIXED : rf::Precision::FP64;
  r.leaf = o->leaf > 0 ? o->leaf : 16;
  r.max_refine = o->max_refine >= 0 ? o->max_refine : (r.prec == rf::Precision::MIXED ? 30 : 0);
  r.refine_tol = o->refine_tol > 0 ? o->refine_tol : 1e-10;
  return r;
}

// f(begin, end) over [0, n) in contiguous parts on the host threads
template <typename F>
void parallel_parts(int64_t n, F&& f) {
  const int64_t t = std::max<int64_t>(1, std::min<int64_t>(std::thread::hardware_concurrency(), n / 65536 + 1));
  if (t == 1) {
    f(int64_t{0}, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + t - 1) / t;
  for (int64_t k = 0; k < t; ++k) {
    const int64_t b = std::min(n, k * chunk), e = std::min(n, b + chunk);
    th.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto& x : th) x.join();
}

// a number per handle that the process never gives twice
uint64_t next_serial() {
  static std::atomic<uint64_t> s{0};
  return ++s;
}

}  // namespace

struct Handle {
  uint64_t serial = next_serial();
  Mode mode = Mode::Pd;
  Options opt;
  int64_t n = 0;                 // caller DOFs
  std::vector<int64_t> eq;       // internal equation per caller DOF, -1 = uncoupled
  std::vector<double> diag;      // caller diagonal (uncoupled DOFs use it)
  std::vector<uint8_t> coupled;
  int64_t n_sym = 0;
  double work = 0.0;
  int64_t factor_size = 0;
  bool has_inertia = false;
  int64_t inertia[3] = {0, 0, 0};
  int64_t nfronts = 0, nlevels = 0;
  // pattern: caller entry p -> Lower value index (-1 = dropped by the filter)
  std::vector<int64_t> indptr, indices, vmap;
  std::vector<int64_t> colptr, rowidx;
  std::vector<int64_t> owner;
  std::unique_ptr<rf::GpuFactor> gpu;
  // mixed precision: the FP64 factor of the same values, built when a solve's
  // refinement misses its target, dropped at the next refactorization
  std::unique_ptr<rf::GpuFactor> gpu64;
  std::unique_ptr<rf::IndefGpu> ind;
  std::unique_ptr<rf::Symbolic> sym;  // kept by host-only analyses and mixed-precision handles
  // solves share the handle's device work buffers and graphs: concurrent
  // rustyfem_solve calls on one handle (allowed by ffi/src/lib.rs:336-371)
  // are serialized here
  mutable std::mutex solve_mu;
};

namespace {

void set_error(const std::string& s) { rf::last_error() = s; }

// ffi/src/lib.rs:100-229 up to the factorization; returns nullptr with *err.
// With gpu == false only the host-side symbolic work is done.
Handle* build(size_t n, size_t nn, const int64_t* loc, const double* coords, const int64_t* indptr,
              const int64_t* indices, const double* data, Mode mode, const Options& opt, int64_t* err,
              bool gpu = true) {
  *err = -1;
  const int64_t nnz = indptr[n];
  for (size_t i = 0; i < n; ++i)
    if (loc[i] < 0 || static_cast<size_t>(loc[i]) >= nn) return nullptr;
  for (int64_t p = 0; p < nnz; ++p)
    if (indices[p] < 0 || static_cast<size_t>(indices[p]) >= n) return nullptr;
  auto h = std::make_unique<Handle>();
  h->mode = mode;
  h->opt = opt;
  h->n = static_cast<int64_t>(n);
  h->coupled.assign(n, 0);
  h->diag.assign(n, 0.0);
  for (size_t c = 0; c < n; ++c)
    for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
      const size_t r = static_cast<size_t>(indices[p]);
      if (r == c) {
        h->diag[c] += data[p];
      } else if (data[p] != 0.0) {
        h->coupled[c] = 1;
        h->coupled[r] = 1;
      }
    }
  for (size_t i = 0; i < n; ++i) {
    if (h->coupled[i]) continue;
    const double d = h->diag[i];
    const bool ok = mode == Mode::Pd ? d > 0.0 : (d != 0.0 && std::isfinite(d));
    if (!ok) {
      *err = static_cast<int64_t>(i);
      return nullptr;
    }
  }
  // nodes with a coupled DOF, compact numbering in first-appearance order
  std::vector<int64_t> nid(nn, -1);
  std::vector<double> xyz;
  int64_t nc = 0;
  for (size_t i = 0; i < n; ++i) {
    if (!h->coupled[i]) continue;
    const int64_t v = loc[i];
    if (nid[v] < 0) {
      nid[v] = nc++;
      xyz.insert(xyz.end(), {coords[3 * v], coords[3 * v + 1], coords[3 * v + 2]});
    }
  }
  std::vector<int64_t> ndof(nc, 0), local(n, 0);
  for (size_t i = 0; i < n; ++i)
    if (h->coupled[i]) local[i] = ndof[nid[loc[i]]]++;
  std::vector<std::vector<int64_t>> adj(nc);
  for (size_t c = 0; c < n; ++c) {
    if (!h->coupled[c]) continue;
    for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
      const size_t r = static_cast<size_t>(indices[p]);
      if (h->coupled[r] && nid[loc[r]] != nid[loc[c]]) {
        adj[nid[loc[c]]].push_back(nid[loc[r]]);
        adj[nid[loc[r]]].push_back(nid[loc[c]]);
      }
    }
  }
  std::vector<int64_t> aptr(nc + 1, 0), aidx;
  for (int64_t v = 0; v < nc; ++v) {
    auto& a = adj[v];
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
    aptr[v + 1] = aptr[v] + static_cast<int64_t>(a.size());
  }
  aidx.reserve(aptr[nc]);
  for (auto& a : adj) {
    aidx.insert(aidx.end(), a.begin(), a.end());
    std::vector<int64_t>().swap(a);
  }
  rf::Symbolic sym = rf::analyse(xyz.data(), nc, aptr, aidx, ndof, opt.leaf);
  h->eq.assign(n, -1);
  for (size_t i = 0; i < n; ++i)
    if (h->coupled[i]) h->eq[i] = sym.first_dof[nid[loc[i]]] + local[i];
  h->n_sym = sym.n;
  h->owner.assign(sym.n, 0);
  for (size_t i = 0; i < n; ++i)
    if (h->eq[i] >= 0) h->owner[h->eq[i]] = static_cast<int64_t>(i);
  sym.pattern(aptr, aidx, ndof, h->colptr, h->rowidx);
  h->work = sym.work();
  h->factor_size = sym.factor_size();
  h->nfronts = static_cast<int64_t>(sym.fronts.size());
  // entry filter per mode (ffi/src/lib.rs:176-186, 193-202, 216-224): vmap[p] is
  // the Lower index the caller entry p adds to, or -1
  h->indptr.assign(indptr, indptr + n + 1);
  h->indices.assign(indices, indices + nnz);
  h->vmap.assign(nnz, -1);
  for (size_t c = 0; c < n; ++c) {
    if (!h->coupled[c]) continue;
    const int64_t qc = h->eq[c];
    for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
      const size_t r = static_cast<size_t>(indices[p]);
      const int64_t qr = h->eq[r];
      if (qr < 0) continue;
      int64_t row, col;
      if (mode == Mode::Pd) {
        if (qr < qc) continue;
        row = qr;
        col = qc;
      } else if (mode == Mode::Indef) {
        if (r < c) continue;
        row = std::max(qr, qc);
        col = std::min(qr, qc);
      } else {
        // LU: lower triangle entries by (qr, qc); upper ones are marked with an offset
        row = std::max(qr, qc);
        col = std::min(qr, qc);
      }
      const auto b = h->rowidx.begin() + h->colptr[col], e = h->rowidx.begin() + h->colptr[col + 1];
      const auto it = std::lower_bound(b, e, row);
      if (it == e || *it != row) throw std::runtime_error("internal: entry missing in the pattern");
      int64_t vi = it - h->rowidx.begin();
      if (mode == Mode::Lu && qr < qc) vi += static_cast<int64_t>(h->rowidx.size());
      h->vmap[p] = vi;
    }
  }
  if (!gpu) {
    h->sym = std::make_unique<rf::Symbolic>(std::move(sym));
    return h.release();
  }
  if (mode == Mode::Pd) {
    rf::Plan plan = rf::build_plan(sym, h->colptr, h->rowidx);
    h->nlevels = static_cast<int64_t>(plan.levels.size());
    h->gpu = rf::make_gpu_factor(std::move(plan), h->colptr, h->rowidx, opt.prec);
    h->gpu->set_caller_map(h->eq, h->diag);
    const char* kt = std::getenv("RUSTYFEM_KERNEL_TIMING");
    h->gpu->set_kernel_timing(kt && *kt && *kt != '0');
    if (opt.prec == rf::Precision::MIXED) {
      h->sym = std::make_unique<rf::Symbolic>(std::move(sym));
      Handle* hp = h.get();
      // refinement of the FP32 factor missed its target: the FP64 factor of
      // the same values (built once, on the values in device memory) solves
      // the same internal right-hand sides with the same refinement
      h->gpu->set_fallback([hp](const double* d_b, double* d_x, int nrhs) -> double {
        if (!hp->gpu64) {
          rf::Plan p64 = rf::build_plan(*hp->sym, hp->colptr, hp->rowidx);
          auto g = rf::make_gpu_factor(std::move(p64), hp->colptr, hp->rowidx, rf::Precision::FP64);
          if (g->factor(hp->gpu->device_lower_values(), true) >= 0) return -1.0;
          hp->gpu64 = std::move(g);
        }
        return hp->gpu64->solve_internal(d_b, d_x, nrhs, hp->opt.refine_tol, std::max(hp->opt.max_refine, 1));
      });
    }
  } else {
    h->ind = rf::make_indef_gpu(std::move(sym), h->colptr, h->rowidx, mode == Mode::Lu);
    h->ind->set_caller_map(h->eq, h->diag);
  }
  return h.release();
}

// Lower values of the handle's pattern from caller data (the filter of the mode).
std::vector<double> lower_values(const Handle* h, const double* data) {
  const size_t nl = h->rowidx.size();
  std::vector<double> v(h->mode == Mode::Lu ? 2 * nl : nl, 0.0);
  for (size_t p = 0; p < h->vmap.size(); ++p)
    if (h->vmap[p] >= 0) v[h->vmap[p]] += data[p];
  return v;
}

// numeric factorization; returns -1 or the caller DOF of the failing pivot
int64_t numeric(Handle* h, const double* data) {
  if (h->mode == Mode::Pd) {
    h->gpu64.reset();
    int64_t e;
    if (data) {
      const std::vector<double> v = lower_values(h, data);
      e = h->gpu->factor(v.data(), false);
    } else {
      e = h->gpu->factor(nullptr, true);
    }
    return e < 0 ? -1 : h->owner[e];
  }
  const std::vector<double> v = lower_values(h, data);
  const int64_t e = h->ind->factor(v.data());
  if (e >= 0) return h->owner[e];
  if (h->mode == Mode::Indef) {
    int64_t ine[3];
    h->ind->inertia(ine);
    for (int64_t i = 0; i < h->n; ++i)
      if (!h->coupled[i]) (h->diag[i] < 0.0 ? ine[0] : ine[2]) += 1;
    std::memcpy(h->inertia, ine, sizeof ine);
    h->has_inertia = true;
  }
  h->factor_size = h->ind->factor_size();
  return -1;
}

Handle* factor_mode(size_t n, size_t nn, const int64_t* loc, const double* coords, const int64_t* indptr,
                    const int64_t* indices, const double* data, Mode mode, const Options& opt, double* d_out,
                    int64_t* err_dof) {
  *err_dof = -1;
  set_error("");
  Handle* h = nullptr;
  try {
    int64_t e = -1;
    h = build(n, nn, loc, coords, indptr, indices, data, mode, opt, &e);
    if (!h) {
      *err_dof = e;
      return nullptr;
    }
    const int64_t q = numeric(h, data);
    if (q >= 0) {
      *err_dof = q;
      delete h;
      return nullptr;
    }
    if (mode == Mode::Pd) {
      h->has_inertia = true;
      h->inertia[0] = 0;
      h->inertia[1] = 0;
      h->inertia[2] = static_cast<int64_t>(n);
      if (d_out) {
        std::vector<double> dq(h->n_sym);
        h->gpu->pivots(dq.data());
        for (size_t i = 0; i < n; ++i) d_out[i] = h->coupled[i] ? dq[h->eq[i]] : h->diag[i];
      }
    }
    return h;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  delete h;
  return nullptr;
}

// Many independent problems as one forest: each problem keeps its own
// analysis (its ordering is the one rustyfem_factor computes for it alone);
// the trees are concatenated (equations, fronts and Lower entries offset), so
// every launch of the level schedule covers the fronts of all problems.
rf::Symbolic merge_symbolic(std::vector<std::unique_ptr<Handle>>& subs, std::vector<int64_t>& colptr,
                            std::vector<int64_t>& rowidx, std::vector<int64_t>& eq_off, std::vector<int64_t>& low_off) {
  rf::Symbolic S;
  colptr.assign(1, 0);
  rowidx.clear();
  eq_off.assign(subs.size() + 1, 0);
  low_off.assign(subs.size() + 1, 0);
  int64_t e0 = 0, f0 = 0;
  for (size_t i = 0; i < subs.size(); ++i) {
    rf::Symbolic& A = *subs[i]->sym;
    eq_off[i] = e0;
    low_off[i] = static_cast<int64_t>(rowidx.size());
    for (const rf::Front& F : A.fronts) {
      rf::Front G = F;
      G.start += e0;
      for (auto& r : G.rows) r += e0;
      for (auto& c : G.children) c += f0;
      if (G.parent >= 0) G.parent += f0;
      S.fronts.push_back(std::move(G));
    }
    for (int64_t r : A.roots) S.roots.push_back(r + f0);
    const int64_t base = static_cast<int64_t>(rowidx.size());
    for (int64_t c = 0; c < A.n; ++c) colptr.push_back(base + subs[i]->colptr[c + 1]);
    for (int64_t r : subs[i]->rowidx) rowidx.push_back(r + e0);
    e0 += A.n;
    f0 += static_cast<int64_t>(A.fronts.size());
  }
  eq_off[subs.size()] = e0;
  low_off[subs.size()] = static_cast<int64_t>(rowidx.size());
  S.n = e0;
  return S;
}

}  // namespace

struct BatchHandle {
  std::vector<std::unique_ptr<Handle>> subs;  // per problem: maps and entry filter (no device state)
  std::vector<int64_t> n_off;                 // caller DOF offsets
  std::vector<int64_t> eq_off, low_off;       // internal equation / Lower entry offsets
  std::unique_ptr<rf::GpuFactor> gpu;
  Options opt;
  int64_t n_total = 0, nl = 0, nfronts = 0, nlevels = 0;
  double work = 0.0;
  mutable std::mutex solve_mu;
};

namespace {

// numeric factorization of the batch; fills err (one per problem) and returns false on a failing pivot
bool batch_numeric(BatchHandle* h, const double* const* data, int64_t* err) {
  std::vector<double> v(h->nl, 0.0);
  for (size_t i = 0; i < h->subs.size(); ++i) {
    const Handle* s = h->subs[i].get();
    for (size_t p = 0; p < s->vmap.size(); ++p)
      if (s->vmap[p] >= 0) v[h->low_off[i] + s->vmap[p]] += data[i][p];
  }
  const int64_t e = h->gpu->factor(v.data(), false);
  if (e < 0) return true;
  const size_t i = static_cast<size_t>(std::upper_bound(h->eq_off.begin(), h->eq_off.end(), e) - h->eq_off.begin()) - 1;
  if (err) err[i] = h->subs[i]->owner[e - h->eq_off[i]];
  return false;
}

// entries of rustyfem_stats / rustyfem_batch_stats that come from the
// factorization's timings (the layout is documented in rustyfem.h)
constexpr size_t NSTATS = 26;
void fill_timings(const rf::Timings& t, double* v) {
  v[0] = t.factor_ms;
  v[1] = t.solve_ms;
  v[2] = t.refine_ms;
  v[3] = t.refine_steps;
  v[4] = t.rel_residual;
  v[10] = t.gemm_ms;
  v[11] = t.gemm_launches;
  v[12] = t.gemm_flops;
  v[13] = t.tune_ms;
  v[14] = t.tuned_launches;
  v[15] = t.agpr_launches;
  for (int i = 0; i < 4; ++i) v[16 + i] = t.refine_history[i];
  v[20] = t.ilv_launches;
  v[21] = t.refine_undone;
  v[22] = t.fell_back;
  v[23] = t.fallback_rel;
  v[24] = t.q_launches;
  v[25] = t.small_levels;
}

}  // namespace

extern "C" {

BatchHandle* rustyfem_factor_batch(size_t nprob, const size_t* n, const size_t* nn, const int64_t* const* loc,
                                   const double* const* coords, const int64_t* const* indptr,
                                   const int64_t* const* indices, const double* const* data,
                                   const rustyfem_options* opt, int64_t* err_dof) {
  set_error("");
  for (size_t i = 0; i < nprob; ++i) err_dof[i] = -1;
  try {
    auto h = std::make_unique<BatchHandle>();
    h->opt = resolve(opt);
    h->n_off.assign(nprob + 1, 0);
    bool bad = false;
    for (size_t i = 0; i < nprob; ++i) {
      int64_t e = -1;
      std::unique_ptr<Handle> s(build(n[i], nn[i], loc[i], coords[i], indptr[i], indices[i], data[i], Mode::Pd,
                                      h->opt, &e, false));
      if (!s) {
        err_dof[i] = e;
        bad = true;
        continue;
      }
      h->n_off[i + 1] = h->n_off[i] + static_cast<int64_t>(n[i]);
      h->work += s->work;
      h->subs.push_back(std::move(s));
    }
    if (bad) return nullptr;
    std::vector<int64_t> colptr, rowidx;
    rf::Symbolic S = merge_symbolic(h->subs, colptr, rowidx, h->eq_off, h->low_off);
    h->n_total = h->n_off[nprob];
    h->nl = static_cast<int64_t>(rowidx.size());
    h->nfronts = static_cast<int64_t>(S.fronts.size());
    std::vector<int64_t> eq(h->n_total, -1);
    std::vector<double> diag(h->n_total, 0.0);
    for (size_t i = 0; i < nprob; ++i) {
      const Handle* s = h->subs[i].get();
      for (int64_t j = 0; j < s->n; ++j) {
        eq[h->n_off[i] + j] = s->eq[j] >= 0 ? s->eq[j] + h->eq_off[i] : -1;
        diag[h->n_off[i] + j] = s->diag[j];
      }
    }
    rf::Plan plan = rf::build_plan(S, colptr, rowidx);
    h->nlevels = static_cast<int64_t>(plan.levels.size());
    h->gpu = rf::make_gpu_factor(std::move(plan), colptr, rowidx, h->opt.prec);
    h->gpu->set_caller_map(eq, diag);
    h->gpu->set_segments(h->eq_off);
    for (auto& s : h->subs) s->sym.reset();
    if (!batch_numeric(h.get(), data, err_dof)) return nullptr;
    return h.release();
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return nullptr;
}

int32_t rustyfem_refactor_batch(BatchHandle* h, const double* const* data, int64_t* err_dof) {
  for (size_t i = 0; i < h->subs.size(); ++i) err_dof[i] = -1;
  try {
    if (!data) {
      const int64_t e = h->gpu->factor(nullptr, true);
      if (e < 0) return 0;
      const size_t i = static_cast<size_t>(std::upper_bound(h->eq_off.begin(), h->eq_off.end(), e) - h->eq_off.begin()) - 1;
      err_dof[i] = h->subs[i]->owner[e - h->eq_off[i]];
      return 1;
    }
    return batch_numeric(h, data, err_dof) ? 0 : 1;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

int32_t rustyfem_solve_batch(const BatchHandle* h, size_t nrhs, double* const* b) {
  if (h->n_total == 0 || nrhs == 0) return 0;
  try {
    std::lock_guard<std::mutex> lock(h->solve_mu);
    const int64_t N = h->n_total;
    std::vector<double> all(static_cast<size_t>(N) * nrhs);
    for (size_t i = 0; i < h->subs.size(); ++i) {
      const int64_t ni = h->n_off[i + 1] - h->n_off[i];
      for (size_t j = 0; j < nrhs; ++j)
        std::memcpy(all.data() + j * N + h->n_off[i], b[i] + j * ni, ni * sizeof(double));
    }
    h->gpu->solve(all.data(), static_cast<int>(nrhs), false, h->opt.refine_tol, h->opt.max_refine);
    for (size_t i = 0; i < h->subs.size(); ++i) {
      const int64_t ni = h->n_off[i + 1] - h->n_off[i];
      for (size_t j = 0; j < nrhs; ++j)
        std::memcpy(b[i] + j * ni, all.data() + j * N + h->n_off[i], ni * sizeof(double));
    }
    return 0;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

int32_t rustyfem_solve_batch_device(const BatchHandle* h, size_t nrhs, double* d_b) {
  if (h->n_total == 0 || nrhs == 0) return 0;
  try {
    std::lock_guard<std::mutex> lock(h->solve_mu);
    h->gpu->solve(d_b, static_cast<int>(nrhs), true, h->opt.refine_tol, h->opt.max_refine);
    return 0;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

void rustyfem_batch_info(const BatchHandle* h, double* out) {
  // [problems, total DOFs, fronts, levels, work flops]
  out[0] = static_cast<double>(h->subs.size());
  out[1] = static_cast<double>(h->n_total);
  out[2] = static_cast<double>(h->nfronts);
  out[3] = static_cast<double>(h->nlevels);
  out[4] = h->work;
}

void rustyfem_batch_stats(const BatchHandle* h, double* out, size_t len) {
  double v[NSTATS] = {0};
  fill_timings(h->gpu->timings(), v);
  v[5] = h->gpu->flops();
  v[6] = static_cast<double>(h->gpu->device_bytes());
  v[7] = static_cast<double>(static_cast<int>(h->opt.prec));
  v[8] = static_cast<double>(h->nfronts);
  v[9] = static_cast<double>(h->nlevels);
  for (size_t i = 0; i < len && i < NSTATS; ++i) out[i] = v[i];
}

void rustyfem_free_batch(BatchHandle* h) { delete h; }

// order of problem i (as rustyfem_order of that problem alone)
void rustyfem_batch_order(const BatchHandle* h, size_t i, int64_t* out) {
  const Handle* s = h->subs[i].get();
  int64_t next = s->n_sym;
  for (int64_t j = 0; j < s->n; ++j) out[j] = s->eq[j] < 0 ? next++ : s->eq[j];
}

Handle* rustyfem_factor(size_t n, size_t nn, const int64_t* loc, const double* coords, const int64_t* indptr,
                        const int64_t* indices, const double* data, double* d_out, int64_t* err_dof) {
  return factor_mode(n, nn, loc, coords, indptr, indices, data, Mode::Pd, Options{}, d_out, err_dof);
}

Handle* rustyfem_factor_ex(size_t n, size_t nn, const int64_t* loc, const double* coords, const int64_t* indptr,
                           const int64_t* indices, const double* data, const rustyfem_options* opt, double* d_out,
                           int64_t* err_dof) {
  return factor_mode(n, nn, loc, coords, indptr, indices, data, Mode::Pd, resolve(opt), d_out, err_dof);
}

Handle* rustyfem_factor_indef(size_t n, size_t nn, const int64_t* loc, const double* coords, const int64_t* indptr,
                              const int64_t* indices, const double* data, int64_t* inertia_out, int64_t* err_dof) {
  Handle* h = factor_mode(n, nn, loc, coords, indptr, indices, data, Mode::Indef, Options{}, nullptr, err_dof);
  if (h && inertia_out) std::memcpy(inertia_out, h->inertia, 3 * sizeof(int64_t));
  return h;
}

Handle* rustyfem_factor_unsym(size_t n, size_t nn, const int64_t* loc, const double* coords, const int64_t* indptr,
                              const int64_t* indices, const double* data, int64_t* err_dof) {
  return factor_mode(n, nn, loc, coords, indptr, indices, data, Mode::Lu, Options{}, nullptr, err_dof);
}

int32_t rustyfem_inertia(const Handle* h, int64_t* out) {
  if (!h->has_inertia) return -1;
  std::memcpy(out, h->inertia, 3 * sizeof(int64_t));
  return 0;
}

int32_t rustyfem_solve(const Handle* h, size_t nrhs, double* b) {
  if (h->n == 0 || nrhs == 0) return 0;
  try {
    std::lock_guard<std::mutex> lock(h->solve_mu);
    if (h->mode == Mode::Pd)
      h->gpu->solve(b, static_cast<int>(nrhs), false, h->opt.refine_tol, h->opt.max_refine);
    else
      h->ind->solve(b, static_cast<int>(nrhs));
    return 0;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

void rustyfem_info(const Handle* h, double* out) {
  out[0] = h->work;
  out[1] = static_cast<double>(h->factor_size);
}

void rustyfem_order(const Handle* h, int64_t* out) {
  int64_t next = h->n_sym;
  for (int64_t i = 0; i < h->n; ++i) out[i] = h->eq[i] < 0 ? next++ : h->eq[i];
}

void rustyfem_free(Handle* h) { delete h; }

int32_t rustyfem_refactor(Handle* h, const double* data, int64_t* err_dof) {
  *err_dof = -1;
  try {
    if (!data && h->mode != Mode::Pd) throw std::runtime_error("refactor without data needs a positive definite handle");
    const int64_t q = numeric(h, data);
    if (q >= 0) {
      *err_dof = q;
      return 1;
    }
    return 0;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

int32_t rustyfem_solve_device(const Handle* h, size_t nrhs, double* d_b) {
  if (h->n == 0 || nrhs == 0) return 0;
  try {
    if (h->mode != Mode::Pd) throw std::runtime_error("device solves need a positive definite handle");
    std::lock_guard<std::mutex> lock(h->solve_mu);
    h->gpu->solve(d_b, static_cast<int>(nrhs), true, h->opt.refine_tol, h->opt.max_refine);
    return 0;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

int32_t rustyfem_solve_device_noalloc(const Handle* h, size_t nrhs, double* d_b) {
  if (h->n == 0 || nrhs == 0) return 0;
  try {
    if (h->mode != Mode::Pd) throw std::runtime_error("device solves need a positive definite handle");
    if (nrhs > static_cast<size_t>(INT32_MAX)) throw std::runtime_error("too many right-hand sides");
    std::lock_guard<std::mutex> lock(h->solve_mu);
    const bool met = h->gpu->solve_noalloc(d_b, static_cast<int>(nrhs), h->opt.refine_tol, h->opt.max_refine);
    return met ? 0 : 1;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

int64_t rustyfem_eigen(const Handle* h, const int64_t* b_indptr, const int64_t* b_indices, const double* b_data,
                       size_t nev, size_t nvec, size_t max_iter, double tol, double* theta_out, double* x_out,
                       int64_t* iterations_out) {
  try {
    if (h->mode != Mode::Pd) throw std::runtime_error("the eigenvalue solver needs a positive definite handle");
    const int64_t n = h->n;
    for (int64_t c = 0; c < n; ++c)
      if (b_indptr[c + 1] < b_indptr[c]) throw std::runtime_error("B: column pointers decrease");
    for (int64_t p = 0; p < b_indptr[n]; ++p)
      if (b_indices[p] < 0 || b_indices[p] >= n) throw std::runtime_error("B: row index out of range");
    rf::EigenOptions o;
    o.nev = static_cast<int>(nev);
    o.vectors = static_cast<int>(nvec);
    if (max_iter > 0) o.max_iter = static_cast<int>(max_iter);
    if (tol > 0.0) o.tol = tol;
    std::lock_guard<std::mutex> lock(h->solve_mu);
    // Op^-1 on device blocks in the caller's numbering (the handle's solve,
    // with its refinement when the handle is mixed precision)
    const rf::BlockSolve solve = [h](double* d_y, int q) {
      h->gpu->solve(d_y, q, true, h->opt.refine_tol, h->opt.max_refine);
    };
    // a symmetric B's CSC arrays are its CSR arrays
    const rf::EigenResult r = rf::subspace_gpu(n, b_indptr, b_indices, b_data, solve, o);
    const size_t k = r.theta.size();
    std::copy(r.theta.begin(), r.theta.end(), theta_out);
    if (x_out) std::copy(r.vectors.begin(), r.vectors.end(), x_out);
    if (iterations_out) *iterations_out = r.iterations;
    return static_cast<int64_t>(k);
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

void rustyfem_stats(const Handle* h, double* out, size_t len) {
  double v[NSTATS] = {0};
  if (h->gpu) {
    fill_timings(h->gpu->timings(), v);
    v[5] = h->gpu->flops();
    v[6] = static_cast<double>(h->gpu->device_bytes());
  }
  v[7] = static_cast<double>(static_cast<int>(h->opt.prec));
  v[8] = static_cast<double>(h->nfronts);
  v[9] = static_cast<double>(h->nlevels);
  for (size_t i = 0; i < len && i < NSTATS; ++i) out[i] = v[i];
}

const char* rustyfem_last_error(void) { return rf::last_error().c_str(); }

}  // extern "C"

// an allocator whose resize leaves new elements uninitialized (the copies
// below fill them in parallel)
template <typename T>
struct NoInitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <typename U>
  NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  template <typename U, typename... Args>
  void construct(U* p, Args&&... args) {
    if constexpr (sizeof...(Args) == 0)
      ::new (static_cast<void*>(p)) U;
    else
      ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
  }
};

extern "C" {

struct PcgHandle {
  std::unique_ptr<rf::PcgGpu> solver;  // null when the handle solves directly
  int64_t n = 0;
  std::mutex mu;
  // IterMatrix's direct solver (src/solver/iterative.rs:138-190): positive
  // definite multifrontal LDL^T of the same values, ND leaves of 48 nodes
  std::unique_ptr<Handle> direct;
  int direct_reason = 0;  // 0 none, 1 non-positive diagonal, 2 AMG breakdown, 3 non-convergence
  bool fallback = true;   // repeat non-converged solves directly
  bool last_direct = false;
  double last_direct_rel = -1.0;
  // the inputs, for a direct factor built when a solve does not converge
  std::vector<int64_t> colptr, ndof;
  std::vector<int64_t, NoInitAlloc<int64_t>> rowidx;
  std::vector<double, NoInitAlloc<double>> val;
  std::vector<double> coords;
  int64_t nn = 0;
};

namespace {

// The lower triangle (natural numbering) as a caller matrix of the PD
// factorization: both triangles, loc from ndof.  Returns the handle, or
// nullptr with *err = the natural equation of the failing pivot (-1 when the
// input is rejected otherwise).
Handle* direct_factor(int64_t n, const int64_t* colptr, const int64_t* rowidx, const double* val, int64_t nn,
                      const double* coords, const int64_t* ndof, int64_t* err) {
  std::vector<int64_t> loc(n, 0);
  for (int64_t v = 0, i = 0; v < nn; ++v)
    for (int64_t l = 0; l < ndof[v]; ++l) loc[i++] = v;
  std::vector<int64_t> cnt(n + 1, 0);
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p) {
      ++cnt[c + 1];
      if (rowidx[p] != c) ++cnt[rowidx[p] + 1];
    }
  for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1), ind(cnt[n]);
  std::vector<double> dat(cnt[n]);
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p) {
      const int64_t r = rowidx[p];
      ind[pos[c]] = r;
      dat[pos[c]++] = val[p];
      if (r != c) {
        ind[pos[r]] = c;
        dat[pos[r]++] = val[p];
      }
    }
  Options opt;
  opt.leaf = 48;  // sparse::analyse, src/sparse.rs:14
  return factor_mode(static_cast<size_t>(n), static_cast<size_t>(nn), loc.data(), coords, cnt.data(), ind.data(),
                     dat.data(), Mode::Pd, opt, nullptr, err);
}

// ||b - A x|| / ||b|| of host vectors for the lower triangle in CSC
double lower_rel_residual(const PcgHandle* h, const double* b, const double* x) {
  const int64_t n = h->n;
  std::vector<double> r(b, b + n);
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = h->colptr[c]; p < h->colptr[c + 1]; ++p) {
      const int64_t q = h->rowidx[p];
      r[q] -= h->val[p] * x[c];
      if (q != c) r[c] -= h->val[p] * x[q];
    }
  double rn = 0.0, bn = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    rn += r[i] * r[i];
    bn += b[i] * b[i];
  }
  return bn > 0.0 ? std::sqrt(rn / bn) : 0.0;
}

// x = A^-1 b with the direct factor (host or device vectors); returns the
// true relative residual, or NaN with x = NaN when no direct factor exists
// (IterMatrix::solve_direct, src/solver/iterative.rs:167-190)
double solve_directly(PcgHandle* h, const double* b, double* x, bool on_device) {
  const int64_t n = h->n;
  std::vector<double> hb(n), hx(n);
  if (on_device) rf::copy_to_host(hb.data(), b, n * sizeof(double));
  else std::memcpy(hb.data(), b, n * sizeof(double));
  if (!h->direct) {
    int64_t e = -1;
    h->direct.reset(direct_factor(n, h->colptr.data(), h->rowidx.data(), h->val.data(), h->nn, h->coords.data(),
                                  h->ndof.data(), &e));
    if (!h->direct) {
      std::fprintf(stderr, "warning: the direct fallback solution failed (singular matrix)\n");
      std::fill(hx.begin(), hx.end(), std::nan(""));
      if (on_device) rf::copy_to_device(x, hx.data(), n * sizeof(double));
      else std::memcpy(x, hx.data(), n * sizeof(double));
      return std::nan("");
    }
    if (!h->direct_reason) h->direct_reason = 3;
  }
  hx = hb;
  if (rustyfem_solve(h->direct.get(), 1, hx.data()) != 0) throw std::runtime_error("direct solve failed: " + rf::last_error());
  if (on_device) rf::copy_to_device(x, hx.data(), n * sizeof(double));
  else std::memcpy(x, hx.data(), n * sizeof(double));
  return lower_rel_residual(h, hb.data(), hx.data());
}

}  // namespace

PcgHandle* rustyfem_pcg_setup_ex(size_t n, const int64_t* colptr, const int64_t* rowidx, const double* val,
                                 int32_t kind, size_t nn, const double* coords, const int64_t* ndof,
                                 const rustyfem_pcg_options* opt, int64_t* err_eq) {
  set_error("");
  int64_t dummy = -1;
  if (!err_eq) err_eq = &dummy;
  *err_eq = -1;
  try {
    // the lower triangle in CSC, natural numbering (src/solver/iterative.rs:74-97)
    {
      std::atomic<int> bad{0};
      parallel_parts(static_cast<int64_t>(n), [&](int64_t c0, int64_t c1) {
        for (int64_t c = c0; c < c1 && !bad.load(std::memory_order_relaxed); ++c) {
          if (colptr[c + 1] < colptr[c]) {
            bad = 1;
            break;
          }
          for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p)
            if (rowidx[p] < c || rowidx[p] >= static_cast<int64_t>(n)) {
              bad = 2;
              break;
            }
        }
      });
      if (bad == 1) throw std::runtime_error("colptr is not monotone");
      if (bad == 2) throw std::runtime_error("rowidx outside the lower triangle");
    }
    int64_t tot = 0;
    for (size_t v = 0; v < nn; ++v) {
      if (ndof[v] < 0) throw std::runtime_error("negative ndof");
      tot += ndof[v];
    }
    if (kind == 0 && tot != static_cast<int64_t>(n)) throw std::runtime_error("sum of ndof differs from n");
    rf::PcgOptions o;
    const int32_t flags = opt ? opt->flags : 0;
    if (opt) {
      o.krylov = opt->krylov;
      o.vcycle_fp32 = opt->vcycle_fp32;
      o.smoother = opt->smoother;
      if (opt->degree > 0) o.degree = opt->degree;
      if (opt->cheb_ratio > 1.0) o.cheb_ratio = opt->cheb_ratio;
    }
    o.keep_hierarchy = (flags & RUSTYFEM_PCG_KEEP_HIERARCHY) != 0;
    o.force_amg_breakdown = (flags & RUSTYFEM_PCG_FORCE_AMG_BREAKDOWN) != 0;
    auto h = std::make_unique<PcgHandle>();
    h->n = static_cast<int64_t>(n);
    h->fallback = (flags & RUSTYFEM_PCG_NO_DIRECT_FALLBACK) == 0;
    if (h->fallback) {
      // a copy of the matrix for the direct fallbacks, in parallel parts
      h->colptr.assign(colptr, colptr + n + 1);
      h->rowidx.resize(colptr[n]);
      h->val.resize(colptr[n]);
      parallel_parts(colptr[n], [&](int64_t b, int64_t e) {
        if (e > b) {
          std::memcpy(h->rowidx.data() + b, rowidx + b, (e - b) * sizeof(int64_t));
          std::memcpy(h->val.data() + b, val + b, (e - b) * sizeof(double));
        }
      });
      h->nn = static_cast<int64_t>(nn);
      h->ndof.assign(ndof, ndof + nn);
      h->coords.assign(coords, coords + 3 * nn);
      // a diagonal entry that is not positive: diagnosis by the direct solver
      // (IterMatrix::factor, src/solver/iterative.rs:211-215)
      bool bad = false;
      for (size_t c = 0; c < n && !bad; ++c)
        bad = !(colptr[c] < colptr[c + 1] && rowidx[colptr[c]] == static_cast<int64_t>(c) && val[colptr[c]] > 0.0);
      if (bad) {
        int64_t e = -1;
        h->direct.reset(direct_factor(h->n, colptr, rowidx, val, h->nn, coords, ndof, &e));
        if (!h->direct) {
          *err_eq = e;
          if (rf::last_error().empty()) set_error("a diagonal entry is not positive; the direct factorization failed");
          return nullptr;
        }
        h->direct_reason = 1;
        return h.release();
      }
    }
    try {
      h->solver = rf::make_pcg_gpu(static_cast<int64_t>(n), colptr, rowidx, val, kind, static_cast<int64_t>(nn),
                                   coords, ndof, o);
    } catch (const rf::AmgBreakdown& ex) {
      if (!h->fallback) throw;
      // AMG could not be built: direct, with diagnosis (src/solver/iterative.rs:237-241)
      int64_t e = -1;
      h->direct.reset(direct_factor(h->n, colptr, rowidx, val, h->nn, coords, ndof, &e));
      if (!h->direct) {
        *err_eq = e;
        set_error(std::string(ex.what()) + "; the direct factorization failed");
        return nullptr;
      }
      h->direct_reason = 2;
    }
    return h.release();
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return nullptr;
}

PcgHandle* rustyfem_pcg_setup(size_t n, const int64_t* colptr, const int64_t* rowidx, const double* val, int32_t kind,
                              size_t nn, const double* coords, const int64_t* ndof, const rustyfem_pcg_options* opt) {
  return rustyfem_pcg_setup_ex(n, colptr, rowidx, val, kind, nn, coords, ndof, opt, nullptr);
}

int64_t rustyfem_pcg_solve(PcgHandle* h, const double* b, double* x, double tol, int64_t maxit, int32_t on_device,
                           double* rel_residual, int32_t* converged) {
  try {
    std::lock_guard<std::mutex> lock(h->mu);
    h->last_direct = false;
    h->last_direct_rel = -1.0;
    if (!h->solver) {
      const double rel = solve_directly(h, b, x, on_device != 0);
      h->last_direct = true;
      h->last_direct_rel = rel;
      if (rel_residual) *rel_residual = rel;
      if (converged) *converged = std::isfinite(rel) ? 1 : 0;
      return 0;
    }
    if (maxit <= 0) maxit = std::max<int64_t>(20000, h->n);
    const rf::PcgResult r = h->solver->solve(b, x, tol, maxit, on_device != 0);
    if (rel_residual) *rel_residual = r.rel_residual;
    if (converged) *converged = r.converged ? 1 : 0;
    if (!r.converged && h->fallback) {
      // not converged: solving directly (src/solver/iterative.rs:262-268)
      std::fprintf(stderr, "warning: PCG did not converge (%lld iterations, relative residual %.2e); solving directly\n",
                   static_cast<long long>(r.iterations), r.rel_residual);
      const double rel = solve_directly(h, b, x, on_device != 0);
      h->last_direct = true;
      h->last_direct_rel = rel;
      if (rel_residual) *rel_residual = rel;
      if (converged) *converged = std::isfinite(rel) ? 1 : 0;
    }
    return r.iterations;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

void rustyfem_pcg_stats(const PcgHandle* h, double* out, size_t len) {
  double v[44];
  for (double& x : v) x = 0.0;
  if (h->solver) h->solver->stats(v, 40);
  v[25] = h->solver ? v[25] : -1.0;
  v[40] = h->direct_reason;
  v[41] = h->last_direct ? 1.0 : 0.0;
  v[42] = h->last_direct_rel;
  v[43] = h->solver && h->solver->setup_on_device() ? 1.0 : 0.0;
  for (size_t i = 0; i < len && i < 44; ++i) out[i] = v[i];
}

int32_t rustyfem_pcg_level(const PcgHandle* h, int32_t level, int32_t what, int64_t* iout, int32_t* idx,
                           double* dout) {
  try {
    const rf::AmgHost* H = h->solver ? h->solver->hierarchy() : nullptr;
    if (!H || level < 0 || level > static_cast<int32_t>(H->levels.size())) return -1;
    const bool coarsest = level == static_cast<int32_t>(H->levels.size());
    const rf::Csr& A = coarsest ? H->coarse : H->levels[level].A;
    const std::vector<int64_t>& bptr = coarsest ? H->coarse_bptr : H->levels[level].bptr;
    const std::vector<double>& B = coarsest ? H->coarse_B : H->levels[level].B;
    const int64_t nbm = A.nrows > 0 ? static_cast<int64_t>(B.size()) / A.nrows : 0;
    auto csr = [&](const rf::Csr& m) {
      if (iout) std::copy(m.ptr.begin(), m.ptr.end(), iout);
      if (idx) std::copy(m.idx.begin(), m.idx.end(), idx);
      if (dout) std::copy(m.val.begin(), m.val.end(), dout);
      return 0;
    };
    switch (what) {
      case -1: {
        int64_t info[8] = {A.nrows, A.nnz(), static_cast<int64_t>(bptr.size()) - 1, 0, nbm, 0, 0, 0};
        if (!coarsest) {
          const rf::AmgHostLevel& L = H->levels[level];
          int64_t nagg = 0;
          for (int64_t g : L.agg) nagg = std::max(nagg, g + 1);
          info[3] = nagg;
          info[5] = L.P.ncols;
          info[6] = L.Pt.nnz();
          info[7] = L.P.nnz();
        }
        if (iout) std::copy(info, info + 8, iout);
        return 0;
      }
      case 0:
        return csr(A);
      case 1:
        return coarsest ? -1 : csr(H->levels[level].Pt);
      case 2:
        return coarsest ? -1 : csr(H->levels[level].P);
      case 3:
        if (coarsest) return -1;
        if (iout) std::copy(H->levels[level].agg.begin(), H->levels[level].agg.end(), iout);
        return 0;
      case 4:
        if (iout) std::copy(bptr.begin(), bptr.end(), iout);
        return 0;
      case 5:
        if (dout) std::copy(B.begin(), B.end(), dout);
        return 0;
      default:
        return -1;
    }
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

int64_t rustyfem_amg_aggregate(size_t n, const int64_t* ptr, const int32_t* idx, const double* val, size_t nblk,
                               const int64_t* bptr, double theta, int64_t* agg) {
  try {
    rf::Csr a;
    a.nrows = a.ncols = static_cast<int64_t>(n);
    a.ptr.assign(ptr, ptr + n + 1);
    a.idx.assign(idx, idx + ptr[n]);
    a.val.assign(val, val + ptr[n]);
    const std::vector<int64_t> bp(bptr, bptr + nblk + 1);
    if (bp.back() != static_cast<int64_t>(n)) throw std::runtime_error("blocks do not cover the rows");
    int64_t nagg = 0;
    const std::vector<int64_t> g = rf::amg_aggregate(a, bp, theta, nagg);
    std::copy(g.begin(), g.end(), agg);
    return nagg;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

int32_t rustyfem_chebyshev(double hi, double ratio, int32_t degree, double* c0, double* c1, double* c2) {
  if (!(hi > 0.0) || !(ratio > 1.0) || degree < 1) return -1;
  std::vector<std::pair<double, double>> steps;
  rf::chebyshev<double>(hi, ratio, degree, *c0, steps);
  for (size_t k = 0; k < steps.size(); ++k) {
    c1[k] = steps[k].first;
    c2[k] = steps[k].second;
  }
  return 0;
}

int32_t rustyfem_pcg_precondition(PcgHandle* h, const double* r, double* z, int32_t on_device) {
  try {
    std::lock_guard<std::mutex> lock(h->mu);
    if (!h->solver) throw std::runtime_error("the handle solves directly: it has no preconditioner");
    h->solver->apply(r, z, on_device != 0);
    return 0;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

int32_t rustyfem_pcg_set_learned_smoother(PcgHandle* h, const double* params, size_t nparams, int32_t sweeps) {
  try {
    if (static_cast<int64_t>(nparams) != rf::mpnn_nparams()) throw std::runtime_error("wrong number of parameters");
    std::lock_guard<std::mutex> lock(h->mu);
    if (!h->solver) throw std::runtime_error("the handle solves directly: it has no AMG hierarchy");
    h->solver->set_learned_smoother(params, sweeps);
    return 0;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

struct MpnnHandle {
  std::unique_ptr<rf::MpnnTrainer> t;
  int64_t n = 0;
};

int64_t rustyfem_mpnn_nparams(void) { return rf::mpnn_nparams(); }

MpnnHandle* rustyfem_mpnn_create(size_t n, const int64_t* colptr, const int64_t* rowidx, const double* val) {
  set_error("");
  try {
    for (size_t c = 0; c < n; ++c) {
      if (colptr[c + 1] < colptr[c]) throw std::runtime_error("colptr is not monotone");
      for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p)
        if (rowidx[p] < static_cast<int64_t>(c) || rowidx[p] >= static_cast<int64_t>(n))
          throw std::runtime_error("rowidx outside the lower triangle");
    }
    auto h = std::make_unique<MpnnHandle>();
    h->n = static_cast<int64_t>(n);
    h->t = rf::make_mpnn_trainer(static_cast<int64_t>(n), colptr, rowidx, val);
    return h.release();
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return nullptr;
}

double rustyfem_mpnn_loss(MpnnHandle* h, const double* params, const double* e0, size_t nvec, int32_t nu,
                          double* grad) {
  try {
    return h->t->loss(params, e0, static_cast<int>(nvec), nu, grad);
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return std::nan("");
}

int64_t rustyfem_mpnn_smoother(MpnnHandle* h, const double* params, double* blocks, int64_t* ptr, int32_t* idx) {
  try {
    h->t->smoother(params, blocks, ptr, idx);
    return h->t->nnzb();
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

void rustyfem_mpnn_free(MpnnHandle* h) { delete h; }

void rustyfem_pcg_free(PcgHandle* h) { delete h; }

int64_t rustyfem_solve_lower(size_t n, const int64_t* colptr, const int64_t* rowidx, const double* val,
                             const double* b, int32_t kind, size_t nn, const double* coords, const int64_t* ndof,
                             double tol, int64_t maxit, double* x, double* rel_residual, int32_t* converged) {
  // solve_lower (src/solver/iterative.rs:275-292) has no direct fallback
  rustyfem_pcg_options o{};
  o.flags = RUSTYFEM_PCG_NO_DIRECT_FALLBACK;
  std::unique_ptr<PcgHandle> h(rustyfem_pcg_setup(n, colptr, rowidx, val, kind, nn, coords, ndof, &o));
  if (!h) return -1;
  return rustyfem_pcg_solve(h.get(), b, x, tol, maxit, 0, rel_residual, converged);
}

int64_t rustyfem_pivots(const Handle* h, int64_t* eq, double* d, double* e, int32_t* kind) {
  try {
    if (h->ind) return h->ind->pivots(eq, d, e, kind);
    if (!h->gpu) return -1;
    h->gpu->pivots(d);
    for (int64_t i = 0; i < h->n_sym; ++i) {
      eq[i] = i;
      e[i] = 0.0;
      kind[i] = 1;
    }
    return h->n_sym;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

void rustyfem_factor_stats(const Handle* h, int64_t* out) {
  if (h->ind) {
    h->ind->stats(out);
    return;
  }
  out[0] = out[1] = out[3] = out[4] = 0;
  out[2] = h->n_sym;
  out[5] = h->factor_size;
}

int32_t rustyfem_check_plan(size_t n, size_t nn, const int64_t* loc, const double* coords, const int64_t* indptr,
                            const int64_t* indices, const double* data, int32_t leaf, int32_t plant, char* msg,
                            size_t msg_len) {
  set_error("");
  try {
    Options opt;
    opt.leaf = leaf > 0 ? leaf : 16;
    int64_t e = -1;
    std::unique_ptr<Handle> h(build(n, nn, loc, coords, indptr, indices, data, Mode::Pd, opt, &e, false));
    if (!h) throw std::runtime_error("input rejected before the factorization");
    rf::Plan plan = rf::build_plan(*h->sym, h->colptr, h->rowidx);
    const std::string r = rf::check_plan(plan, plant);
    if (msg && msg_len) {
      std::strncpy(msg, r.c_str(), msg_len - 1);
      msg[msg_len - 1] = 0;
    }
    return r.empty() ? 0 : 1;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

int32_t rustyfem_analyse(size_t n, size_t nn, const int64_t* loc, const double* coords, const int64_t* indptr,
                         const int64_t* indices, const double* data, int32_t leaf, int64_t* order_out,
                         double* info_out, int64_t* err_dof) {
  *err_dof = -1;
  set_error("");
  try {
    Options opt;
    opt.leaf = leaf > 0 ? leaf : 16;
    int64_t e = -1;
    std::unique_ptr<Handle> h(build(n, nn, loc, coords, indptr, indices, data, Mode::Pd, opt, &e, false));
    if (!h) {
      *err_dof = e;
      return 1;
    }
    if (order_out) rustyfem_order(h.get(), order_out);
    if (info_out) rustyfem_info(h.get(), info_out);
    return 0;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

struct AssemblyHandle {
  std::unique_ptr<rf::GpuAssembly> a;
};

AssemblyHandle* rustyfem_assembly_create(int32_t etype, size_t nelem, const int64_t* conn, size_t nn,
                                         const double* coords, const int64_t* eq, size_t n) {
  set_error("");
  try {
    if (etype < RUSTYFEM_C3D8 || etype > RUSTYFEM_CPE4) throw std::runtime_error("unknown element type");
    auto h = std::make_unique<AssemblyHandle>();
    h->a = rf::make_gpu_assembly(static_cast<int64_t>(nelem), static_cast<rf::ElemType>(etype), conn,
                                 static_cast<int64_t>(nn), coords, eq, static_cast<int64_t>(n));
    return h.release();
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return nullptr;
}

int64_t rustyfem_assembly_nnz(const AssemblyHandle* a) { return a->a->nnz(); }

void rustyfem_assembly_pattern(const AssemblyHandle* a, int64_t* indptr, int64_t* indices) {
  const auto& p = a->a->indptr();
  const auto& i = a->a->indices();
  std::copy(p.begin(), p.end(), indptr);
  std::copy(i.begin(), i.end(), indices);
}

int64_t rustyfem_assembly_values(AssemblyHandle* a, const double* young, const double* nu, const double* thickness,
                                 double* data, int32_t on_device) {
  set_error("");
  try {
    return a->a->values(young, nu, thickness, data, on_device != 0);
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -2;
}

int32_t rustyfem_assembly_refactor(AssemblyHandle* a, Handle* h, const double* young, const double* nu,
                                   const double* thickness, int64_t* err_dof, int64_t* err_elem) {
  *err_dof = -1;
  *err_elem = -1;
  set_error("");
  try {
    if (h->mode != Mode::Pd || !h->gpu) throw std::runtime_error("assembly refactor needs a positive definite handle");
    if (a->a->bound_owner() != h->serial) {
      if (h->indptr != a->a->indptr() || h->indices != a->a->indices())
        throw std::runtime_error("the handle was factored on another pattern than the assembly's");
      for (int64_t i = 0; i < h->n; ++i)
        if (!h->coupled[i]) throw std::runtime_error("the handle has an uncoupled DOF");
      a->a->bind(h->vmap, static_cast<int64_t>(h->rowidx.size()), h->serial);
    }
    const int64_t bad = a->a->values_bound(young, nu, thickness);
    if (bad >= 0) {
      *err_elem = bad;
      return 2;
    }
    // the numeric step of a positive definite refactorization from device values
    h->gpu64.reset();
    const int64_t e = h->gpu->factor(a->a->bound_values(), true);
    if (e >= 0) {
      *err_dof = h->owner[e];
      return 1;
    }
    return 0;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

void rustyfem_assembly_stats(const AssemblyHandle* a, double* out, size_t len) {
  a->a->stats(out, static_cast<int>(std::min<size_t>(len, 6)));
}

void rustyfem_assembly_free(AssemblyHandle* a) { delete a; }

}  // extern "C"
