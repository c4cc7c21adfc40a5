// The C ABI of ffi/src/lib.rs, served by the MI355X multifrontal solver.
//
// build() restates ffi/src/lib.rs:100-229 (validation, uncoupled DOFs,
// node compaction, adjacency, nested dissection with leaf 16, equation map,
// entry filter per mode); the numeric factorization and the solves run on
// the GPU (mf_gpu.hip, indef_gpu.hip).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rustyfem.h"
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

}  // namespace

struct Handle {
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
  std::unique_ptr<rf::IndefGpu> ind;
  std::unique_ptr<rf::Symbolic> sym;  // kept only by host-only analyses
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
constexpr size_t NSTATS = 21;
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

struct PcgHandle {
  std::unique_ptr<rf::PcgGpu> solver;
  int64_t n = 0;
  std::mutex mu;
};

PcgHandle* rustyfem_pcg_setup(size_t n, const int64_t* colptr, const int64_t* rowidx, const double* val, int32_t kind,
                              size_t nn, const double* coords, const int64_t* ndof, const rustyfem_pcg_options* opt) {
  set_error("");
  try {
    // the lower triangle in CSC, natural numbering (src/solver/iterative.rs:74-97)
    for (size_t c = 0; c < n; ++c) {
      if (colptr[c + 1] < colptr[c]) throw std::runtime_error("colptr is not monotone");
      for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p)
        if (rowidx[p] < static_cast<int64_t>(c) || rowidx[p] >= static_cast<int64_t>(n))
          throw std::runtime_error("rowidx outside the lower triangle");
    }
    int64_t tot = 0;
    for (size_t v = 0; v < nn; ++v) {
      if (ndof[v] < 0) throw std::runtime_error("negative ndof");
      tot += ndof[v];
    }
    if (kind == 0 && tot != static_cast<int64_t>(n)) throw std::runtime_error("sum of ndof differs from n");
    rf::PcgOptions o;
    if (opt) {
      o.krylov = opt->krylov;
      o.vcycle_fp32 = opt->vcycle_fp32;
      o.smoother = opt->smoother;
      if (opt->degree > 0) o.degree = opt->degree;
      if (opt->cheb_ratio > 1.0) o.cheb_ratio = opt->cheb_ratio;
    }
    auto h = std::make_unique<PcgHandle>();
    h->n = static_cast<int64_t>(n);
    h->solver = rf::make_pcg_gpu(static_cast<int64_t>(n), colptr, rowidx, val, kind, static_cast<int64_t>(nn), coords,
                                 ndof, o);
    return h.release();
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return nullptr;
}

int64_t rustyfem_pcg_solve(PcgHandle* h, const double* b, double* x, double tol, int64_t maxit, int32_t on_device,
                           double* rel_residual, int32_t* converged) {
  try {
    std::lock_guard<std::mutex> lock(h->mu);
    if (maxit <= 0) maxit = std::max<int64_t>(20000, h->n);
    const rf::PcgResult r = h->solver->solve(b, x, tol, maxit, on_device != 0);
    if (rel_residual) *rel_residual = r.rel_residual;
    if (converged) *converged = r.converged ? 1 : 0;
    return r.iterations;
  } catch (const std::exception& ex) {
    set_error(ex.what());
  } catch (...) {
    set_error("unknown error");
  }
  return -1;
}

void rustyfem_pcg_stats(const PcgHandle* h, double* out, size_t len) { h->solver->stats(out, len); }

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
  std::unique_ptr<PcgHandle> h(rustyfem_pcg_setup(n, colptr, rowidx, val, kind, nn, coords, ndof, nullptr));
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

}  // extern "C"
