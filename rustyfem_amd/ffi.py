"""ctypes binding of librustyfem_ffi.so -- the drop-in C ABI of RustyFEM's ffi/
library (ffi/src/lib.rs:64-407) served by HIP kernels on an MI355X.

The class mirrors what a ctypes caller of the reference library does: the
same eight ``rustyfem_*`` functions with the same argument meaning, error
codes and in-place solves, plus the additive ``_ex``/``refactor``/``stats``
entry points.  The library is built by ``rustyfem_amd.build``; importing the
binding without it raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librustyfem_ffi.so")

EXPORTS = ["rustyfem_factor", "rustyfem_factor_indef", "rustyfem_factor_unsym", "rustyfem_inertia",
           "rustyfem_solve", "rustyfem_info", "rustyfem_order", "rustyfem_free", "rustyfem_factor_ex",
           "rustyfem_refactor", "rustyfem_solve_device", "rustyfem_solve_device_noalloc", "rustyfem_stats", "rustyfem_last_error", "rustyfem_analyse",
           "rustyfem_check_plan", "rustyfem_pivots", "rustyfem_factor_stats", "rustyfem_pcg_setup",
           "rustyfem_pcg_solve", "rustyfem_pcg_stats", "rustyfem_pcg_free", "rustyfem_solve_lower",
           "rustyfem_factor_batch", "rustyfem_refactor_batch", "rustyfem_solve_batch", "rustyfem_solve_batch_device",
           "rustyfem_batch_info", "rustyfem_batch_stats", "rustyfem_batch_order", "rustyfem_free_batch",
           "rustyfem_pcg_set_learned_smoother", "rustyfem_mpnn_nparams", "rustyfem_mpnn_create", "rustyfem_mpnn_loss",
           "rustyfem_mpnn_smoother", "rustyfem_mpnn_free", "rustyfem_pcg_precondition", "rustyfem_chebyshev",
           "rustyfem_pcg_setup_ex", "rustyfem_pcg_level", "rustyfem_amg_aggregate", "rustyfem_eigen",
           "rustyfem_assembly_create", "rustyfem_assembly_nnz", "rustyfem_assembly_pattern", "rustyfem_assembly_values",
           "rustyfem_assembly_refactor", "rustyfem_assembly_stats", "rustyfem_assembly_free"]

_lib = None


class PcgOptions(C.Structure):
    """rustyfem_pcg_options (include/rustyfem.h)."""
    _fields_ = [("krylov", C.c_int32), ("vcycle_fp32", C.c_int32), ("smoother", C.c_int32), ("degree", C.c_int32),
                ("cheb_ratio", C.c_double), ("flags", C.c_int32), ("reserved", C.c_int32)]


PCG_KEEP_HIERARCHY = 1
PCG_NO_DIRECT_FALLBACK = 2
PCG_FORCE_AMG_BREAKDOWN = 4


class Options(C.Structure):
    """rustyfem_options (include/rustyfem.h)."""
    _fields_ = [("precision", C.c_int32), ("leaf", C.c_int32), ("max_refine", C.c_int32), ("reserved", C.c_int32),
                ("refine_tol", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m rustyfem_amd.build` "
                               "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, sz, i64p, f64p = C.c_void_p, C.c_size_t, C.POINTER(C.c_int64), C.POINTER(C.c_double)
        fac = [sz, sz, vp, vp, vp, vp, vp]
        L.rustyfem_factor.restype = vp
        L.rustyfem_factor.argtypes = fac + [vp, i64p]
        L.rustyfem_factor_ex.restype = vp
        L.rustyfem_factor_ex.argtypes = fac + [C.POINTER(Options), vp, i64p]
        L.rustyfem_factor_indef.restype = vp
        L.rustyfem_factor_indef.argtypes = fac + [vp, i64p]
        L.rustyfem_factor_unsym.restype = vp
        L.rustyfem_factor_unsym.argtypes = fac + [i64p]
        L.rustyfem_inertia.restype = C.c_int32
        L.rustyfem_inertia.argtypes = [vp, vp]
        L.rustyfem_solve.restype = C.c_int32
        L.rustyfem_solve.argtypes = [vp, sz, vp]
        L.rustyfem_solve_device.restype = C.c_int32
        L.rustyfem_solve_device.argtypes = [vp, sz, vp]
        L.rustyfem_eigen.restype = C.c_int64
        L.rustyfem_eigen.argtypes = [vp, vp, vp, vp, sz, sz, sz, C.c_double, vp, vp, i64p]
        L.rustyfem_solve_device_noalloc.restype = C.c_int32
        L.rustyfem_solve_device_noalloc.argtypes = [vp, sz, vp]
        L.rustyfem_info.restype = None
        L.rustyfem_info.argtypes = [vp, vp]
        L.rustyfem_order.restype = None
        L.rustyfem_order.argtypes = [vp, vp]
        L.rustyfem_free.restype = None
        L.rustyfem_free.argtypes = [vp]
        L.rustyfem_refactor.restype = C.c_int32
        L.rustyfem_refactor.argtypes = [vp, vp, i64p]
        L.rustyfem_stats.restype = None
        L.rustyfem_stats.argtypes = [vp, vp, sz]
        L.rustyfem_last_error.restype = C.c_char_p
        L.rustyfem_last_error.argtypes = []
        L.rustyfem_analyse.restype = C.c_int32
        L.rustyfem_analyse.argtypes = fac + [C.c_int32, vp, vp, i64p]
        L.rustyfem_check_plan.restype = C.c_int32
        L.rustyfem_check_plan.argtypes = fac + [C.c_int32, C.c_int32, C.c_char_p, sz]
        L.rustyfem_pivots.restype = C.c_int64
        L.rustyfem_pivots.argtypes = [vp, vp, vp, vp, vp]
        L.rustyfem_factor_stats.restype = None
        L.rustyfem_factor_stats.argtypes = [vp, vp]
        L.rustyfem_pcg_setup.restype = vp
        L.rustyfem_pcg_setup.argtypes = [sz, vp, vp, vp, C.c_int32, sz, vp, vp, C.POINTER(PcgOptions)]
        L.rustyfem_pcg_setup_ex.restype = vp
        L.rustyfem_pcg_setup_ex.argtypes = [sz, vp, vp, vp, C.c_int32, sz, vp, vp, C.POINTER(PcgOptions), i64p]
        L.rustyfem_pcg_level.restype = C.c_int32
        L.rustyfem_pcg_level.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp]
        L.rustyfem_amg_aggregate.restype = C.c_int64
        L.rustyfem_amg_aggregate.argtypes = [sz, vp, vp, vp, sz, vp, C.c_double, vp]
        L.rustyfem_pcg_solve.restype = C.c_int64
        L.rustyfem_pcg_solve.argtypes = [vp, vp, vp, C.c_double, C.c_int64, C.c_int32, f64p, C.POINTER(C.c_int32)]
        L.rustyfem_pcg_stats.restype = None
        L.rustyfem_pcg_stats.argtypes = [vp, vp, sz]
        L.rustyfem_chebyshev.restype = C.c_int32
        L.rustyfem_chebyshev.argtypes = [C.c_double, C.c_double, C.c_int32, vp, vp, vp]
        L.rustyfem_pcg_precondition.restype = C.c_int32
        L.rustyfem_pcg_precondition.argtypes = [vp, vp, vp, C.c_int32]
        L.rustyfem_pcg_free.restype = None
        L.rustyfem_pcg_free.argtypes = [vp]
        L.rustyfem_pcg_set_learned_smoother.restype = C.c_int32
        L.rustyfem_pcg_set_learned_smoother.argtypes = [vp, vp, sz, C.c_int32]
        L.rustyfem_mpnn_nparams.restype = C.c_int64
        L.rustyfem_mpnn_nparams.argtypes = []
        L.rustyfem_mpnn_create.restype = vp
        L.rustyfem_mpnn_create.argtypes = [sz, vp, vp, vp]
        L.rustyfem_mpnn_loss.restype = C.c_double
        L.rustyfem_mpnn_loss.argtypes = [vp, vp, vp, sz, C.c_int32, vp]
        L.rustyfem_mpnn_smoother.restype = C.c_int64
        L.rustyfem_mpnn_smoother.argtypes = [vp, vp, vp, vp, vp]
        L.rustyfem_mpnn_free.restype = None
        L.rustyfem_mpnn_free.argtypes = [vp]
        pp = C.POINTER(C.c_void_p)
        L.rustyfem_factor_batch.restype = vp
        L.rustyfem_factor_batch.argtypes = [sz, vp, vp, pp, pp, pp, pp, pp, C.POINTER(Options), vp]
        L.rustyfem_refactor_batch.restype = C.c_int32
        L.rustyfem_refactor_batch.argtypes = [vp, pp, vp]
        L.rustyfem_solve_batch.restype = C.c_int32
        L.rustyfem_solve_batch.argtypes = [vp, sz, pp]
        L.rustyfem_solve_batch_device.restype = C.c_int32
        L.rustyfem_solve_batch_device.argtypes = [vp, sz, vp]
        L.rustyfem_batch_info.restype = None
        L.rustyfem_batch_info.argtypes = [vp, vp]
        L.rustyfem_batch_stats.restype = None
        L.rustyfem_batch_stats.argtypes = [vp, vp, sz]
        L.rustyfem_batch_order.restype = None
        L.rustyfem_batch_order.argtypes = [vp, sz, vp]
        L.rustyfem_free_batch.restype = None
        L.rustyfem_free_batch.argtypes = [vp]
        L.rustyfem_solve_lower.restype = C.c_int64
        L.rustyfem_solve_lower.argtypes = [sz, vp, vp, vp, vp, C.c_int32, sz, vp, vp, C.c_double, C.c_int64, vp, f64p,
                                           C.POINTER(C.c_int32)]
        L.rustyfem_assembly_create.restype = vp
        L.rustyfem_assembly_create.argtypes = [C.c_int32, sz, vp, sz, vp, vp, sz]
        L.rustyfem_assembly_nnz.restype = C.c_int64
        L.rustyfem_assembly_nnz.argtypes = [vp]
        L.rustyfem_assembly_pattern.restype = None
        L.rustyfem_assembly_pattern.argtypes = [vp, vp, vp]
        L.rustyfem_assembly_values.restype = C.c_int64
        L.rustyfem_assembly_values.argtypes = [vp, vp, vp, vp, vp, C.c_int32]
        L.rustyfem_assembly_refactor.restype = C.c_int32
        L.rustyfem_assembly_refactor.argtypes = [vp, vp, vp, vp, vp, i64p, i64p]
        L.rustyfem_assembly_stats.restype = None
        L.rustyfem_assembly_stats.argtypes = [vp, vp, sz]
        L.rustyfem_assembly_free.restype = None
        L.rustyfem_assembly_free.argtypes = [vp]
        _lib = L
    return _lib


def last_error() -> str:
    return lib().rustyfem_last_error().decode()


def analyse(p, leaf=16):
    """Host-only symbolic analysis (rustyfem_analyse): (order, info, err_dof)."""
    keep, args = _arrays(p)
    order = np.zeros(max(int(p.n), 1), dtype=np.int64)
    info = np.zeros(2)
    err = C.c_int64(0)
    rc = lib().rustyfem_analyse(*args, int(leaf), order.ctypes.data, info.ctypes.data, C.byref(err))
    if rc < 0:
        raise RuntimeError("rustyfem_analyse failed: " + last_error())
    return order[:int(p.n)], info, int(err.value)


def check_plan(p, leaf=16, plant=0):
    """Static hazard check of the GPU schedule (rustyfem_check_plan): '' when clean."""
    keep, args = _arrays(p)
    buf = C.create_string_buffer(512)
    rc = lib().rustyfem_check_plan(*args, int(leaf), int(plant), buf, 512)
    if rc < 0:
        raise RuntimeError("rustyfem_check_plan failed: " + last_error())
    return buf.value.decode()


def _arrays(p):
    keep = (np.ascontiguousarray(p.loc, dtype=np.int64), np.ascontiguousarray(np.asarray(p.coords).reshape(-1),
                                                                              dtype=np.float64),
            np.ascontiguousarray(p.indptr, dtype=np.int64), np.ascontiguousarray(p.indices, dtype=np.int64),
            np.ascontiguousarray(p.data, dtype=np.float64))
    return keep, [int(p.n), int(p.nn)] + [a.ctypes.data for a in keep]


_STATS_KEYS = ["factor_ms", "solve_ms", "refine_ms", "refine_steps", "rel_residual", "flops", "device_bytes",
               "precision", "fronts", "levels", "gemm_ms", "gemm_launches", "gemm_flops", "tune_ms",
               "tuned_launches", "agpr_launches"]


def _stats_dict(fn, h):
    """rustyfem_stats / rustyfem_batch_stats as a dict (26 entries, rustyfem.h)."""
    o = np.zeros(26)
    fn(h, o.ctypes.data, 26)
    d = dict(zip(_STATS_KEYS, o.tolist()))
    d["refine_history"] = [float(v) for v in o[16:20] if v >= 0]
    d["ilv_launches"] = float(o[20])
    d["refine_undone"] = bool(o[21])
    d["fp64_fallback"] = bool(o[22])
    d["fallback_rel_residual"] = float(o[23])
    d["q_launches"] = float(o[24])
    d["small_levels"] = int(o[25])
    return d


class Handle:
    """A factorization held by the library (ffi/src/lib.rs:44-55)."""

    def __init__(self, ptr, n, indptr=None, indices=None):
        self._h = ptr
        self.n = n
        self.indptr = indptr
        self.indices = indices

    @classmethod
    def factor(cls, p, want_d=False, precision=None, leaf=0, max_refine=-1, refine_tol=0.0):
        """rustyfem_factor (precision None) or rustyfem_factor_ex.  Returns (handle|None, err_dof, d)."""
        keep, args = _arrays(p)
        err = C.c_int64(0)
        d = np.zeros(max(int(p.n), 1)) if want_d else None
        dptr = d.ctypes.data if want_d else None
        if precision is None:
            h = lib().rustyfem_factor(*args, dptr, C.byref(err))
        else:
            o = Options(1 if precision in (1, "mixed") else 0, int(leaf), int(max_refine), 0, float(refine_tol))
            h = lib().rustyfem_factor_ex(*args, C.byref(o), dptr, C.byref(err))
        hd = cls(h, int(p.n), keep[2], keep[3]) if h else None
        return hd, int(err.value), (d[:int(p.n)] if want_d else None)

    @classmethod
    def factor_indef(cls, p):
        keep, args = _arrays(p)
        err = C.c_int64(0)
        ine = np.zeros(3, dtype=np.int64)
        h = lib().rustyfem_factor_indef(*args, ine.ctypes.data, C.byref(err))
        return (cls(h, int(p.n)) if h else None), int(err.value), ine

    @classmethod
    def factor_unsym(cls, p):
        keep, args = _arrays(p)
        err = C.c_int64(0)
        h = lib().rustyfem_factor_unsym(*args, C.byref(err))
        return (cls(h, int(p.n)) if h else None), int(err.value)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().rustyfem_free(self._h)
            self._h = None

    def free(self):
        self.__del__()

    def solve(self, b):
        b = np.asarray(b, dtype=np.float64)
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        flat = np.array(b.reshape(-1, order="F"), dtype=np.float64, copy=True)
        rc = lib().rustyfem_solve(self._h, nrhs, flat.ctypes.data)
        if rc != 0:
            raise RuntimeError("rustyfem_solve failed: " + last_error())
        return flat.reshape(b.shape, order="F")

    def solve_device(self, ptr, nrhs):
        rc = lib().rustyfem_solve_device(self._h, int(nrhs), ptr)
        if rc != 0:
            raise RuntimeError("rustyfem_solve_device failed: " + last_error())

    def solve_device_noalloc(self, ptr, nrhs):
        """rustyfem_solve_device_noalloc; returns True when the refinement met
        its target (always True without refinement)."""
        rc = lib().rustyfem_solve_device_noalloc(self._h, int(nrhs), ptr)
        if rc < 0:
            raise RuntimeError("rustyfem_solve_device_noalloc failed: " + last_error())
        return rc == 0

    def eigen(self, b, nev, nvec=0, max_iter=0, tol=0.0, vectors=True):
        """rustyfem_eigen with B a scipy sparse symmetric matrix (caller numbering):
        (theta, x or None, iterations)."""
        import scipy.sparse as sp
        bc = sp.csc_matrix(b)
        bc.sort_indices()
        ip = np.ascontiguousarray(bc.indptr, dtype=np.int64)
        ix = np.ascontiguousarray(bc.indices, dtype=np.int64)
        dd = np.ascontiguousarray(bc.data, dtype=np.float64)
        th = np.zeros(max(int(nev), 1))
        x = np.zeros((self.n, max(int(nev), 1)), order="F") if vectors else None
        its = C.c_int64(0)
        k = lib().rustyfem_eigen(self._h, ip.ctypes.data, ix.ctypes.data, dd.ctypes.data, int(nev), int(nvec),
                                 int(max_iter), float(tol), th.ctypes.data, x.ctypes.data if vectors else None,
                                 C.byref(its))
        if k < 0:
            raise RuntimeError("rustyfem_eigen failed: " + last_error())
        return th[:k], (x[:, :k] if vectors else None), int(its.value)

    def refactor(self, data=None):
        err = C.c_int64(0)
        if data is None:
            rc = lib().rustyfem_refactor(self._h, None, C.byref(err))
        else:
            d = np.ascontiguousarray(data, dtype=np.float64)
            rc = lib().rustyfem_refactor(self._h, d.ctypes.data, C.byref(err))
        if rc < 0:
            raise RuntimeError("rustyfem_refactor failed: " + last_error())
        return int(err.value) if rc == 1 else -1

    def inertia(self):
        o = np.zeros(3, dtype=np.int64)
        rc = lib().rustyfem_inertia(self._h, o.ctypes.data)
        return o if rc == 0 else None

    def info(self):
        o = np.zeros(2)
        lib().rustyfem_info(self._h, o.ctypes.data)
        return o

    def order(self):
        o = np.zeros(max(self.n, 1), dtype=np.int64)
        lib().rustyfem_order(self._h, o.ctypes.data)
        return o[:self.n]

    def pivots(self):
        """(equation, d, e, kind) of every pivot in front postorder."""
        n = max(self.n, 1)
        eq = np.zeros(n, dtype=np.int64)
        d = np.zeros(n)
        e = np.zeros(n)
        kind = np.zeros(n, dtype=np.int32)
        c = lib().rustyfem_pivots(self._h, eq.ctypes.data, d.ctypes.data, e.ctypes.data, kind.ctypes.data)
        if c < 0:
            raise RuntimeError(last_error())
        return eq[:c], d[:c], e[:c], kind[:c]

    def factor_stats(self):
        """neg, zero, pos, delayed, two_by_two, factor_size."""
        o = np.zeros(6, dtype=np.int64)
        lib().rustyfem_factor_stats(self._h, o.ctypes.data)
        return o

    def stats(self):
        return _stats_dict(lib().rustyfem_stats, self._h)


KIND = {"amg": 0, "ic0": 1, "jacobi": 2}


class Pcg:
    """GPU preconditioned Krylov solver on a lower-triangle CSC matrix in natural
    numbering (rustyfem_pcg_setup / _solve)."""

    def __init__(self, n, colptr, rowidx, val, kind="amg", coords=None, ndof=None, krylov="pcg", vcycle="fp64",
                 smoother="chebyshev", degree=2, cheb_ratio=20.0, flags=0):
        self.n = int(n)
        self._keep = [np.ascontiguousarray(colptr, dtype=np.int64), np.ascontiguousarray(rowidx, dtype=np.int64),
                      np.ascontiguousarray(val, dtype=np.float64)]
        if coords is None:
            coords = np.zeros((self.n, 3))
            ndof = np.ones(self.n, dtype=np.int64)
        c = np.ascontiguousarray(coords, dtype=np.float64)
        nd = np.ascontiguousarray(ndof, dtype=np.int64)
        self._keep += [c, nd]
        o = PcgOptions(1 if krylov == "fcg" else 0, 1 if vcycle == "fp32" else 0, 1 if smoother == "jacobi" else 0,
                       int(degree), float(cheb_ratio), int(flags), 0)
        err = C.c_int64(-1)
        self._h = lib().rustyfem_pcg_setup_ex(self.n, self._keep[0].ctypes.data, self._keep[1].ctypes.data,
                                              self._keep[2].ctypes.data, KIND[kind], c.shape[0], c.ctypes.data,
                                              nd.ctypes.data, C.byref(o), C.byref(err))
        self.err_eq = int(err.value)
        if not self._h:
            raise RuntimeError(last_error())

    def solve(self, b, tol=1e-10, maxit=0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n)
        rel = C.c_double(0.0)
        conv = C.c_int32(0)
        it = lib().rustyfem_pcg_solve(self._h, b.ctypes.data, x.ctypes.data, float(tol), int(maxit), 0, C.byref(rel),
                                      C.byref(conv))
        if it < 0:
            raise RuntimeError(last_error())
        return x, int(it), float(rel.value), bool(conv.value)

    def solve_device(self, d_b, d_x, tol=1e-10, maxit=0):
        rel = C.c_double(0.0)
        conv = C.c_int32(0)
        it = lib().rustyfem_pcg_solve(self._h, d_b, d_x, float(tol), int(maxit), 1, C.byref(rel), C.byref(conv))
        if it < 0:
            raise RuntimeError(last_error())
        return int(it), float(rel.value), bool(conv.value)

    def level(self, l):
        """Host copy of AMG level l (RUSTYFEM_PCG_KEEP_HIERARCHY): A, P~, P as
        scipy CSR matrices, the aggregates, block boundaries and near-null space
        B; the coarsest level carries only A, bptr and B."""
        import scipy.sparse as sp
        info = np.zeros(8, dtype=np.int64)
        if lib().rustyfem_pcg_level(self._h, int(l), -1, info.ctypes.data, None, None) != 0:
            return None
        rows, nnz_a, nblk, nagg, nbm, ncols, nnz_pt, nnz_p = (int(v) for v in info)

        def csr(what, nr, nc, nnz):
            ptr = np.zeros(nr + 1, dtype=np.int64)
            idx = np.zeros(max(nnz, 1), dtype=np.int32)
            val = np.zeros(max(nnz, 1))
            if lib().rustyfem_pcg_level(self._h, int(l), what, ptr.ctypes.data, idx.ctypes.data, val.ctypes.data):
                return None
            return sp.csr_matrix((val[:nnz], idx[:nnz], ptr), shape=(nr, nc))

        out = {"A": csr(0, rows, rows, nnz_a), "nblk": nblk, "nagg": nagg, "nbm": nbm}
        bptr = np.zeros(nblk + 1, dtype=np.int64)
        lib().rustyfem_pcg_level(self._h, int(l), 4, bptr.ctypes.data, None, None)
        out["bptr"] = bptr
        B = np.zeros(max(rows * nbm, 1))
        lib().rustyfem_pcg_level(self._h, int(l), 5, None, None, B.ctypes.data)
        out["B"] = B[:rows * nbm].reshape(rows, nbm)
        if ncols > 0:
            out["Pt"] = csr(1, rows, ncols, nnz_pt)
            out["P"] = csr(2, rows, ncols, nnz_p)
            agg = np.zeros(nblk, dtype=np.int64)
            lib().rustyfem_pcg_level(self._h, int(l), 3, agg.ctypes.data, None, None)
            out["agg"] = agg
        return out

    def stats(self):
        o = np.zeros(44)
        lib().rustyfem_pcg_stats(self._h, o.ctypes.data, 44)
        lv = int(o[2])
        return {"setup_ms": o[0], "solve_ms": o[1], "levels": lv, "complexity": o[3], "spmv_ms": o[4],
                "spmv_bytes": o[5], "coarse": int(o[6]), "iterations": int(o[7]),
                "sizes": [int(v) for v in o[8:8 + min(lv, 12)]], "ic0_shift": o[20],
                "ic0_forward_levels": int(o[21]), "ic0_backward_levels": int(o[22]), "ic0_attempts": int(o[23]),
                "fallback": bool(o[24]), "true_rel_residual": o[25], "abandoned_iterations": int(o[26]),
                "learned_sweeps": int(o[27]), "blocks": [int(b) for b in o[28:28 + min(lv, 12)]],
                "direct_reason": int(o[40]), "last_direct": bool(o[41]), "direct_rel_residual": o[42],
                "setup_on_device": bool(o[43])}

    def precondition(self, r):
        """z = M r, one application of the preconditioner."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.n)
        if lib().rustyfem_pcg_precondition(self._h, r.ctypes.data, z.ctypes.data, 0) != 0:
            raise RuntimeError(last_error())
        return z

    def precondition_device(self, d_r, d_z):
        """z = M r on device pointers."""
        if lib().rustyfem_pcg_precondition(self._h, d_r, d_z, 1) != 0:
            raise RuntimeError(last_error())

    def set_learned_smoother(self, params, sweeps=1):
        """Finest-level smoother x += M (b - A x) with M from the learned network."""
        params = np.ascontiguousarray(params, dtype=np.float64)
        if lib().rustyfem_pcg_set_learned_smoother(self._h, params.ctypes.data, params.size, int(sweeps)) != 0:
            raise RuntimeError(last_error())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().rustyfem_pcg_free(self._h)
            self._h = None


def chebyshev(hi, ratio, degree):
    """(c0, c1, c2) of the V-cycle's Chebyshev smoother (rustyfem_chebyshev)."""
    c0 = np.zeros(1)
    c1 = np.zeros(max(1, degree - 1))
    c2 = np.zeros(max(1, degree - 1))
    if lib().rustyfem_chebyshev(float(hi), float(ratio), int(degree), c0.ctypes.data, c1.ctypes.data,
                                c2.ctypes.data) != 0:
        raise ValueError("invalid Chebyshev interval or degree")
    return float(c0[0]), c1[:degree - 1], c2[:degree - 1]


def mpnn_nparams():
    return int(lib().rustyfem_mpnn_nparams())


class Mpnn:
    """Training handle of the learned smoother for one matrix (lower-triangle CSC,
    natural node-major numbering, 3 DOFs per node)."""

    def __init__(self, n, colptr, rowidx, val):
        self.n = int(n)
        self._keep = [np.ascontiguousarray(colptr, np.int64), np.ascontiguousarray(rowidx, np.int64),
                      np.ascontiguousarray(val, np.float64)]
        self._h = lib().rustyfem_mpnn_create(self.n, *[a.ctypes.data for a in self._keep])
        self._keep = None
        if not self._h:
            raise RuntimeError(last_error())

    def loss(self, params, e0, nu=2, grad=False):
        """Smoothing loss for the columns of e0 (n x nvec); with grad=True also dL/dparams."""
        params = np.ascontiguousarray(params, dtype=np.float64)
        e0 = np.asfortranarray(e0, dtype=np.float64).reshape(self.n, -1, order="F")
        g = np.zeros(params.size) if grad else None
        flat = np.ascontiguousarray(e0.reshape(-1, order="F"))
        val = lib().rustyfem_mpnn_loss(self._h, params.ctypes.data, flat.ctypes.data, e0.shape[1], int(nu),
                                       g.ctypes.data if grad else None)
        if not np.isfinite(val):
            raise RuntimeError(last_error())
        return (val, g) if grad else val

    def smoother(self, params):
        """(ptr, idx, blocks (nnzb, 3, 3)) of M."""
        params = np.ascontiguousarray(params, dtype=np.float64)
        nnzb = lib().rustyfem_mpnn_smoother(self._h, params.ctypes.data, None, None, None)
        if nnzb < 0:
            raise RuntimeError(last_error())
        blocks = np.zeros(nnzb * 9)
        ptr = np.zeros(self.n // 3 + 1, dtype=np.int64)
        idx = np.zeros(nnzb, dtype=np.int32)
        lib().rustyfem_mpnn_smoother(self._h, params.ctypes.data, blocks.ctypes.data, ptr.ctypes.data, idx.ctypes.data)
        return ptr, idx, blocks.reshape(nnzb, 3, 3)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().rustyfem_mpnn_free(self._h)
            self._h = None


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


class Batch:
    """Independent positive definite problems factored and solved together
    (rustyfem_factor_batch)."""

    def __init__(self, problems, precision=None, max_refine=-1, refine_tol=0.0):
        self.problems = list(problems)
        k = len(self.problems)
        self.ns = np.array([int(p.n) for p in self.problems], dtype=np.uint64)
        nns = np.array([int(p.nn) for p in self.problems], dtype=np.uint64)
        self._keep = [_arrays(p)[0] for p in self.problems]
        cols = list(zip(*self._keep))  # loc, coords, indptr, indices, data
        self._ptr = [_ptrs(list(c)) for c in cols]
        self._nns = nns
        o = Options(1 if precision in (1, "mixed") else 0, 0, int(max_refine), 0, float(refine_tol))
        self.err = np.full(max(k, 1), -1, dtype=np.int64)
        self._h = lib().rustyfem_factor_batch(k, self.ns.ctypes.data, nns.ctypes.data, *self._ptr, C.byref(o),
                                              self.err.ctypes.data)
        self.err = self.err[:k]
        self.off = np.concatenate([[0], np.cumsum(self.ns.astype(np.int64))])

    @property
    def ok(self):
        return bool(self._h)

    def refactor(self, datas=None):
        err = np.full(len(self.problems), -1, dtype=np.int64)
        if datas is None:
            rc = lib().rustyfem_refactor_batch(self._h, None, err.ctypes.data)
        else:
            arrs = [np.ascontiguousarray(d, dtype=np.float64) for d in datas]
            rc = lib().rustyfem_refactor_batch(self._h, _ptrs(arrs), err.ctypes.data)
        if rc < 0:
            raise RuntimeError(last_error())
        return rc, err

    def solve(self, rhs):
        """rhs: list of (n_i, nrhs) arrays; returns the solutions."""
        nrhs = rhs[0].shape[1] if rhs[0].ndim == 2 else 1
        bs = [np.asfortranarray(np.asarray(b, dtype=np.float64).reshape(len(b), -1)).copy(order="F") for b in rhs]
        if lib().rustyfem_solve_batch(self._h, nrhs, _ptrs(bs)) != 0:
            raise RuntimeError(last_error())
        return bs

    def solve_device(self, d_b, nrhs):
        if lib().rustyfem_solve_batch_device(self._h, int(nrhs), d_b) != 0:
            raise RuntimeError(last_error())

    def info(self):
        o = np.zeros(5)
        lib().rustyfem_batch_info(self._h, o.ctypes.data)
        return {"problems": int(o[0]), "n": int(o[1]), "fronts": int(o[2]), "levels": int(o[3]), "work": o[4]}

    def stats(self):
        return _stats_dict(lib().rustyfem_batch_stats, self._h)

    def order(self, i):
        o = np.zeros(max(int(self.ns[i]), 1), dtype=np.int64)
        lib().rustyfem_batch_order(self._h, i, o.ctypes.data)
        return o[:int(self.ns[i])]

    def __del__(self):
        if getattr(self, "_h", None):
            lib().rustyfem_free_batch(self._h)
            self._h = None


def amg_aggregate(a, bptr, theta):
    """Vanek's three passes of one level (rustyfem_amg_aggregate) on a full
    symmetric scipy matrix with DOF blocks bptr: (aggregate per block, count)."""
    import scipy.sparse as sp
    m = sp.csr_matrix(a)
    m.sort_indices()
    ptr = np.ascontiguousarray(m.indptr, dtype=np.int64)
    idx = np.ascontiguousarray(m.indices, dtype=np.int32)
    val = np.ascontiguousarray(m.data, dtype=np.float64)
    bp = np.ascontiguousarray(bptr, dtype=np.int64)
    agg = np.zeros(bp.size - 1, dtype=np.int64)
    n = lib().rustyfem_amg_aggregate(m.shape[0], ptr.ctypes.data, idx.ctypes.data, val.ctypes.data, bp.size - 1,
                                     bp.ctypes.data, float(theta), agg.ctypes.data)
    if n < 0:
        raise RuntimeError(last_error())
    return agg, int(n)


ELEMENT_TYPES = {"C3D8": 1, "C3D4": 2, "CPS4": 3, "CPE4": 4}


class DistortedElement(ValueError):
    """An element whose Jacobian determinant fails the reference's test
    (src/elements/common.rs:15); .element is its index."""

    def __init__(self, element):
        super().__init__(f"element {element} is distorted or inverted (negative Jacobian)")
        self.element = element


class Assembly:
    """GPU element assembly of a mesh (rustyfem_assembly_*): the pattern once,
    then the stiffness values for any per-element material."""

    def __init__(self, etype, conn, coords, eq, n):
        code = ELEMENT_TYPES[etype]
        self.etype = etype
        self._conn = np.ascontiguousarray(conn, dtype=np.int64)
        self._coords = np.ascontiguousarray(coords, dtype=np.float64)
        self._eq = np.ascontiguousarray(eq, dtype=np.int64)
        self.nelem = self._conn.shape[0]
        self.n = int(n)
        self._h = lib().rustyfem_assembly_create(code, self.nelem, self._conn.ctypes.data, self._coords.shape[0],
                                                 self._coords.ctypes.data, self._eq.ctypes.data, self.n)
        if not self._h:
            raise RuntimeError("rustyfem_assembly_create failed: " + last_error())
        self.nnz = int(lib().rustyfem_assembly_nnz(self._h))
        self.indptr = np.zeros(self.n + 1, dtype=np.int64)
        self.indices = np.zeros(max(self.nnz, 1), dtype=np.int64)
        lib().rustyfem_assembly_pattern(self._h, self.indptr.ctypes.data, self.indices.ctypes.data)
        self.indices = self.indices[:self.nnz]

    def _material(self, young, nu, thickness):
        y = np.ascontiguousarray(np.broadcast_to(np.asarray(young, dtype=np.float64), (self.nelem,)))
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(nu, dtype=np.float64), (self.nelem,)))
        t = None if thickness is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(thickness, dtype=np.float64), (self.nelem,)))
        return y, v, t

    def values(self, young, nu, thickness=None, out=None):
        """The nnz values (host array), or into device memory at the address
        `out` (nnz doubles)."""
        y, v, t = self._material(young, nu, thickness)
        tp = None if t is None else t.ctypes.data
        if out is not None:
            addr = out.value if isinstance(out, C.c_void_p) else int(out)
            r = lib().rustyfem_assembly_values(self._h, y.ctypes.data, v.ctypes.data, tp, addr, 1)
            data = out
        else:
            data = np.zeros(max(self.nnz, 1))
            r = lib().rustyfem_assembly_values(self._h, y.ctypes.data, v.ctypes.data, tp, data.ctypes.data, 0)
            data = data[:self.nnz]
        if r == -2:
            raise RuntimeError("rustyfem_assembly_values failed: " + last_error())
        if r >= 0:
            raise DistortedElement(int(r))
        return data

    def refactor(self, handle, young, nu, thickness=None):
        """Assemble into handle's factorization layout and refactor it: -1, or
        the caller DOF of a failing pivot."""
        y, v, t = self._material(young, nu, thickness)
        ed, ee = C.c_int64(0), C.c_int64(0)
        rc = lib().rustyfem_assembly_refactor(self._h, handle._h, y.ctypes.data, v.ctypes.data,
                                              None if t is None else t.ctypes.data, C.byref(ed), C.byref(ee))
        if rc < 0:
            raise RuntimeError("rustyfem_assembly_refactor failed: " + last_error())
        if rc == 2:
            raise DistortedElement(int(ee.value))
        return int(ed.value) if rc == 1 else -1

    def stats(self):
        o = np.zeros(6)
        lib().rustyfem_assembly_stats(self._h, o.ctypes.data, 6)
        return {"element_ms": o[0], "gather_ms": o[1], "setup_s": o[2], "contributions": int(o[3]),
                "doubles_per_element": int(o[4]), "bound_gather_ms": o[5]}

    def __del__(self):
        if getattr(self, "_h", None):
            lib().rustyfem_assembly_free(self._h)
            self._h = None
