"""ctypes wrapper around the C oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle is the CPU restatement of the reference
algorithms that the tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use as the checker or as the timed CPU baseline.  Nothing in
the product (``rustyfem_amd``) imports it.

Parity status of each oracle function (what pins it) is listed in the header of
``ora.h`` and in the tests under ``tests/test_oracle_*.py``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
vp = C.c_void_p
sz = C.c_size_t


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sigs = {
            "ora_analyse": (vp, [sz, f64p, i64p, i64p, i64p, sz]),
            "ora_sym_free": (None, [vp]),
            "ora_sym_n": (sz, [vp]),
            "ora_sym_nfronts": (sz, [vp]),
            "ora_sym_work": (C.c_double, [vp]),
            "ora_sym_factor_size": (sz, [vp]),
            "ora_sym_fronts": (None, [vp, i64p, i64p, i64p, i64p, i64p]),
            "ora_sym_rows": (None, [vp, sz, i64p]),
            "ora_sym_first_dof": (None, [vp, i64p]),
            "ora_sym_nodes": (sz, [vp, sz, vp]),
            "ora_sym_pattern": (sz, [vp, i64p, i64p, i64p, vp, vp]),
            "ora_factor_pd": (vp, [vp, i64p, i64p, f64p, C.POINTER(C.c_int64)]),
            "ora_factor_free": (None, [vp]),
            "ora_factor_pivots": (None, [vp, f64p]),
            "ora_factor_solve": (None, [vp, f64p]),
            "ora_factor_front": (None, [vp, sz, vp, vp, vp]),
            "ora_factor_indef": (vp, [vp, i64p, i64p, f64p, C.c_double, C.POINTER(C.c_int64)]),
            "ora_ifactor_free": (None, [vp]),
            "ora_ifactor_solve": (None, [vp, f64p]),
            "ora_ifactor_stats": (None, [vp, i64p]),
            "ora_factor_lu": (vp, [vp, i64p, i64p, f64p, f64p, C.c_double, C.POINTER(C.c_int64)]),
            "ora_lufactor_free": (None, [vp]),
            "ora_lufactor_solve": (None, [vp, f64p]),
            "ora_lufactor_stats": (None, [vp, i64p]),
            "ora_ffi_factor": (vp, [sz, sz, i64p, f64p, i64p, i64p, f64p, vp, C.POINTER(C.c_int64)]),
            "ora_ffi_factor_indef": (vp, [sz, sz, i64p, f64p, i64p, i64p, f64p, vp, C.POINTER(C.c_int64)]),
            "ora_ffi_factor_unsym": (vp, [sz, sz, i64p, f64p, i64p, i64p, f64p, C.POINTER(C.c_int64)]),
            "ora_ffi_inertia": (C.c_int32, [vp, i64p]),
            "ora_ffi_solve": (C.c_int32, [vp, sz, f64p]),
            "ora_ffi_info": (None, [vp, f64p]),
            "ora_ffi_order": (None, [vp, i64p]),
            "ora_ffi_free": (None, [vp]),
            "ora_ffi_sym": (vp, [vp]),
            "ora_ffi_pivots": (sz, [vp, i64p, f64p, f64p, vp]),
            "ora_ffi_stats": (None, [vp, i64p]),
            "ora_indef_last_forced": (C.c_int64, []),
            "ora_ffi_leaf": (sz, []),
            "ora_set_ffi_leaf": (None, [sz]),
            "ora_solve_lower": (C.c_int64, [sz, i64p, i64p, f64p, f64p, C.c_int, sz, f64p, i64p, C.c_double,
                                            C.c_int64, f64p, C.POINTER(C.c_double), C.POINTER(C.c_int32), vp,
                                            C.POINTER(C.c_double)]),
            "ora_ic0_factor": (C.c_int64, [sz, i64p, i64p, f64p, vp, vp, vp, vp, C.POINTER(C.c_double)]),
            "ora_precond_apply": (None, [sz, i64p, i64p, f64p, C.c_int, f64p, f64p]),
            "ora_natural_pattern": (sz, [sz, i64p, i64p, i64p, vp, vp, vp]),
            "ora_sym_matvec": (None, [sz, i64p, i64p, f64p, f64p, f64p]),
            "ora_rigid_body_modes": (sz, [sz, f64p, i64p, f64p]),
            "ora_set_threads": (None, [C.c_int]),
            "ora_threads": (C.c_int, []),
            "ora_rsort_by_ikey": (None, [i64p, sz, vp]),
            "ora_amg_aggregate": (C.c_int64, [sz, i64p, i64p, f64p, sz, i64p, C.c_double, i64p]),
            "ora_amg_build": (vp, [sz, i64p, i64p, f64p, sz, f64p, i64p]),
            "ora_amg_nlev": (sz, [vp]),
            "ora_amg_level": (C.c_int32, [vp, sz, C.c_int32, vp, vp, vp]),
            "ora_amg_free": (None, [vp]),
            "ora_rsort_net9": (None, [i64p, vp]),
            "ora_rsort_net13": (None, [i64p, vp]),
            "ora_eigen": (C.c_int64, [vp, sz, i64p, i64p, f64p, sz, sz, sz, C.c_double, f64p, vp,
                                      C.POINTER(C.c_int64)]),
            "ora_jacobi_eigen": (None, [sz, f64p, f64p, f64p]),
            "ora_element_stiffness": (C.c_int32, [C.c_int32, f64p, C.c_double, C.c_double, C.c_double, f64p]),
            "ora_assemble": (C.c_int64, [C.c_int32, C.c_int64, i64p, C.c_int64, f64p, i64p, C.c_int64, f64p, f64p,
                                         vp, i64p, i64p, f64p]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(t: int) -> None:
    lib().ora_set_threads(int(t))


def amg_aggregate(a, bptr, theta):
    """The oracle's three-pass aggregation of one level (src/solver/iterative/amg.rs:108-181)
    on a full symmetric scipy matrix: (aggregate per block, count)."""
    import scipy.sparse as sp
    m = sp.csr_matrix(a)
    m.sort_indices()
    bp = np.ascontiguousarray(bptr, dtype=np.int64)
    agg = np.zeros(bp.size - 1, dtype=np.int64)
    n = lib().ora_amg_aggregate(m.shape[0], np.ascontiguousarray(m.indptr, dtype=np.int64),
                                np.ascontiguousarray(m.indices, dtype=np.int64),
                                np.ascontiguousarray(m.data, dtype=np.float64), bp.size - 1, bp, float(theta), agg)
    return agg, int(n)


class AmgHierarchy:
    """The oracle's Amg::new (src/solver/iterative/amg.rs:282-341) with every level recorded."""

    def __init__(self, n, colptr, rowidx, val, coords, ndof):
        c = np.ascontiguousarray(coords, dtype=np.float64)
        self._h = lib().ora_amg_build(int(n), _i64(colptr), _i64(rowidx), _f64(val), c.shape[0], c,
                                      _i64(ndof))
        if not self._h:
            raise RuntimeError("oracle AMG setup failed")
        self.nlev = int(lib().ora_amg_nlev(self._h))

    def level(self, l):
        import scipy.sparse as sp
        L = lib()
        info = np.zeros(8, dtype=np.int64)
        if L.ora_amg_level(self._h, int(l), -1, info.ctypes.data, None, None) != 0:
            return None
        rows, nnz_a, nblk, nagg, nbm, ncols, nnz_pt, nnz_p = (int(v) for v in info)

        def csr(what, nr, nc, nnz):
            ptr = np.zeros(nr + 1, dtype=np.int64)
            idx = np.zeros(max(nnz, 1), dtype=np.int64)
            val = np.zeros(max(nnz, 1))
            L.ora_amg_level(self._h, int(l), what, ptr.ctypes.data, idx.ctypes.data, val.ctypes.data)
            return sp.csr_matrix((val[:nnz], idx[:nnz], ptr), shape=(nr, nc))

        out = {"A": csr(0, rows, rows, nnz_a), "nblk": nblk, "nagg": nagg, "nbm": nbm}
        bptr = np.zeros(nblk + 1, dtype=np.int64)
        L.ora_amg_level(self._h, int(l), 4, bptr.ctypes.data, None, None)
        out["bptr"] = bptr
        B = np.zeros(max(rows * nbm, 1))
        L.ora_amg_level(self._h, int(l), 5, None, None, B.ctypes.data)
        out["B"] = B[:rows * nbm].reshape(rows, nbm)
        if ncols > 0:
            out["Pt"] = csr(1, rows, ncols, nnz_pt)
            out["P"] = csr(2, rows, ncols, nnz_p)
            agg = np.zeros(nblk, dtype=np.int64)
            L.ora_amg_level(self._h, int(l), 3, agg.ctypes.data, None, None)
            out["agg"] = agg
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ora_amg_free(self._h)
            self._h = None


def rust_sort_unstable(items, key=None):
    """Items (int64) sorted by key[item] (by their own value when key is None)
    with the restated unstable sort of Rust's core library (ora_rsort.c)."""
    v = np.array(items, dtype=np.int64, copy=True)
    k = None if key is None else np.ascontiguousarray(key, dtype=np.int64)
    lib().ora_rsort_by_ikey(v, v.size, None if k is None else k.ctypes.data)
    return v


def rust_sort_network(items, key=None):
    """One of the small sort's networks (9 or 13 inputs) applied alone."""
    v = np.array(items, dtype=np.int64, copy=True)
    k = None if key is None else np.ascontiguousarray(key, dtype=np.int64)
    f = {9: lib().ora_rsort_net9, 13: lib().ora_rsort_net13}[v.size]
    f(v, None if k is None else k.ctypes.data)
    return v


def threads() -> int:
    return int(lib().ora_threads())


class Symbolic:
    """Nested-dissection analysis (src/sparse.rs:167-232) of a node graph."""

    def __init__(self, coords, adj_ptr, adj_idx, ndof, leaf):
        self.coords = _f64(np.asarray(coords).reshape(-1, 3))
        self.adj_ptr = _i64(adj_ptr)
        self.adj_idx = _i64(adj_idx)
        self.ndof = _i64(ndof)
        self.nn = len(self.ndof)
        self._h = lib().ora_analyse(self.nn, self.coords.ravel(), self.adj_ptr, self.adj_idx, self.ndof, int(leaf))
        self._owned = True

    @classmethod
    def _borrow(cls, ptr, nn):
        s = cls.__new__(cls)
        s._h = ptr
        s._owned = False
        s.nn = nn
        return s

    def __del__(self):
        if getattr(self, "_owned", False) and self._h:
            lib().ora_sym_free(self._h)
            self._h = None

    @property
    def n(self):
        return int(lib().ora_sym_n(self._h))

    @property
    def nfronts(self):
        return int(lib().ora_sym_nfronts(self._h))

    def work(self):
        return float(lib().ora_sym_work(self._h))

    def factor_size(self):
        return int(lib().ora_sym_factor_size(self._h))

    def fronts(self):
        nf = self.nfronts
        out = [np.zeros(nf, dtype=np.int64) for _ in range(5)]
        lib().ora_sym_fronts(self._h, *out)
        return dict(zip(["start", "k", "u", "nchild", "parent"], out))

    def rows(self, f):
        u = int(self.fronts()["u"][f])
        r = np.zeros(max(u, 1), dtype=np.int64)
        lib().ora_sym_rows(self._h, int(f), r)
        return r[:u]

    def all_rows(self):
        fr = self.fronts()
        out = []
        for f in range(self.nfronts):
            r = np.zeros(max(int(fr["u"][f]), 1), dtype=np.int64)
            lib().ora_sym_rows(self._h, f, r)
            out.append(r[:int(fr["u"][f])])
        return out

    def nodes(self, f):
        cnt = int(lib().ora_sym_nodes(self._h, int(f), None))
        a = np.zeros(max(cnt, 1), dtype=np.int64)
        lib().ora_sym_nodes(self._h, int(f), a.ctypes.data)
        return a[:cnt]

    def first_dof(self):
        a = np.zeros(self.nn, dtype=np.int64)
        lib().ora_sym_first_dof(self._h, a)
        return a

    def pattern(self, adj_ptr=None, adj_idx=None, ndof=None):
        adj_ptr = self.adj_ptr if adj_ptr is None else _i64(adj_ptr)
        adj_idx = self.adj_idx if adj_idx is None else _i64(adj_idx)
        ndof = self.ndof if ndof is None else _i64(ndof)
        nnz = int(lib().ora_sym_pattern(self._h, adj_ptr, adj_idx, ndof, None, None))
        colptr = np.zeros(self.n + 1, dtype=np.int64)
        rowidx = np.zeros(max(nnz, 1), dtype=np.int64)
        lib().ora_sym_pattern(self._h, adj_ptr, adj_idx, ndof, colptr.ctypes.data, rowidx.ctypes.data)
        return colptr, rowidx[:nnz]


class PdFactor:
    """Multifrontal LDL^T of a lower CSC matrix in internal numbering (src/sparse.rs:595-653)."""

    def __init__(self, sym: Symbolic, colptr, rowidx, val):
        self.sym = sym
        err = C.c_int64(-1)
        self._h = lib().ora_factor_pd(sym._h, _i64(colptr), _i64(rowidx), _f64(val), C.byref(err))
        self.err = int(err.value)

    @property
    def ok(self):
        return bool(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ora_factor_free(self._h)
            self._h = None

    def pivots(self):
        d = np.zeros(self.sym.n)
        lib().ora_factor_pivots(self._h, d)
        return d

    def solve(self, b):
        x = _f64(b).copy()
        lib().ora_factor_solve(self._h, x)
        return x

    def front(self, f):
        fr = self.sym.fronts()
        k, u = int(fr["k"][f]), int(fr["u"][f])
        l11 = np.zeros(max(k * (k - 1) // 2, 1))
        d = np.zeros(max(k, 1))
        l21 = np.zeros(max(u * k, 1))
        lib().ora_factor_front(self._h, int(f), l11.ctypes.data, d.ctypes.data, l21.ctypes.data)
        return l11[:k * (k - 1) // 2], d[:k], l21[:u * k].reshape(u, k)


class IndefFactor:
    def __init__(self, sym: Symbolic, colptr, rowidx, val, threshold=0.01):
        self.sym = sym
        err = C.c_int64(-1)
        self._h = lib().ora_factor_indef(sym._h, _i64(colptr), _i64(rowidx), _f64(val), float(threshold),
                                         C.byref(err))
        self.err = int(err.value)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ora_ifactor_free(self._h)
            self._h = None

    def stats(self):
        o = np.zeros(6, dtype=np.int64)
        lib().ora_ifactor_stats(self._h, o)
        return dict(zip(["neg", "zero", "pos", "delayed", "two_by_two", "factor_size"], o.tolist()))

    def solve(self, b):
        x = _f64(b).copy()
        lib().ora_ifactor_solve(self._h, x)
        return x


class Handle:
    """The oracle's restatement of the ffi/ C ABI (ffi/src/lib.rs:64-407)."""

    def __init__(self, ptr, n):
        self._h = ptr
        self.n = n

    @staticmethod
    def _args(p):
        return (int(p.n), int(p.nn), _i64(p.loc), _f64(np.asarray(p.coords).reshape(-1)), _i64(p.indptr),
                _i64(p.indices), _f64(p.data))

    @classmethod
    def factor(cls, p, want_d=False):
        err = C.c_int64(0)
        d = np.zeros(max(int(p.n), 1)) if want_d else None
        h = lib().ora_ffi_factor(*cls._args(p), d.ctypes.data if want_d else None, C.byref(err))
        return (cls(h, int(p.n)) if h else None), int(err.value), (d[:int(p.n)] if want_d else None)

    @classmethod
    def factor_indef(cls, p):
        err = C.c_int64(0)
        ine = np.zeros(3, dtype=np.int64)
        h = lib().ora_ffi_factor_indef(*cls._args(p), ine.ctypes.data, C.byref(err))
        return (cls(h, int(p.n)) if h else None), int(err.value), ine

    @classmethod
    def factor_unsym(cls, p):
        err = C.c_int64(0)
        h = lib().ora_ffi_factor_unsym(*cls._args(p), C.byref(err))
        return (cls(h, int(p.n)) if h else None), int(err.value)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ora_ffi_free(self._h)
            self._h = None

    def solve(self, b):
        b = np.asarray(b, dtype=np.float64)
        x = np.array(b, dtype=np.float64, order="F", copy=True)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        flat = np.ascontiguousarray(x.reshape(-1, order="F"))
        rc = lib().ora_ffi_solve(self._h, nrhs, flat)
        assert rc == 0
        return flat.reshape(x.shape, order="F")

    def eigen(self, b, nev, nvec=0, max_iter=2000, tol=1e-10):
        """Shift-invert subspace iteration (src/solver/eigen.rs:258-387) on this
        PD handle (Op) with B a scipy sparse symmetric matrix: (theta, x, iterations)."""
        import scipy.sparse as sp
        bc = sp.csc_matrix(b)
        bc.sort_indices()
        th = np.zeros(max(int(nev), 1))
        x = np.zeros(self.n * max(int(nev), 1))
        its = C.c_int64(0)
        k = lib().ora_eigen(self._h, self.n, _i64(bc.indptr), _i64(bc.indices), _f64(bc.data), int(nev), int(nvec),
                            int(max_iter), float(tol), th, x.ctypes.data, C.byref(its))
        if k < 0:
            raise RuntimeError(f"ora_eigen failed ({k})")
        return th[:k], x[:self.n * k].reshape((self.n, k), order="F"), int(its.value)

    def inertia(self):
        o = np.zeros(3, dtype=np.int64)
        rc = lib().ora_ffi_inertia(self._h, o)
        return (o if rc == 0 else None)

    def info(self):
        o = np.zeros(2)
        lib().ora_ffi_info(self._h, o)
        return o

    def pivots(self):
        """(equation, d, e, kind) of every pivot in front postorder."""
        eq = np.zeros(max(self.n, 1), dtype=np.int64)
        d = np.zeros(max(self.n, 1))
        e = np.zeros(max(self.n, 1))
        kind = np.zeros(max(self.n, 1), dtype=np.int32)
        c = lib().ora_ffi_pivots(self._h, eq, d, e, kind.ctypes.data)
        return eq[:c], d[:c], e[:c], kind[:c]

    def stats(self):
        """neg, zero, pos, delayed, two_by_two, factor_size."""
        o = np.zeros(6, dtype=np.int64)
        lib().ora_ffi_stats(self._h, o)
        return o

    def order(self):
        o = np.zeros(max(self.n, 1), dtype=np.int64)
        lib().ora_ffi_order(self._h, o)
        return o[:self.n]

    def sym(self, nn):
        return Symbolic._borrow(lib().ora_ffi_sym(self._h), nn)


def jacobi_eigen(a):
    """The oracle's cyclic Jacobi (src/solver/eigen.rs:159-214): (d, v)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    d = np.zeros(max(n, 1))
    v = np.zeros(max(n * n, 1))
    lib().ora_jacobi_eigen(n, a.reshape(-1), d, v)
    return d[:n], v[:n * n].reshape(n, n)


ELEMENT_TYPES = {"C3D8": 1, "C3D4": 2, "CPS4": 3, "CPE4": 4}


def element_stiffness(etype, x, young, nu, thickness=1.0):
    """The oracle's element stiffness (nd x nd) of a C3D8, C3D4, CPS4 or CPE4
    with node coordinates x (nodes x 3); None for a distorted element."""
    code = ELEMENT_TYPES[etype]
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    nodes = len(x) // 3
    nd = nodes * (3 if code <= 2 else 2)
    ke = np.zeros(nd * nd)
    r = lib().ora_element_stiffness(code, x, float(young), float(nu), float(thickness), ke)
    if r == -1:
        return None
    if r != 0:
        raise ValueError(f"ora_element_stiffness failed ({r})")
    return ke.reshape(nd, nd)


def assemble(etype, conn, coords, eq, n, young, nu, indptr, indices, thickness=None):
    """The reference's element loop and scatter into the given both-triangle
    CSC pattern: the data array, or the index of the first distorted element
    (an int)."""
    code = ELEMENT_TYPES[etype]
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    eq = np.ascontiguousarray(eq, dtype=np.int64)
    young = np.ascontiguousarray(young, dtype=np.float64)
    nu = np.ascontiguousarray(nu, dtype=np.float64)
    th = None if thickness is None else np.ascontiguousarray(thickness, dtype=np.float64)
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    indices = np.ascontiguousarray(indices, dtype=np.int64)
    data = np.zeros(max(int(indptr[-1]), 1))
    r = lib().ora_assemble(code, conn.shape[0], conn.reshape(-1), coords.shape[0], coords.reshape(-1),
                           eq.reshape(-1), int(n), young, nu, None if th is None else th.ctypes.data, indptr,
                           indices, data)
    if r >= 0:
        return int(r)
    if r != -1:
        raise ValueError(f"ora_assemble failed ({r})")
    return data[:int(indptr[-1])]


def solve_lower(n, colptr, rowidx, val, b, kind="amg", coords=None, ndof=None, tol=1e-10, maxit=0):
    """PCG on a lower CSC matrix (src/solver/iterative.rs:275-292)."""
    k = {"amg": 0, "ic0": 1, "jacobi": 2}[kind]
    x = np.zeros(n)
    rel = C.c_double(0.0)
    conv = C.c_int32(0)
    sizes = np.zeros(16, dtype=np.int64)
    cx = C.c_double(0.0)
    coords = _f64(np.zeros((0, 3)) if coords is None else np.asarray(coords).reshape(-1))
    ndof = _i64(np.zeros(0) if ndof is None else ndof)
    it = lib().ora_solve_lower(n, _i64(colptr), _i64(rowidx), _f64(val), _f64(b), k, len(ndof), coords, ndof,
                               float(tol), int(maxit), x, C.byref(rel), C.byref(conv), sizes.ctypes.data,
                               C.byref(cx))
    return x, {"iterations": int(it), "rel_residual": rel.value, "converged": bool(conv.value),
               "amg_sizes": [int(s) for s in sizes if s], "complexity": cx.value}


def ic0_factor(n, colptr, rowidx, val):
    """IC(0) factor (src/solver/iterative/csr.rs:229-257): (ptr, idx, lval, d,
    shift) with the strictly lower L row-wise."""
    colptr, rowidx, val = _i64(colptr), _i64(rowidx), _f64(val)
    nnz = int(lib().ora_ic0_factor(n, colptr, rowidx, val, None, None, None, None, None))
    ptr = np.zeros(n + 1, dtype=np.int64)
    idx = np.zeros(max(nnz, 1), dtype=np.int64)
    lval = np.zeros(max(nnz, 1))
    d = np.zeros(max(n, 1))
    shift = C.c_double(0.0)
    lib().ora_ic0_factor(n, colptr, rowidx, val, ptr.ctypes.data, idx.ctypes.data, lval.ctypes.data, d.ctypes.data,
                         C.byref(shift))
    return ptr, idx[:nnz], lval[:nnz], d[:n], shift.value


def precond_apply(n, colptr, rowidx, val, r, kind="ic0"):
    """z = M r (Precond::apply, src/solver/iterative/csr.rs:193-217, 296-316)."""
    k = {"ic0": 1, "jacobi": 2}[kind]
    z = np.zeros(n)
    lib().ora_precond_apply(n, _i64(colptr), _i64(rowidx), _f64(val), k, _f64(r), z)
    return z


def natural_pattern(ndof, adj_ptr, adj_idx):
    nn = len(ndof)
    nnz = int(lib().ora_natural_pattern(nn, _i64(adj_ptr), _i64(adj_idx), _i64(ndof), None, None, None))
    n = int(np.sum(ndof))
    first = np.zeros(max(nn, 1), dtype=np.int64)
    colptr = np.zeros(n + 1, dtype=np.int64)
    rowidx = np.zeros(max(nnz, 1), dtype=np.int64)
    lib().ora_natural_pattern(nn, _i64(adj_ptr), _i64(adj_idx), _i64(ndof), first.ctypes.data, colptr.ctypes.data,
                              rowidx.ctypes.data)
    return first[:nn], colptr, rowidx[:nnz]


def sym_matvec(colptr, rowidx, val, x):
    n = len(colptr) - 1
    y = np.zeros(n)
    lib().ora_sym_matvec(n, _i64(colptr), _i64(rowidx), _f64(val), _f64(x), y)
    return y


def rigid_body_modes(coords, ndof):
    ndof = _i64(ndof)
    n = int(ndof.sum())
    b = np.zeros(n * 6)
    nb = int(lib().ora_rigid_body_modes(len(ndof), _f64(np.asarray(coords).reshape(-1)), ndof, b))
    return b[:n * nb].reshape(n, nb)
