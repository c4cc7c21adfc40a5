"""The failure contract of the solver path, forced from the host.

Iterative path (IterMatrix, src/solver/iterative.rs:208-268): a diagonal
entry that is not positive sends the matrix to the direct solver for
diagnosis (:211-215), an AMG hierarchy that cannot be built makes the handle
solve directly (:237-241), and a solve that does not converge is repeated
with the direct solver (:262-268).  The direct solver here is the GPU
multifrontal LDL^T with 48-node leaves (sparse::analyse, src/sparse.rs:14),
so every fallback result is compared with rustyfem_factor_ex at leaf 48.

Direct path: a mixed-precision solve whose refinement cannot reach its
target (a tolerance below what FP64 residuals can show) is repeated with an
FP64 factorization of the same values, and reported in rustyfem_stats.
"""
import numpy as np
import pytest

from rustyfem_amd import meshes
from tests import _amg_checks as ac
from tests._problems import dense_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rf():
    from rustyfem_amd import ffi
    ffi.lib()
    return ffi


def _pcg(rf, p, **kw):
    nodes, coords, ndof, ap, ai = meshes.node_graph(p)
    cp, ri, v = ac.lower_of(p)
    return rf.Pcg(p.n, cp, ri, v, coords=coords, ndof=ndof, **kw)


def _direct48(rf, p, b):
    h, err, _ = rf.Handle.factor(p, precision="fp64", leaf=48, max_refine=0)
    assert h is not None, (err, rf.last_error())
    return h.solve(b)


def test_forced_amg_breakdown_solves_directly(rf):
    p = meshes.hex_block(10, nrhs=1)
    print(f"case amg breakdown n={p.n}", flush=True)
    s = _pcg(rf, p, flags=rf.PCG_FORCE_AMG_BREAKDOWN)
    b = p.rhs[:, 0].copy()
    x, it, rel, conv = s.solve(b, tol=1e-10)
    st = s.stats()
    assert st["direct_reason"] == 2 and st["last_direct"]
    assert conv and it == 0 and rel <= 1e-12
    xd = _direct48(rf, p, b)
    assert np.max(np.abs(x - xd)) <= 1e-13 * np.max(np.abs(xd))
    with pytest.raises(RuntimeError):
        s.precondition(b)
    # without the fallback the breakdown is an error, as in solve_lower
    with pytest.raises(RuntimeError):
        _pcg(rf, p, flags=rf.PCG_FORCE_AMG_BREAKDOWN | rf.PCG_NO_DIRECT_FALLBACK)


def test_non_convergence_falls_back_to_direct(rf):
    p = meshes.hex_block(10, nrhs=1)
    print(f"case maxit=1 n={p.n}", flush=True)
    b = p.rhs[:, 0].copy()
    s = _pcg(rf, p)
    x, it, rel, conv = s.solve(b, tol=1e-10, maxit=1)
    st = s.stats()
    assert it == 1 and conv and st["last_direct"] and st["direct_reason"] == 3
    xd = _direct48(rf, p, b)
    assert np.max(np.abs(x - xd)) <= 1e-13 * np.max(np.abs(xd))
    # the next solve converges by PCG again, with the direct factor kept
    x2, it2, rel2, conv2 = s.solve(b, tol=1e-10)
    assert conv2 and it2 > 1 and not s.stats()["last_direct"]
    assert np.max(np.abs(x2 - xd)) <= 1e-8 * np.max(np.abs(xd))
    # solve_lower's contract: the non-converged iterate is returned as such
    s2 = _pcg(rf, p, flags=rf.PCG_NO_DIRECT_FALLBACK)
    x3, it3, rel3, conv3 = s2.solve(b, tol=1e-10, maxit=1)
    assert it3 == 1 and not conv3 and not s2.stats()["last_direct"] and rel3 > 1e-10


def test_non_positive_diagonal_is_diagnosed_directly(rf, ora):
    # a zero diagonal entry: the setup reports the equation at which the
    # direct factorization stops, as the PD factorization at leaf 48 does
    import dataclasses
    p = meshes.hex_block(6, nrhs=1)
    q = 40
    data = p.data.copy()
    cols = np.repeat(np.arange(p.n), np.diff(p.indptr))
    data[(cols == q) & (p.indices == q)] = 0.0
    bad = dataclasses.replace(p, data=data)
    print(f"case zero diagonal n={p.n} eq={q}", flush=True)
    nodes, coords, ndof, ap, ai = meshes.node_graph(bad)
    cp, ri, v = ac.lower_of(bad)
    # lower_of drops explicit zeros through scipy: put the diagonal back
    import scipy.sparse as sp
    low = sp.tril(bad.to_scipy()).tocsc()
    low.sort_indices()
    assert low[q, q] == 0.0
    e48_h, e48, _ = rf.Handle.factor(bad, precision="fp64", leaf=48)
    assert e48_h is None and e48 >= 0
    s_err = _setup_err(rf, p.n, cp, ri, v, coords, ndof)
    assert s_err == e48, (s_err, e48, rf.last_error())
    # the oracle's PD factorization at leaf 48 stops at the same equation
    lib = ora.lib()
    lib.ora_set_ffi_leaf(48)
    try:
        ho, eo, _ = ora.Handle.factor(bad)
    finally:
        lib.ora_set_ffi_leaf(16)
    assert ho is None and eo == e48


def _setup_err(rf, n, cp, ri, v, coords, ndof):
    import ctypes as C
    o = rf.PcgOptions(0, 0, 0, 2, 20.0, 0, 0)
    e = C.c_int64(-1)
    keep = [np.ascontiguousarray(cp), np.ascontiguousarray(ri), np.ascontiguousarray(v),
            np.ascontiguousarray(coords, dtype=np.float64), np.ascontiguousarray(ndof, dtype=np.int64)]
    h = rf.lib().rustyfem_pcg_setup_ex(n, keep[0].ctypes.data, keep[1].ctypes.data, keep[2].ctypes.data, 0,
                                       keep[3].shape[0], keep[3].ctypes.data, keep[4].ctypes.data, C.byref(o),
                                       C.byref(e))
    assert not h
    return int(e.value)


def test_mixed_refinement_below_fp32_floor_falls_back_to_fp64(rf):
    p = meshes.hex_block(10, nrhs=2)
    print(f"case mixed fallback n={p.n}", flush=True)
    hm, em, _ = rf.Handle.factor(p, precision="mixed", refine_tol=1e-17, max_refine=30)
    assert hm is not None, rf.last_error()
    x = hm.solve(p.rhs)
    st = hm.stats()
    assert st["fp64_fallback"]
    h64, _, _ = rf.Handle.factor(p, precision="fp64", refine_tol=1e-17, max_refine=30)
    x64 = h64.solve(p.rhs)
    assert np.max(np.abs(x - x64)) <= 1e-13 * np.max(np.abs(x64))
    a = p.to_scipy()
    rel = np.linalg.norm(p.rhs - a @ x, axis=0) / np.linalg.norm(p.rhs, axis=0)
    assert np.all(rel <= 1e-13)
    # the refinement history never ends above its best value: a correction
    # that raised the residual is taken back
    assert st["rel_residual"] <= min(st["refine_history"]) * (1 + 1e-12)
    # a reachable target needs no fallback
    hm2, _, _ = rf.Handle.factor(p, precision="mixed")
    hm2.solve(p.rhs)
    assert not hm2.stats()["fp64_fallback"]
