"""Indefinite LDL^T and LU on the GPU through the C ABI, against the oracle and
the reference's known answers (ffi/src/lib.rs:429-477,
src/steps/nlgeom_solver.rs:144-183, examples/solver_bench.rs:122-141).

The GPU factorization keeps the CPU algorithm's operation order for every
matrix entry, so the pivot sequence (equations, D and 2x2 off-diagonals or the
U diagonal, kinds, the roots' row permutations), the inertia, the delayed and
2x2 counts and the factor size must equal the oracle's bit for bit.  The
solves sum in another order: solutions are compared to the residual
(relative 1e-9 or the reference's own 1e-10 absolute bound) and to the
oracle's solution relative to the conditioning."""
import math
from types import SimpleNamespace

import numpy as np
import pytest

from rustyfem_amd import meshes
from tests._problems import (bipartite_grid, chain_problem, csc, dense_problem, laplacian3d, laplacian_sigma,
                             unsym_grid)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rf():
    from rustyfem_amd import ffi
    ffi.lib()
    return ffi


def same_bits(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def assert_same_factor(hg, ho):
    eg, dg, xg, kg = hg.pivots()
    eo, do, xo, ko = ho.pivots()
    assert np.array_equal(eg, eo), "pivot order differs"
    assert np.array_equal(kg, ko), "pivot kinds differ"
    assert same_bits(dg, do), f"D differs: max rel {np.max(np.abs(dg - do) / np.maximum(np.abs(do), 1e-300))}"
    assert same_bits(xg, xo), "2x2 off-diagonals / permutation differ"
    assert np.array_equal(hg.factor_stats(), ho.stats())
    assert np.array_equal(hg.info(), ho.info())


def relres(p, x, b):
    a = csc(p)
    return np.linalg.norm(b - a @ x, axis=0) / np.linalg.norm(b, axis=0)


def test_ffi_known_answers(rf, ora):
    """ffi/src/lib.rs:429-477 on the GPU: inertia of tridiag(-1, 2 - 1.3, -1),
    the unsymmetric convection-diffusion chain, the singular DOF 7."""
    n, sigma = 40, 1.3
    p = chain_problem(n, lambda c: 2.0 - sigma, lambda r, c: -1.0)
    h, err, ine = rf.Handle.factor_indef(p)
    assert h is not None, rf.last_error()
    below = sum(1 for k in range(1, n + 1) if 2.0 - 2.0 * math.cos(k * math.pi / (n + 1)) < sigma)
    assert list(ine) == [below, 0, n - below]
    assert list(h.inertia()) == list(ine)
    ho, _, _ = ora.Handle.factor_indef(p)
    assert_same_factor(h, ho)
    b0 = np.array([math.sin(i * 0.3) for i in range(n)])
    x = h.solve(b0)
    for i in range(n):
        ax = (2.0 - sigma) * x[i] - (x[i - 1] if i > 0 else 0.0) - (x[i + 1] if i + 1 < n else 0.0)
        assert abs(ax - b0[i]) < 1e-10
    p = chain_problem(n, lambda c: 2.0, lambda r, c: -1.3 if r > c else -0.7)
    h, err = rf.Handle.factor_unsym(p)
    assert h is not None, rf.last_error()
    assert h.inertia() is None
    ho, _ = ora.Handle.factor_unsym(p)
    assert_same_factor(h, ho)
    x = h.solve(b0)
    for i in range(n):
        ax = 2.0 * x[i] - (1.3 * x[i - 1] if i > 0 else 0.0) - (0.7 * x[i + 1] if i + 1 < n else 0.0)
        assert abs(ax - b0[i]) < 1e-10
    p = chain_problem(n, lambda c: 0.0 if c == 7 else 1.0, lambda r, c: 0.0)
    h, err, _ = rf.Handle.factor_indef(p)
    assert h is None and err == 7


def test_nlgeom_dual(rf, ora):
    """src/steps/nlgeom_solver.rs:144-183: two negative eigenvalues, x recovered."""
    a = np.array([[2.0, 1.0, 0.0, 0.0], [1.0, -3.0, 1.0, 0.0], [0.0, 1.0, 1.0, 2.0], [0.0, 0.0, 2.0, 1.0]])
    p = dense_problem(a)
    h, err, ine = rf.Handle.factor_indef(p)
    assert h is not None, rf.last_error()
    assert ine[0] == 2
    x = np.array([1.0, -2.0, 0.5, 3.0])
    assert np.max(np.abs(h.solve(a @ x) - x)) < 1e-12
    ho, _, _ = ora.Handle.factor_indef(p)
    assert_same_factor(h, ho)


@pytest.mark.parametrize("g", [6, 8, 12])
def test_laplacian_inertia(rf, ora, g):
    """Shifted 3D Laplacian: analytic inertia (examples/solver_bench.rs:122-141)."""
    print(f"case laplacian g={g}", flush=True)
    sigma, below = laplacian_sigma(g)
    p = laplacian3d(g, sigma)
    h, err, ine = rf.Handle.factor_indef(p)
    assert h is not None, rf.last_error()
    assert list(ine) == [below, 0, g ** 3 - below]
    ho, _, _ = ora.Handle.factor_indef(p)
    assert_same_factor(h, ho)
    b = np.sin(np.arange(p.n) * 0.37)
    x = h.solve(b)
    assert relres(p, x, b) < 1e-9
    xo = ho.solve(b)
    assert np.max(np.abs(x - xo)) / np.max(np.abs(xo)) < 1e-8


def test_laplacian_40_pivot_sequence_and_inertia(rf, ora):
    """The reference's indefinite benchmark family at 40^3 (64,000 equations,
    examples/solver_bench.rs:87-157): the pivot sequence equals the oracle's
    bit for bit and the inertia is the analytic count of eigenvalues below
    sigma (6,406)."""
    import time
    sigma, below = meshes.laplacian_sigma(40)
    p = meshes.laplacian3d(40, sigma)
    print(f"case laplacian g=40 n={p.n} sigma={sigma:.6f} analytic negatives {below}", flush=True)
    t0 = time.perf_counter()
    h, err, ine = rf.Handle.factor_indef(p)
    print(f"GPU analysis + factorization {time.perf_counter() - t0:.2f}s", flush=True)
    assert h is not None, rf.last_error()
    assert list(ine) == [below, 0, p.n - below] and below == 6406
    ho, _, ineo = ora.Handle.factor_indef(p)
    assert list(ineo) == list(ine)
    assert_same_factor(h, ho)
    x = h.solve(p.rhs)
    assert relres(p, x, p.rhs)[0] < 1e-9


@pytest.mark.parametrize("g,dscale", [(5, 1e-3), (6, 1e-2), (8, 1e-3)])
def test_delayed_and_2x2_pivots(rf, ora, g, dscale):
    """Bipartite grid with a tiny diagonal: 2x2 pivots and delayed pivots."""
    print(f"case bipartite g={g} dscale={dscale}", flush=True)
    p = bipartite_grid(g, 1, dscale)
    h, err, ine = rf.Handle.factor_indef(p)
    assert h is not None, rf.last_error()
    ho, _, ineo = ora.Handle.factor_indef(p)
    assert list(ine) == list(ineo)
    st = h.factor_stats()
    assert st[4] > 0
    assert_same_factor(h, ho)
    b = np.cos(np.arange(p.n) * 0.21)
    x = h.solve(b)
    assert relres(p, x, b) < 1e-9


def test_delays_reach_many_levels(rf, ora):
    """A larger strongly indefinite grid with delays at several levels."""
    p = bipartite_grid(10, 3, 1e-3)
    h, _, _ = rf.Handle.factor_indef(p)
    assert h is not None, rf.last_error()
    ho, _, _ = ora.Handle.factor_indef(p)
    assert h.factor_stats()[3] > 0
    assert_same_factor(h, ho)
    b = np.sin(np.arange(p.n) * 0.11)
    assert relres(p, h.solve(b), b) < 1e-9


def test_singular_root_reports_oracle_equation(rf, ora):
    """Zero diagonal on an odd bipartite grid is singular: the roots' Bunch-Kaufman
    step reports the same equation (as a caller DOF) as the oracle."""
    p = bipartite_grid(5, 0, 0.0)
    h, err, _ = rf.Handle.factor_indef(p)
    ho, erro, _ = ora.Handle.factor_indef(p)
    assert h is None and ho is None
    assert err == erro >= 0


@pytest.mark.parametrize("g,zf", [(6, 0.0), (6, 0.6), (6, 1.0), (8, 0.3), (4, 1.0)])
def test_lu_matches_oracle(rf, ora, g, zf):
    """LU with diagonal threshold pivoting, delays, and partial pivoting at the root."""
    print(f"case lu g={g} zero_frac={zf}", flush=True)
    p = unsym_grid(g, 0, zf)
    h, err = rf.Handle.factor_unsym(p)
    assert h is not None, rf.last_error()
    ho, _ = ora.Handle.factor_unsym(p)
    assert_same_factor(h, ho)
    if zf == 1.0:
        assert (h.pivots()[3] == 3).any()
    b = np.sin(np.arange(p.n) * 0.29)
    x = h.solve(b)
    assert relres(p, x, b) < 1e-9
    xo = ho.solve(b)
    assert np.max(np.abs(x - xo)) / np.max(np.abs(xo)) < 1e-8


def test_lu_singular(rf, ora):
    p = unsym_grid(5, 0, 1.0)
    h, err = rf.Handle.factor_unsym(p)
    ho, erro = ora.Handle.factor_unsym(p)
    assert h is None and ho is None
    assert err == erro >= 0


def test_shifted_stiffness_multi_rhs_and_refactor(rf, ora):
    """C3D8 B-bar stiffness minus a shift between two eigenvalues (3 DOFs per
    node): inertia by eigenvalue count, identical factor, 5 right-hand sides,
    and a refactorization with doubled values."""
    p0 = meshes.hex_block(5, nrhs=5, seed=7)
    a0 = p0.to_scipy()
    ev = np.linalg.eigvalsh(a0.toarray())
    q = len(ev) // 8
    while ev[q + 1] - ev[q] < 1e-6 * abs(ev[q]):
        q += 1
    sigma = 0.5 * (ev[q] + ev[q + 1])
    import scipy.sparse as sp
    a = (a0 - sigma * sp.identity(p0.n, format="csc")).tocsc()
    a.sort_indices()
    p = SimpleNamespace(n=p0.n, nn=p0.nn, loc=p0.loc, coords=p0.coords, indptr=a.indptr.astype(np.int64),
                        indices=a.indices.astype(np.int64), data=a.data.copy())
    h, err, ine = rf.Handle.factor_indef(p)
    assert h is not None, rf.last_error()
    assert list(ine) == [q + 1, 0, p.n - q - 1]
    ho, _, _ = ora.Handle.factor_indef(p)
    assert_same_factor(h, ho)
    x = h.solve(p0.rhs)
    assert np.all(relres(p, x, p0.rhs) < 1e-9)
    for j in range(p0.rhs.shape[1]):
        xj = h.solve(p0.rhs[:, j])
        assert np.allclose(xj, x[:, j], rtol=0, atol=1e-13 * np.max(np.abs(x[:, j])))
    assert h.refactor(p.data * 2.0) == -1
    x2 = h.solve(p0.rhs)
    assert np.allclose(x2, 0.5 * x, rtol=1e-10, atol=0)
