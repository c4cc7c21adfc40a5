"""The oracle's C-ABI restatement against the reference's own known-answer tests
(ffi/src/lib.rs:429-477, src/steps/nlgeom_solver.rs:144-183,
examples/solver_bench.rs:87-157) and against dense linear algebra."""
import math

import numpy as np
import pytest

from tests._problems import chain_problem, dense_problem, laplacian3d, laplacian_sigma


def test_ffi_indefinite_unsymmetric_singular(ora):
    """ffi/src/lib.rs:429-477, restated."""
    n = 40
    sigma = 1.3
    p = chain_problem(n, lambda c: 2.0 - sigma, lambda r, c: -1.0)
    # the positive definite path must refuse
    h, err, _ = ora.Handle.factor(p)
    assert h is None
    h, err, ine = ora.Handle.factor_indef(p)
    assert h is not None
    below = sum(1 for k in range(1, n + 1) if 2.0 - 2.0 * math.cos(k * math.pi / (n + 1)) < sigma)
    assert list(ine) == [below, 0, n - below]
    b0 = np.array([math.sin(i * 0.3) for i in range(n)])
    x = h.solve(b0)
    for i in range(n):
        ax = (2.0 - sigma) * x[i] - (x[i - 1] if i > 0 else 0.0) - (x[i + 1] if i + 1 < n else 0.0)
        assert abs(ax - b0[i]) < 1e-10
    assert list(h.inertia()) == list(ine)
    # unsymmetric convection-diffusion
    p = chain_problem(n, lambda c: 2.0, lambda r, c: -1.3 if r > c else -0.7)
    h, err = ora.Handle.factor_unsym(p)
    assert h is not None
    x = h.solve(b0)
    for i in range(n):
        ax = 2.0 * x[i] - (1.3 * x[i - 1] if i > 0 else 0.0) - (0.7 * x[i + 1] if i + 1 < n else 0.0)
        assert abs(ax - b0[i]) < 1e-10
    assert h.inertia() is None
    # singular: zero matrix in one degree of freedom
    p = chain_problem(n, lambda c: 0.0 if c == 7 else 1.0, lambda r, c: 0.0)
    h, err, _ = ora.Handle.factor_indef(p)
    assert h is None and err == 7


def test_nlgeom_dual_indefinite(ora):
    """src/steps/nlgeom_solver.rs:144-183: leading minors 2, -7, -9, 19 -> PD fails,
    indefinite gives 2 negative eigenvalues and recovers x."""
    a = np.array([[2.0, 1.0, 0.0, 0.0], [1.0, -3.0, 1.0, 0.0], [0.0, 1.0, 1.0, 2.0], [0.0, 0.0, 2.0, 1.0]])
    p = dense_problem(a)
    h, err, _ = ora.Handle.factor(p)
    assert h is None and err >= 0
    h, err, ine = ora.Handle.factor_indef(p)
    assert h is not None
    assert ine[0] == 2
    x = np.array([1.0, -2.0, 0.5, 3.0])
    y = h.solve(a @ x)
    assert np.max(np.abs(y - x)) < 1e-12


@pytest.mark.parametrize("g", [6, 8, 10])
def test_laplacian_inertia_analytic(ora, g):
    """Inertia of the shifted 3D Laplacian equals the analytic count
    (examples/solver_bench.rs:122-141)."""
    sigma, below = laplacian_sigma(g)
    p = laplacian3d(g, sigma)
    h, err, ine = ora.Handle.factor_indef(p)
    assert h is not None
    assert list(ine) == [below, 0, g ** 3 - below]
    b = np.sin(np.arange(p.n) * 0.37)
    x = h.solve(b)
    import scipy.sparse as sp
    a = sp.csc_matrix((p.data, p.indices, p.indptr), shape=(p.n, p.n))
    assert np.linalg.norm(a @ x - b) / np.linalg.norm(b) < 1e-9


def test_pd_matches_dense_and_pivot_reconstruction(ora):
    """PD handle on a random SPD matrix with node blocks: solution vs numpy, pivots
    D equal to the dense LDL^T pivots in the ABI's internal order (any correct
    LDL^T of P A P^T has the same D), and log det matches slogdet."""
    rng = np.random.default_rng(5)
    nn = 30
    coords = rng.uniform(0, 1, (nn, 3))
    # couple nearby nodes, two DOFs per node
    loc = np.repeat(np.arange(nn), 2)
    n = 2 * nn
    a = np.zeros((n, n))
    for i in range(nn):
        for j in range(i + 1, nn):
            if np.linalg.norm(coords[i] - coords[j]) < 0.45:
                blk = rng.standard_normal((2, 2)) * 0.3
                a[2 * i:2 * i + 2, 2 * j:2 * j + 2] = blk
                a[2 * j:2 * j + 2, 2 * i:2 * i + 2] = blk.T
    a += np.eye(n) * (np.abs(a).sum(axis=1).max() + 1.0)
    p = dense_problem(a, coords=coords, loc=loc)
    h, err, d = ora.Handle.factor(p, want_d=True)
    assert h is not None
    b = rng.standard_normal(n)
    x = h.solve(b)
    assert np.linalg.norm(x - np.linalg.solve(a, b)) / np.linalg.norm(x) < 1e-12
    order = h.order()
    # dense LDL^T in the internal order
    perm = np.argsort(order)
    ap = a[np.ix_(perm, perm)]
    dd = np.zeros(n)
    m = ap.copy()
    for j in range(n):
        dd[j] = m[j, j]
        m[j + 1:, j + 1:] -= np.outer(m[j + 1:, j], m[j, j + 1:]) / m[j, j]
    assert np.allclose(d[perm], dd, rtol=1e-12, atol=0)
    assert abs(np.sum(np.log(d)) - np.linalg.slogdet(a)[1]) < 1e-9 * abs(np.linalg.slogdet(a)[1])


def test_pd_first_failing_pivot(ora):
    """The PD path reports the first non-positive pivot in elimination order
    (src/sparse.rs:528-531, first failure in DFS order, mapped to the caller's DOF),
    checked by a dense elimination in the internal order."""
    rng = np.random.default_rng(11)
    nn = 40
    coords = rng.uniform(0, 1, (nn, 3))
    n = nn
    a = np.zeros((n, n))
    for i in range(nn):
        for j in range(i + 1, nn):
            if np.linalg.norm(coords[i] - coords[j]) < 0.4:
                a[i, j] = a[j, i] = -rng.uniform(0.1, 1.0)
    deg = -a.sum(axis=1)
    a += np.diag(deg + 0.05)
    # make it indefinite: shift by roughly the 30 % eigenvalue
    ev = np.linalg.eigvalsh(a)
    a -= np.eye(n) * ev[n // 3]
    a[np.diag_indices(n)] += 1e-3
    p = dense_problem(a, coords=coords)
    h, err, _ = ora.Handle.factor(p)
    assert h is None
    # internal order from the indefinite handle (same ordering, PD filter is irrelevant here)
    hi, _, _ = ora.Handle.factor_indef(p)
    order = hi.order()
    perm = np.argsort(order)
    m = a[np.ix_(perm, perm)].copy()
    first = None
    for j in range(n):
        dj = m[j, j]
        orig = a[perm[j], perm[j]]
        if not (orig > 0.0) or not (dj > 1e-12 * orig):
            first = perm[j]
            break
        m[j + 1:, j + 1:] -= np.outer(m[j + 1:, j], m[j, j + 1:]) / dj
    assert first is not None and err == first


def test_uncoupled_dofs_and_order(ora):
    """Uncoupled DOFs are solved as b/diag and numbered last in caller order
    (ffi/src/lib.rs:105-128, 353-355, 390-398); a non-positive uncoupled diagonal
    is reported as its DOF (ffi/src/lib.rs:122-126)."""
    a = np.array([[4.0, 1.0, 0.0, 0.0, 0.0], [1.0, 3.0, 0.0, 0.0, 0.0], [0.0, 0.0, 2.0, 0.0, 0.0],
                  [0.0, 0.0, 0.0, 5.0, -1.0], [0.0, 0.0, 0.0, -1.0, 6.0]])
    p = dense_problem(a, coords=np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]], float), loc=[0, 0, 1, 2, 2])
    h, err, d = ora.Handle.factor(p, want_d=True)
    assert h is not None
    order = h.order()
    assert order[2] == 4  # uncoupled DOF numbered after the 4 coupled ones
    assert sorted(order[[0, 1, 3, 4]]) == [0, 1, 2, 3]
    assert d[2] == 2.0
    b = np.arange(1.0, 6.0)
    assert np.allclose(h.solve(b), np.linalg.solve(a, b), rtol=1e-13)
    a2 = a.copy()
    a2[2, 2] = -1.0
    h2, err2, _ = ora.Handle.factor(dense_problem(a2, coords=p.coords, loc=p.loc))
    assert h2 is None and err2 == 2


def test_invalid_input_reports_minus_one(ora):
    """ffi/src/lib.rs:101-104: a DOF with an out-of-range node is invalid input."""
    a = np.array([[2.0, -1.0], [-1.0, 2.0]])
    p = dense_problem(a)
    p.loc = np.array([0, 5], dtype=np.int64)
    h, err, _ = ora.Handle.factor(p)
    assert h is None and err == -1


def test_multiple_rhs(ora):
    from rustyfem_amd import meshes
    q = meshes.hex_block(6, nrhs=3)
    h, err, _ = ora.Handle.factor(q)
    x = h.solve(q.rhs)
    a = q.to_scipy()
    for j in range(3):
        xs = h.solve(q.rhs[:, j])
        assert np.array_equal(xs, x[:, j])
        assert np.linalg.norm(a @ x[:, j] - q.rhs[:, j]) / np.linalg.norm(q.rhs[:, j]) < 1e-12
