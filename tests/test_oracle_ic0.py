"""The oracle's IC(0) preconditioner (src/solver/iterative/csr.rs:219-324),
pinned by what the mathematics fixes: IC(0) reproduces A on its own pattern,
it is the exact LDL^T factor when the pattern has no fill, Kershaw's matrix
breaks it without a shift, and its application inverts L D L^T."""
import numpy as np
import scipy.sparse as sp

from rustyfem_amd import meshes


def lower_csc(a):
    low = sp.csc_matrix(sp.tril(sp.csc_matrix(a)))
    low.sort_indices()
    return low.indptr.astype(np.int64), low.indices.astype(np.int64), low.data.copy()


def ldlt(n, ptr, idx, lval, d):
    l = sp.csr_matrix((lval, idx, ptr), shape=(n, n)).toarray() + np.eye(n)
    return l @ np.diag(d) @ l.T


def kershaw():
    """Kershaw (1978): SPD, and incomplete Cholesky without fill meets a
    negative pivot (d_3 = -5 by hand: d = 3, 5/3, 3/5, 3 - 4/3 - 20/3)."""
    return np.array([[3., -2., 0., 2.], [-2., 3., -2., 0.], [0., -2., 3., -2.], [2., 0., -2., 3.]])


def test_ic0_reproduces_a_on_its_pattern(ora):
    p = meshes.tet_block(3, seed=5)
    a = p.to_scipy().tocsr()
    n = p.n
    ptr, idx, lval, d, shift = ora.ic0_factor(n, *lower_csc(a))
    assert shift == 0.0 and np.all(d > 0)
    m = ldlt(n, ptr, idx, lval, d)
    dense = a.toarray()
    rows = np.repeat(np.arange(n), np.diff(ptr))
    scale = np.abs(dense).max()
    # strictly lower pattern and diagonal: (L D L^T)_ij = A_ij
    assert np.max(np.abs(m[rows, idx] - dense[rows, idx])) < 1e-12 * scale
    assert np.max(np.abs(np.diag(m) - np.diag(dense))) < 1e-12 * scale
    # and the pattern of L is exactly the strictly lower pattern of A
    la = sp.tril(a, -1).tocsr()
    la.sort_indices()
    assert np.array_equal(ptr, la.indptr) and np.array_equal(idx, la.indices)


def test_ic0_is_exact_without_fill(ora):
    rng = np.random.default_rng(7)
    n = 200
    off = -rng.uniform(0.1, 1.0, n - 1)
    diag = np.abs(np.concatenate([[0], off])) + np.abs(np.concatenate([off, [0]])) + rng.uniform(0.01, 0.1, n)
    a = sp.diags([off, diag, off], [-1, 0, 1]).toarray()
    r = rng.standard_normal(n)
    z = ora.precond_apply(n, *lower_csc(a), r, kind="ic0")
    x = np.linalg.solve(a, r)
    assert np.max(np.abs(z - x)) < 1e-10 * np.max(np.abs(x))


def test_ic0_kershaw_breakdown_shift(ora):
    a = kershaw()
    assert np.linalg.eigvalsh(a).min() > 0
    ptr, idx, lval, d, shift = ora.ic0_factor(4, *lower_csc(a))
    # retried with A + shift diag(A), shift = 1e-3 4^k (csr.rs:250-256)
    k = np.log(shift / 1e-3) / np.log(4.0)
    assert shift > 0 and abs(k - round(k)) < 1e-9
    assert np.all(d > 0)
    m = ldlt(4, ptr, idx, lval, d)
    want = a + shift * np.diag(np.diag(a))
    rows = np.repeat(np.arange(4), np.diff(ptr))
    assert np.allclose(m[rows, idx], want[rows, idx], rtol=0, atol=1e-12)
    assert np.allclose(np.diag(m), np.diag(want), rtol=0, atol=1e-12)
    # without the shift the last pivot is negative: 3 - (2/3)^2 3 - (10/3)^2 (3/5)
    assert abs((3 - 4 / 9 * 3 - 100 / 9 * 3 / 5) - (-5.0)) < 1e-12


def test_ic0_apply_inverts_the_factor(ora):
    p = meshes.hex_block(3)
    a = p.to_scipy().tocsr()
    n = p.n
    low = lower_csc(a)
    ptr, idx, lval, d, shift = ora.ic0_factor(n, *low)
    r = np.random.default_rng(2).standard_normal(n)
    z = ora.precond_apply(n, *low, r, kind="ic0")
    assert np.max(np.abs(ldlt(n, ptr, idx, lval, d) @ z - r)) < 1e-9 * np.max(np.abs(r))


def test_jacobi_apply(ora):
    p = meshes.hex_block(3)
    a = p.to_scipy().tocsr()
    r = np.random.default_rng(1).standard_normal(p.n)
    z = ora.precond_apply(p.n, *lower_csc(a), r, kind="jacobi")
    assert np.allclose(z, r / a.diagonal(), rtol=1e-15, atol=0)
