"""rustyfem_eigen (shift-invert block subspace iteration, src/solver/eigen.rs:
258-387, on the GPU factorization and multi-RHS device solve) against the
oracle's restatement of the same algorithm and against closed forms.

The GPU sums its Gram matrices and dot products in another order than the
oracle, so the Ritz values agree to rounding (relative 1e-10) and the
iteration counts agree to two (the residual test of the last iterations
compares values near the tolerance); the eigenvalues match the analytic spectrum of
the 3D Laplacian (a triple eigenvalue included) and scipy's shift-invert
Lanczos on an elastic block with a lumped mass matrix."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from tests._eigen_problems import elastic_case, laplacian_case, laplacian_lowest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rf():
    from rustyfem_amd import ffi
    ffi.lib()
    return ffi


@pytest.mark.parametrize("precision", [None, "mixed"])
def test_laplacian_against_oracle_and_closed_form(rf, ora, precision):
    g = 10
    p, b = laplacian_case(g)
    print(f"case eigen laplacian g={g} precision={precision}", flush=True)
    h, err, _ = rf.Handle.factor(p, precision=precision)
    assert h is not None, rf.last_error()
    # the lowest eigenvalue and the triple one above it (a clean gap to the
    # next triple, so that the iteration counts do not hinge on rounding)
    theta, x, its = h.eigen(b, nev=4)
    ho, _, _ = ora.Handle.factor(p)
    tho, xo, itso = ho.eigen(b, nev=4)
    lam = laplacian_lowest(g, 4)
    print(f"theta {theta} iterations {its} (oracle {itso})", flush=True)
    assert len(theta) == 4
    assert np.allclose(1.0 / theta, lam, rtol=1e-9, atol=0)
    assert np.allclose(theta, tho, rtol=1e-10, atol=0)
    assert abs(its - itso) <= 2
    a = p.to_scipy()
    for i in range(4):
        assert abs(x[:, i] @ (a @ x[:, i]) - 1.0) < 1e-8
        r = a @ x[:, i] - x[:, i] / theta[i]
        assert np.linalg.norm(r) <= 1e-7 * np.linalg.norm(a @ x[:, i])


def test_elastic_block_against_lanczos_and_oracle(rf, ora):
    p, m = elastic_case(6)
    print(f"case eigen elastic hex6 n={p.n}", flush=True)
    h, err, _ = rf.Handle.factor(p)
    theta, x, its = h.eigen(m, nev=8)
    ho, _, _ = ora.Handle.factor(p)
    tho, _, itso = ho.eigen(m, nev=8)
    k = p.to_scipy()
    lam_ref = np.sort(sla.eigsh(k.tocsc(), k=8, M=m.tocsc(), sigma=0.0, which="LM", return_eigenvectors=False))
    assert np.allclose(np.sort(1.0 / theta), lam_ref, rtol=1e-8, atol=0)
    assert np.allclose(theta, tho, rtol=1e-10, atol=0)
    print(f"iterations {its} (oracle {itso})", flush=True)
    assert abs(its - itso) <= 2
    for i in range(8):
        r = k @ x[:, i] - (1.0 / theta[i]) * (m @ x[:, i])
        assert np.linalg.norm(r) <= 1e-7 * np.linalg.norm(k @ x[:, i])


def test_rejects_indefinite_handle(rf):
    from tests._problems import laplacian3d, laplacian_sigma
    sigma, _ = laplacian_sigma(6)
    p = laplacian3d(6, sigma)
    import scipy.sparse as sp
    h, err, _ = rf.Handle.factor_indef(p)
    with pytest.raises(RuntimeError, match="positive definite"):
        h.eigen(sp.identity(p.n, format="csc"), nev=2)
