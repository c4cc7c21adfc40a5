"""The oracle's shift-invert block subspace iteration (src/solver/eigen.rs:
258-387) and its cyclic Jacobi (src/solver/eigen.rs:159-214), pinned by
closed forms and library routines: Jacobi against numpy's symmetric
eigensolver; the subspace iteration against the analytic spectrum of the 3D
Laplacian (a triple eigenvalue included) and against scipy's shift-invert
Lanczos on an elastic block with a lumped mass matrix."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from tests._eigen_problems import elastic_case, laplacian_case, laplacian_lowest


@pytest.mark.parametrize("n", [1, 2, 5, 12])
def test_jacobi_against_numpy(ora, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    a = a + a.T
    d, v = ora.jacobi_eigen(a)
    assert np.allclose(np.sort(d), np.linalg.eigvalsh(a), rtol=0, atol=1e-12 * max(1.0, np.abs(a).max()))
    assert np.allclose(v.T @ v, np.eye(n), atol=1e-12)
    assert np.allclose(a @ v, v * d, atol=1e-11 * max(1.0, np.abs(a).max()))


def test_laplacian_closed_form(ora):
    g = 8
    p, b = laplacian_case(g)
    h, err, _ = ora.Handle.factor(p)
    assert h is not None
    theta, x, its = h.eigen(b, nev=6)
    lam = laplacian_lowest(g, 6)
    # largest theta first: theta = 1 / lambda for the smallest lambda
    assert np.allclose(1.0 / theta, lam, rtol=1e-9, atol=0)
    # x^T Op x = 1 and Op x = B x / theta
    a = p.to_scipy()
    for i in range(len(theta)):
        assert abs(x[:, i] @ (a @ x[:, i]) - 1.0) < 1e-9
        r = a @ x[:, i] - x[:, i] / theta[i]
        assert np.linalg.norm(r) <= 1e-7 * np.linalg.norm(a @ x[:, i])
    assert its >= 1


def test_elastic_lumped_mass_against_lanczos(ora):
    p, m = elastic_case(4)
    h, err, _ = ora.Handle.factor(p)
    theta, x, its = h.eigen(m, nev=5)
    k = p.to_scipy()
    lam_ref = np.sort(sla.eigsh(k.tocsc(), k=5, M=m.tocsc(), sigma=0.0, which="LM", return_eigenvectors=False))
    assert np.allclose(np.sort(1.0 / theta), lam_ref, rtol=1e-8, atol=0)


def test_dropped_mass_term_is_detected(ora):
    """A plausible slip -- taking B = I instead of the mass matrix -- moves the
    eigenvalues far outside the tolerance of the check above."""
    import scipy.sparse as sp
    p, m = elastic_case(4)
    h, err, _ = ora.Handle.factor(p)
    theta, _, _ = h.eigen(sp.identity(p.n, format="csc"), nev=5)
    k = p.to_scipy()
    lam_ref = np.sort(sla.eigsh(k.tocsc(), k=5, M=m.tocsc(), sigma=0.0, which="LM", return_eigenvectors=False))
    assert not np.allclose(np.sort(1.0 / theta), lam_ref, rtol=1e-3, atol=0)
