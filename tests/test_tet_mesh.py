"""The C3D4 (linear tetrahedron) test problems (src/elements/solid_higher.rs:109).

Pinned by what the element's mathematics fixes: rigid body motions give no
forces, a linear displacement field stores the strain energy V/2 eps^T D eps
of its constant strain (this fixes B, the engineering shear order of
src/elements/common.rs:160-170 and D of src/materials/elastic.rs:47-70), the
assembled matrix passes the patch test (K u = 0 at interior nodes for a
linear field on a homogeneous block), is exactly symmetric and positive
definite once clamped, and the body load integrates the cube's volume."""
import numpy as np
import pytest

from rustyfem_amd import meshes


def _rand_tets(rng, ne):
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
    return base[None] * rng.uniform(0.5, 2.0, (ne, 1, 3)) + rng.uniform(-0.1, 0.1, (ne, 4, 3))


def test_rigid_body_modes_in_null_space():
    rng = np.random.default_rng(0)
    x = _rand_tets(rng, 20)
    k, _ = meshes.tet4_stiffness(x, np.full(20, 210e9), 0.3)
    for e in range(20):
        modes = []
        for d in range(3):
            u = np.zeros((4, 3)); u[:, d] = 1.0
            modes.append(u.ravel())
        for a, b in ((0, 1), (1, 2), (0, 2)):
            u = np.zeros((4, 3)); u[:, a] = -x[e, :, b]; u[:, b] = x[e, :, a]
            modes.append(u.ravel())
        for u in modes:
            assert np.linalg.norm(k[e] @ u) <= 1e-9 * np.linalg.norm(k[e]) * np.linalg.norm(u)


def test_strain_energy_of_linear_field():
    rng = np.random.default_rng(1)
    x = _rand_tets(rng, 10)
    e, nu = 3.0, 0.27
    k, vol = meshes.tet4_stiffness(x, np.full(10, e), nu)
    d = meshes.dmat_3d(e, nu)
    for t in range(10):
        g = rng.standard_normal((3, 3))
        u = (x[t] @ g.T).ravel()          # u_i = G x
        eps = np.array([g[0, 0], g[1, 1], g[2, 2], g[0, 1] + g[1, 0], g[0, 2] + g[2, 0], g[1, 2] + g[2, 1]])
        assert np.isclose(0.5 * u @ k[t] @ u, 0.5 * vol[t] * eps @ d @ eps, rtol=1e-12)
        assert vol[t] > 0


def test_patch_test_and_symmetry():
    p = meshes.tet_block(5, log_std=0.0, seed=3, clamp=None)
    a = p.to_scipy()
    assert abs(a - a.T).max() == 0.0
    g = np.array([[1.0, 0.3, -0.2], [0.1, -0.5, 0.4], [0.25, 0.05, 0.7]])
    xyz = p.coords[p.loc[::3]]
    u = (xyz @ g.T).ravel()
    f = (a @ u).reshape(-1, 3)
    interior = np.all((xyz > 1e-12) & (xyz < 1 - 1e-12), axis=1)
    assert interior.sum() > 0
    assert np.max(np.abs(f[interior])) <= 1e-9 * np.max(np.abs(f))
    assert np.isclose(p.rhs.sum(), -1e4, rtol=1e-12)     # body 1e4 N/m^3 over the unit cube


@pytest.mark.parametrize("nx", [4, 7])
def test_clamped_block_is_positive_definite(nx):
    p = meshes.tet_block(nx, seed=11)
    a = p.to_scipy().toarray()
    assert np.array_equal(a, a.T)
    np.linalg.cholesky(a)
    assert 0.5 <= p.meta["log_std"] <= 2.0
    assert p.n == 3 * (nx + 1) ** 2 * nx
