"""Pins the oracle's element stiffness and assembly (oracle/ora_elem.c).

The element matrices are checked against the Hessian of the strain energy
written in tensor form (displacement gradients of the element's own
interpolation, the isotropic elasticity tensor, the B-bar dilatation
average) -- a formulation that shares no Voigt B matrix, D matrix or
operation order with the oracle, so a dropped term, a wrong shear factor, a
wrong weight or a transposed operand fails it.  The B-bar elements are also
checked for the rank-one volumetric part that selective reduced integration
leaves, every element for its rigid-body null space, and the assembled
matrices against the independent structured assemblies of
rustyfem_amd/meshes.py (same pattern, values to rounding).
"""
import numpy as np
import pytest

import oracle
from rustyfem_amd import meshes

HEX_NAT = meshes.HEX8_NAT
G2 = 1.0 / np.sqrt(3.0)


def _lame(e, nu):
    return e * nu / ((1 + nu) * (1 - 2 * nu)), e / (2 * (1 + nu))


def _hex_grads(x):
    """(points, weights, grad N (8 x 3)) at the 2x2x2 Gauss points."""
    out = []
    for z in (-G2, G2):
        for y in (-G2, G2):
            for xx in (-G2, G2):
                p = np.array([xx, y, z])
                f = 1 + p[None, :] * HEX_NAT
                dxi = np.stack([HEX_NAT[:, 0] * f[:, 1] * f[:, 2], HEX_NAT[:, 1] * f[:, 0] * f[:, 2],
                                HEX_NAT[:, 2] * f[:, 0] * f[:, 1]], axis=1) / 8.0
                jac = x.T @ dxi             # dx_c / dxi_r
                out.append((np.linalg.det(jac), dxi @ np.linalg.inv(jac)))
    return out


def _quad_grads(x):
    nat = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=float)
    out = []
    for eta in (-G2, G2):
        for xi in (-G2, G2):
            dxi = np.stack([nat[:, 0] * (1 + eta * nat[:, 1]), nat[:, 1] * (1 + xi * nat[:, 0])], axis=1) / 4.0
            jac = x[:, :2].T @ dxi
            out.append((np.linalg.det(jac), dxi @ np.linalg.inv(jac)))
    return out


def _strains(grad, dim):
    """Symmetric strain tensors of the unit nodal displacements: (nodes*dim, dim, dim)."""
    nn = grad.shape[0]
    eps = np.zeros((nn * dim, dim, dim))
    for a in range(nn):
        for k in range(dim):
            g = np.zeros((dim, dim))
            g[k, :] = grad[a, :dim]          # du_k/dx_c
            eps[dim * a + k] = 0.5 * (g + g.T)
    return eps


def energy_hessian(etype, x, e, nu, t=1.0):
    """Stiffness from the tensor-form strain energy."""
    lam, mu = _lame(e, nu)
    if etype == "C3D4":
        m = np.hstack([np.ones((4, 1)), x])
        grad = np.linalg.inv(m)[1:, :].T        # grad N_a = rows 1..3 of M^-1, column a
        vol = abs(np.linalg.det(m)) / 6.0
        eps = _strains(grad, 3)
        tr = np.trace(eps, axis1=1, axis2=2)
        return vol * (lam * np.outer(tr, tr) + 2 * mu * np.einsum("iab,jab->ij", eps, eps))
    if etype == "C3D8":
        pts = _hex_grads(x)
        eps = [_strains(g, 3) for _, g in pts]
        w = np.array([d for d, _ in pts])
        trs = [np.trace(ep, axis1=1, axis2=2) for ep in eps]
        tbar = sum(wi * tr for wi, tr in zip(w, trs)) / w.sum()
        k = np.zeros((24, 24))
        for wi, ep, tr in zip(w, eps, trs):
            eb = ep + ((tbar - tr) / 3.0)[:, None, None] * np.eye(3)[None]
            k += wi * (lam * np.outer(tbar, tbar) + 2 * mu * np.einsum("iab,jab->ij", eb, eb))
        return k
    pts = _quad_grads(x)
    eps = [_strains(g, 2) for _, g in pts]
    w = np.array([d for d, _ in pts]) * t
    trs = [np.trace(ep, axis1=1, axis2=2) for ep in eps]
    k = np.zeros((8, 8))
    if etype == "CPS4":
        c = e / (1 - nu * nu)
        for wi, ep, tr in zip(w, eps, trs):
            k += wi * c * ((1 - nu) * np.einsum("iab,jab->ij", ep, ep) + nu * np.outer(tr, tr))
        return k
    tbar = sum(wi * tr for wi, tr in zip(w, trs)) / w.sum()
    for wi, ep, tr in zip(w, eps, trs):
        eb = ep + ((tbar - tr) / 2.0)[:, None, None] * np.eye(2)[None]
        k += wi * (lam * np.outer(tbar, tbar) + 2 * mu * np.einsum("iab,jab->ij", eb, eb))
    return k


def random_element(etype, rng, distort=0.2):
    if etype == "C3D8":
        x = (HEX_NAT + 1) / 2 * np.array([1.0, 0.7, 1.3])
    elif etype == "C3D4":
        x = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    else:
        x = np.array([[0.0, 0, 0], [1.2, 0, 0], [1.2, 0.8, 0], [0, 0.8, 0]])
    x = x + rng.uniform(-distort, distort, x.shape)
    if etype in ("CPS4", "CPE4"):
        x[:, 2] = 0.0
    return x


TYPES = ["C3D8", "C3D4", "CPS4", "CPE4"]


@pytest.mark.parametrize("etype", TYPES)
def test_element_matches_tensor_energy(etype):
    rng = np.random.default_rng(7)
    for trial in range(10):
        x = random_element(etype, rng)
        e, nu, t = rng.uniform(1e9, 3e11), rng.uniform(0.0, 0.45), rng.uniform(0.5, 2.0)
        k = oracle.element_stiffness(etype, x, e, nu, t)
        ref = energy_hessian(etype, x, e, nu, t)
        assert k is not None
        np.testing.assert_allclose(k, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("etype", TYPES)
def test_rigid_body_null_space(etype):
    rng = np.random.default_rng(3)
    x = random_element(etype, rng)
    k = oracle.element_stiffness(etype, x, 2e11, 0.3, 1.0)
    if etype in ("CPS4", "CPE4"):
        modes = [np.tile([1, 0], 4), np.tile([0, 1], 4), np.column_stack([-x[:, 1], x[:, 0]]).ravel()]
    else:
        modes = [np.tile(v, x.shape[0]) for v in np.eye(3)]
        for axis in np.eye(3):
            modes.append(np.cross(axis[None, :], x).ravel())
    for m in modes:
        assert np.abs(k @ m).max() < 1e-12 * np.abs(k).max() * np.abs(m).max()
    # and nothing else: the element has full rank otherwise
    s = np.linalg.svd(k, compute_uv=False)
    assert (s > 1e-10 * s[0]).sum() == k.shape[0] - len(modes)


@pytest.mark.parametrize("etype", ["C3D8", "CPE4"])
def test_bbar_volumetric_part_is_rank_one(etype):
    """With the dilatation averaged over the element, the part of K that
    grows with lambda at fixed mu is V lam a a^T: rank one (full integration
    would leave rank 8 or 4)."""
    rng = np.random.default_rng(11)
    x = random_element(etype, rng)
    mu = 8e10
    ks = []
    for nu in (0.1, 0.3):
        e = 2 * mu * (1 + nu)
        ks.append(oracle.element_stiffness(etype, x, e, nu, 1.0))
    s = np.linalg.svd(ks[1] - ks[0], compute_uv=False)
    assert s[1] < 1e-10 * s[0]


def test_plane_stress_has_no_bbar():
    """CPS4 is fully integrated: its lambda-like part is not rank one."""
    rng = np.random.default_rng(2)
    x = random_element("CPS4", rng)
    e = 2e11
    d = oracle.element_stiffness("CPS4", x, e, 0.3) - oracle.element_stiffness("CPS4", x, e, 0.1)
    s = np.linalg.svd(d, compute_uv=False)
    assert s[1] > 1e-3 * s[0]


def test_constant_strain_energy_of_box_hex():
    """u = E x reproduces a uniform strain exactly: u^T K u = V eps:C:eps."""
    x = (HEX_NAT + 1) / 2 * np.array([0.5, 1.0, 2.0])
    e, nu = 1e9, 0.25
    lam, mu = _lame(e, nu)
    k = oracle.element_stiffness("C3D8", x, e, nu)
    rng = np.random.default_rng(5)
    g = rng.standard_normal((3, 3))
    u = (x @ g.T).ravel()
    eps = 0.5 * (g + g.T)
    want = 1.0 * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))
    assert u @ k @ u == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("etype", TYPES)
def test_distorted_element_is_reported(etype):
    rng = np.random.default_rng(1)
    x = random_element(etype, rng, distort=0.05)
    order = {"C3D8": [1, 0, 3, 2, 5, 4, 7, 6], "C3D4": [0, 2, 1, 3]}.get(etype, [1, 0, 3, 2])
    assert oracle.element_stiffness(etype, x[order], 2e11, 0.3) is None


def _hex_problem(nx, ny, nz):
    p = meshes.hex_block(nx, ny, nz, e=210e9, nu=0.3)
    conn = meshes.grid_elements((nx, ny, nz))
    free = np.ones((nz + 1, ny + 1, nx + 1), dtype=bool)
    free[:, :, 0] = False
    eq = meshes.equation_numbers(free, 3)
    return p, conn, eq


def test_assembled_hex_block_matches_structured_assembly():
    p, conn, eq = _hex_problem(4, 3, 2)
    ne = conn.shape[0]
    data = oracle.assemble("C3D8", conn, p.coords, eq, p.n, np.full(ne, 210e9), np.full(ne, 0.3), p.indptr,
                           p.indices)
    assert not isinstance(data, int)
    np.testing.assert_allclose(data, p.data, rtol=0, atol=1e-12 * np.abs(p.data).max())
    # the pattern holds nothing the elements do not couple
    assert np.all(data != 0.0) or np.count_nonzero(data == 0) < 0.05 * data.size


def test_assembled_tet_block_matches_structured_assembly():
    p = meshes.tet_block(4, seed=3, elements=True)
    m = p.meta
    ne = m["conn"].shape[0]
    data = oracle.assemble("C3D4", m["conn"], p.coords, m["eq"], p.n, m["young"], np.full(ne, m["nu"]), p.indptr,
                           p.indices)
    assert not isinstance(data, int)
    np.testing.assert_allclose(data, p.data, rtol=0, atol=1e-11 * np.abs(p.data).max())


def test_assembled_plate_matches_structured_assembly():
    nx, ny = 6, 3
    p = meshes.plate_cpe4(nx, ny, 7e10, 0.33, nrhs=1)
    conn = meshes.grid_elements((nx, ny))
    free = np.ones((ny + 1, nx + 1), dtype=bool)
    free[:, 0] = False
    eq = meshes.equation_numbers(free, 2)
    ne = conn.shape[0]
    data = oracle.assemble("CPE4", conn, p.coords, eq, p.n, np.full(ne, 7e10), np.full(ne, 0.33), p.indptr,
                           p.indices)
    assert not isinstance(data, int)
    np.testing.assert_allclose(data, p.data, rtol=0, atol=1e-12 * np.abs(p.data).max())


def test_assembly_reports_first_distorted_element():
    p, conn, eq = _hex_problem(3, 2, 2)
    conn = conn.copy()
    for e in (7, 4):
        conn[e] = conn[e][[1, 0, 3, 2, 5, 4, 7, 6]]
    ne = conn.shape[0]
    r = oracle.assemble("C3D8", conn, p.coords, eq, p.n, np.full(ne, 1e9), np.full(ne, 0.3), p.indptr, p.indices)
    assert r == 4
