"""The oracle's iterative path (src/solver/iterative*.rs): PCG with AMG, IC(0) and
Jacobi against the direct solve, and the AMG null space."""
import numpy as np
import pytest

from rustyfem_amd import meshes


def lower_of(p, first=None):
    """Lower triangle in node-block natural order (node order already in the problem)."""
    import scipy.sparse as sp
    a = p.to_scipy().tocoo()
    m = a.row >= a.col
    low = sp.csc_matrix((a.data[m], (a.row[m], a.col[m])), shape=a.shape)
    low.sort_indices()
    return low.indptr.astype(np.int64), low.indices.astype(np.int64), low.data.copy()


@pytest.mark.parametrize("kind", ["amg", "ic0", "jacobi"])
def test_pcg_matches_direct(ora, kind):
    p = meshes.hex_block(8, nrhs=1)
    nodes, coords, ndof, ap, ai = meshes.node_graph(p)
    colptr, rowidx, val = lower_of(p)
    b = p.rhs[:, 0].copy()
    x, rep = ora.solve_lower(p.n, colptr, rowidx, val, b, kind, coords, ndof, tol=1e-10)
    assert rep["converged"] and rep["rel_residual"] <= 1e-10
    a = p.to_scipy()
    assert np.linalg.norm(a @ x - b) / np.linalg.norm(b) < 1e-9
    h, err, _ = ora.Handle.factor(p)
    xd = h.solve(b)
    # the reference reports max relative deviation 4e-12 at tol 1e-10 (docs/REFERENCE.md:418-419)
    assert np.max(np.abs(x - xd)) / np.max(np.abs(xd)) < 1e-8
    if kind == "amg":
        assert len(rep["amg_sizes"]) >= 2 and rep["complexity"] >= 1.0


def test_natural_pattern_is_node_lower(ora):
    p = meshes.hex_block(4)
    nodes, coords, ndof, ap, ai = meshes.node_graph(p)
    first, colptr, rowidx = ora.natural_pattern(ndof, ap, ai)
    c2, r2, _ = lower_of(p)
    assert np.array_equal(colptr, c2) and np.array_equal(rowidx, r2)
    assert np.array_equal(first, np.concatenate([[0], np.cumsum(ndof)[:-1]]))


def test_rigid_body_modes_span_null_space(ora):
    """K B = 0 for the stiffness of an unconstrained block (6 modes in 3D) and a
    plane plate (3 modes in 2D) -- src/solver/iterative/amg.rs:39-95."""
    import scipy.sparse as sp
    for kind in ("3d", "2d"):
        if kind == "3d":
            p = meshes.hex_block(3)
            # unconstrained: rebuild with every node free
            ke = meshes.hex8_stiffness((meshes.HEX8_NAT + 1) / 2 / 3, meshes.dmat_3d(1.0, 0.3))
            n, fid, ip, ix, d = meshes._assemble_structured((3, 3, 3), ke, None, np.ones(64, bool), 3)
            coords = p.coords
            ndof = np.full(64, 3, dtype=np.int64)
            want = 6
        else:
            ke = meshes.quad4_stiffness(np.array([[0, 0], [.5, 0], [.5, .5], [0, .5]], float),
                                        meshes.dmat_plane_stress(1.0, 0.3), 1.0, False)
            n, fid, ip, ix, d = meshes._assemble_structured((4, 2), ke, None, np.ones(15, bool), 2)
            jj, ii = np.meshgrid(np.arange(3), np.arange(5), indexing="ij")
            coords = np.stack([ii.ravel() * .5, jj.ravel() * .5, 0 * ii.ravel()], axis=1).astype(float)
            ndof = np.full(15, 2, dtype=np.int64)
            want = 3
        k = sp.csc_matrix((d, ix, ip), shape=(n, n))
        b = ora.rigid_body_modes(coords, ndof)
        assert b.shape[1] == want
        assert np.linalg.matrix_rank(b) == want
        assert np.abs(k @ b).max() < 1e-12 * np.abs(k).max() * np.abs(b).max() * 10


def test_pcg_breakdown_on_indefinite(ora):
    """pcg reports non-convergence when p^T A p <= 0 (csr.rs:358-361)."""
    from tests._problems import laplacian3d, laplacian_sigma
    sigma, _ = laplacian_sigma(6)
    p = laplacian3d(6, sigma)
    import scipy.sparse as sp
    a = sp.csc_matrix((p.data, p.indices, p.indptr), shape=(p.n, p.n)).tocoo()
    m = a.row >= a.col
    low = sp.csc_matrix((a.data[m], (a.row[m], a.col[m])), shape=a.shape)
    low.sort_indices()
    b = np.ones(p.n)
    x, rep = ora.solve_lower(p.n, low.indptr.astype(np.int64), low.indices.astype(np.int64), low.data, b,
                             "jacobi", tol=1e-10)
    assert not rep["converged"]
