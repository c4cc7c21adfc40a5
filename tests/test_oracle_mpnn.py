"""The CPU restatement of the learned smoother (oracle/mpnn.py), pinned by what
its definition fixes: M is block-symmetric (M_ij = M_ji^T), a zero output
layer gives scaled damped Jacobi (M_ii = omega (3 / trace A_ii) I, zero off
the diagonal), and the sparse loss equals the dense evaluation
mean_v ||(I - M A)^nu e_v||_A^2 / ||e_v||_A^2 written independently here."""
import numpy as np

from oracle import mpnn
from rustyfem_amd import meshes


def lower_of(p):
    import scipy.sparse as sp
    a = p.to_scipy().tocoo()
    k = a.row >= a.col
    low = sp.csc_matrix((a.data[k], (a.row[k], a.col[k])), shape=a.shape)
    low.sort_indices()
    return low.indptr.astype(np.int64), low.indices.astype(np.int64), low.data.copy()


def _case(nx=3, seed=2):
    p = meshes.tet_block(nx, seed=seed)
    ptr, idx, blocks = mpnn.bsr_from_lower(p.n, *lower_of(p))
    return p, ptr, idx, blocks


def test_block_symmetry():
    p, ptr, idx, blocks = _case()
    m = mpnn.smoother(mpnn.init_params(1, out_scale=0.3), ptr, idx, blocks)
    mm = mpnn.bsr_matrix(ptr, idx, m).toarray()
    assert np.allclose(mm, mm.T, rtol=0, atol=1e-15 * np.abs(mm).max())


def test_zero_output_layer_is_scaled_jacobi():
    p, ptr, idx, blocks = _case()
    prm = mpnn.init_params(3)
    prm[-(9 * mpnn.H + 9):] = 0.0
    m = mpnn.smoother(prm, ptr, idx, blocks)
    nb = len(ptr) - 1
    row = np.repeat(np.arange(nb), np.diff(ptr))
    for e in range(len(idx)):
        if row[e] == idx[e]:
            want = mpnn.OMEGA * 3.0 / np.trace(blocks[e]) * np.eye(3)
            assert np.allclose(m[e], want, rtol=1e-14, atol=0)
        else:
            assert not np.any(m[e])


def test_loss_matches_dense_evaluation():
    p, ptr, idx, blocks = _case(nx=3, seed=5)
    prm = mpnn.init_params(4, out_scale=0.05)
    rng = np.random.default_rng(0)
    e0 = rng.standard_normal((p.n, 3))
    a = p.to_scipy().toarray()
    m = mpnn.bsr_matrix(ptr, idx, mpnn.smoother(prm, ptr, idx, blocks)).toarray()
    s = np.eye(p.n) - m @ a
    ratios = []
    for v in range(3):
        e = s @ (s @ e0[:, v])
        ratios.append((e @ a @ e) / (e0[:, v] @ a @ e0[:, v]))
    assert np.isclose(mpnn.loss(prm, ptr, idx, blocks, e0, nu=2), np.mean(ratios), rtol=1e-12)
    assert mpnn.NPARAMS == 3753
