"""CPU restatement (numpy, FP64) of the learned smoother's network and of its
smoothing loss.

TEST INFRASTRUCTURE ONLY: the checker of the GPU kernels in
rustyfem_amd/csrc/mpnn_gpu.hip.  The learned smoother has no counterpart in
the reference (a new piece of the solver path); this file restates the
definition that the product's header documents:

    s_i   = sqrt(3 / trace(A_ii))
    x_e   = [s_i s_j vec(A_ij), [i == j]]            e = (i, j) a block of A
    m_e   = relu(W1 x_e + b1)
    n_i   = mean over the blocks of row i of m_e
    h_e   = relu(W2 [m_e; n_i; n_j] + b2)
    O_e   = W3 h_e + b3                              (3x3, row-major)
    M_ij  = s_i s_j ((O_e + O_rev(e)^T) / 2 + omega [i == j] I),  omega = 2/3

    L     = mean_v e_nu^T A e_nu / e_0^T A e_0,    e_{k+1} = (I - M A) e_k

Parity: pinned by the block symmetry of M, by the closed form of M when the
output layer is zero (scaled damped Jacobi), and by a dense-matrix evaluation
of the loss (tests/test_oracle_mpnn.py); the GPU forward, loss and gradient
are checked against this restatement, the gradient by central finite
differences (tests/test_gpu_mpnn.py).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

H, F = 32, 10
OMEGA = 2.0 / 3.0
NPARAMS = H * F + H + H * 3 * H + H + 9 * H + 9


def split(params):
    p = np.asarray(params, dtype=np.float64)
    assert p.size == NPARAMS
    o = 0
    out = []
    for shape in ((H, F), (H,), (H, 3 * H), (H,), (9, H), (9,)):
        k = int(np.prod(shape))
        out.append(p[o:o + k].reshape(shape))
        o += k
    return out


def init_params(seed=0, scale=0.1, out_scale=1e-3):
    """Random hidden layers, a small output layer (M starts near scaled damped Jacobi)."""
    rng = np.random.default_rng(seed)
    w1 = rng.standard_normal((H, F)) * scale
    b1 = rng.standard_normal(H) * scale
    w2 = rng.standard_normal((H, 3 * H)) * scale
    b2 = rng.standard_normal(H) * scale
    w3 = rng.standard_normal((9, H)) * out_scale
    b3 = np.zeros(9)
    return np.concatenate([a.ravel() for a in (w1, b1, w2, b2, w3, b3)])


def bsr_from_lower(n, colptr, rowidx, val):
    """3x3 block pattern (row pointers, sorted block columns, blocks) of the
    full symmetric matrix whose lower triangle is given in CSC."""
    low = sp.csc_matrix((val, rowidx, colptr), shape=(n, n))
    a = (low + low.T - sp.diags(low.diagonal())).tocsr()
    b = a.tobsr(blocksize=(3, 3))
    b.sort_indices()
    return b.indptr.astype(np.int64), b.indices.astype(np.int64), b.data.copy()


def smoother(params, ptr, idx, blocks):
    """The blocks of M (nnzb, 3, 3) on A's block pattern."""
    w1, b1, w2, b2, w3, b3 = split(params)
    nb = len(ptr) - 1
    deg = np.diff(ptr)
    row = np.repeat(np.arange(nb), deg)
    diag = row == idx
    tr = np.zeros(nb)
    tr[row[diag]] = np.trace(blocks[diag], axis1=1, axis2=2)
    s = np.sqrt(3.0 / tr)
    ss = s[row] * s[idx]
    x = np.concatenate([ss[:, None] * blocks.reshape(-1, 9), diag[:, None].astype(np.float64)], axis=1)
    m = np.maximum(x @ w1.T + b1, 0.0)
    node = np.zeros((nb, H))
    np.add.at(node, row, m)
    node /= np.maximum(deg, 1)[:, None]
    z = np.concatenate([m, node[row], node[idx]], axis=1)
    hid = np.maximum(z @ w2.T + b2, 0.0)
    o = (hid @ w3.T + b3).reshape(-1, 3, 3)
    key = row * nb + idx
    rev = np.searchsorted(key, idx * nb + row)
    assert np.array_equal(key[rev], idx * nb + row), "pattern not structurally symmetric"
    g = 0.5 * (o + np.swapaxes(o[rev], 1, 2))
    g[diag] += OMEGA * np.eye(3)
    return ss[:, None, None] * g


def bsr_matrix(ptr, idx, blocks):
    nb = len(ptr) - 1
    return sp.bsr_matrix((blocks, idx, ptr), shape=(3 * nb, 3 * nb))


def loss(params, ptr, idx, blocks, e0, nu=2):
    """Smoothing loss for the columns of e0 (n x nvec)."""
    a = bsr_matrix(ptr, idx, blocks).tocsr()
    m = bsr_matrix(ptr, idx, smoother(params, ptr, idx, blocks)).tocsr()
    e = np.array(e0, dtype=np.float64, copy=True).reshape(a.shape[0], -1)
    den = np.einsum("iv,iv->v", e, a @ e)
    for _ in range(nu):
        e = e - m @ (a @ e)
    num = np.einsum("iv,iv->v", e, a @ e)
    return float(np.mean(num / den))
