"""Mathematical checks of one smoothed-aggregation AMG level, shared by the
oracle's tests and the GPU handle's tests (src/solver/iterative/amg.rs:108-341).

For level l with matrix A, blocks bptr, aggregates agg, near-null space B,
tentative prolongator P~, smoothed P and the next level's A_c and B_c:
  * the aggregates partition the blocks and every aggregate is non-empty;
  * P~ only couples a DOF to the coarse columns of its own aggregate, and the
    coarse columns of one aggregate are contiguous, in aggregate order;
  * P~^T P~ = I (orthonormal columns per aggregate, disjoint aggregates);
  * P~ B_c = B (the QR of the near-null space on every aggregate);
  * P = P~ - omega D^-1 A P~ for one omega with omega * lambda_max(D^-1 A) in
    [4 / (3 * 1.05) * (1 - 1e-3), 4 / 3]: 4 / (3 rho) with rho 1.05 x the
    power-method estimate, which never exceeds lambda_max;
  * A_c = P^T A P, symmetric.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def block_of(bptr):
    return np.repeat(np.arange(bptr.size - 1), np.diff(bptr))


def check_level(lv, nxt, tol=1e-11):
    A, Pt, P, agg, bptr, B = lv["A"], lv["Pt"], lv["P"], lv["agg"], lv["bptr"], lv["B"]
    n = A.shape[0]
    nagg = lv["nagg"]
    # partition
    assert agg.shape == (bptr.size - 1,)
    assert agg.min() == 0 and agg.max() == nagg - 1
    assert np.all(np.bincount(agg, minlength=nagg) > 0)
    # column ownership: coarse column -> aggregate, contiguous in aggregate order
    dof_agg = agg[block_of(bptr)]
    coo = Pt.tocoo()
    col_agg = np.full(Pt.shape[1], -1)
    col_agg[coo.col] = dof_agg[coo.row]
    assert np.all(col_agg >= 0)
    assert np.all(np.diff(col_agg) >= 0)
    assert np.array_equal(dof_agg[coo.row], col_agg[coo.col])
    # orthonormal columns
    g = (Pt.T @ Pt).toarray()
    assert np.max(np.abs(g - np.eye(g.shape[0]))) < tol * 10
    # the near-null space is reproduced: P~ B_c = B
    Bc = nxt["B"]
    assert Bc.shape == (Pt.shape[1], B.shape[1])
    scale = np.max(np.abs(B))
    assert np.max(np.abs(Pt @ Bc - B)) <= 1e-12 * scale * 10
    # the smoothing step
    d = A.diagonal()
    w = sp.diags(1.0 / d) @ (A @ Pt)
    diff = (Pt - P).toarray().ravel()
    wv = w.toarray().ravel()
    omega = float(diff @ wv / (wv @ wv))
    assert np.max(np.abs(diff - omega * wv)) <= 1e-12 * max(np.max(np.abs(diff)), 1e-300) * 10
    lam = float(sla.eigsh(sp.diags(1.0 / np.sqrt(d)) @ A @ sp.diags(1.0 / np.sqrt(d)), k=1, which="LA",
                          return_eigenvectors=False, tol=1e-10)[0])
    assert 4.0 / (3.0 * 1.05) * (1.0 - 1e-3) <= omega * lam <= 4.0 / 3.0 + 1e-9, omega * lam
    # Galerkin product
    Ac = nxt["A"]
    ref = (P.T @ A @ P).toarray()
    got = Ac.toarray()
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)) * 100
    assert np.max(np.abs(got - got.T)) <= 1e-13 * np.max(np.abs(got)) * 100
    return omega


def lower_of(p):
    a = p.to_scipy().tocoo()
    m = a.row >= a.col
    low = sp.csc_matrix((a.data[m], (a.row[m], a.col[m])), shape=a.shape)
    low.sort_indices()
    return low.indptr.astype(np.int64), low.indices.astype(np.int64), low.data.copy()


def chain_permuted():
    """A 6-node chain whose labels run 0-1-2-5-4-3 (edges 0-1, 1-2, 2-5, 5-4,
    4-3), scalar Laplacian.  Pass 1: 0 starts {0, 1}, 2 and 5 see assigned
    neighbours, 3 starts {3, 4}.  Pass 2 with the pass-1 snapshot: 2 joins
    aggregate 0 through 1; 5's strong list is [2, 4] and 2 had no aggregate
    in the snapshot, so 5 joins aggregate 1 through 4 -> [0, 0, 0, 1, 1, 1].
    Reading the live map instead would put 5 into aggregate 0 through 2."""
    edges = [(0, 1), (1, 2), (2, 5), (5, 4), (4, 3)]
    a = np.diag(np.full(6, 2.0))
    for i, j in edges:
        a[i, j] = a[j, i] = -1.0
    return sp.csr_matrix(a), np.arange(7), np.array([0, 0, 0, 1, 1, 1])


def grid3_weak():
    """3 x 3 grid (row-major), 2D Laplacian with diagonal 4; the 0-1 coupling is
    weak (0.1^2 <= 0.08^2 * 4, so not strong).  Pass 1: 0 starts {0, 3}, 1
    starts {1, 2, 4}, 8 starts {8, 5, 7}; pass 2: 6 joins 3's aggregate ->
    [0, 1, 1, 0, 1, 2, 0, 2, 2].  With the coupling strong (-1) the result is
    {0, 1, 3, 6}, {2, 4, 5, 7, 8} -> [0, 0, 1, 0, 1, 1, 0, 1, 1]."""
    def grid(c01):
        a = np.diag(np.full(9, 4.0))
        for r in range(3):
            for c in range(3):
                i = 3 * r + c
                if c < 2:
                    a[i, i + 1] = a[i + 1, i] = -1.0
                if r < 2:
                    a[i, i + 3] = a[i + 3, i] = -1.0
        a[0, 1] = a[1, 0] = c01
        return sp.csr_matrix(a)
    return (grid(-0.1), np.arange(10), np.array([0, 1, 1, 0, 1, 2, 0, 2, 2]),
            grid(-1.0), np.array([0, 0, 1, 0, 1, 1, 0, 1, 1]))


def chain7():
    """7-node chain in label order: {0, 1}, {2, 3, 4}, {5, 6} from pass 1 alone."""
    a = sp.diags([-np.ones(6), 2 * np.ones(7), -np.ones(6)], [-1, 0, 1]).tocsr()
    return a, np.arange(8), np.array([0, 0, 1, 1, 1, 2, 2])


def blocks3_pair():
    """Two 3-DOF nodes whose coupling block has Frobenius norm^2 just above and
    just below theta^2 sqrt(|D_0|^2 |D_1|^2): the block strength sums squares
    over the whole 3 x 3 coupling block."""
    d = np.eye(3) * 2.0  # |D|_F^2 = 12, theta^2 sqrt(12 * 12) = 0.0768 at theta = 0.08
    out = []
    for s, expect in ((0.17, [0, 0]), (0.15, [0, 1])):  # 3 s^2 = 0.0867 (strong), 0.0675 (weak)
        a = np.zeros((6, 6))
        a[:3, :3] = d
        a[3:, 3:] = d
        a[:3, 3:] = -s * np.eye(3)
        a[3:, :3] = -s * np.eye(3)
        out.append((sp.csr_matrix(a), np.array([0, 3, 6]), np.array(expect)))
    return out
