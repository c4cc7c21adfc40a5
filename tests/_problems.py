"""Small problem builders shared by the tests (CSC both triangles, caller numbering)."""
import math
from types import SimpleNamespace

import numpy as np


def chain(n, diag, off):
    """1D chain with one DOF per node, CSC with both triangles (ffi/src/lib.rs:415-427)."""
    indptr = [0]
    indices = []
    data = []
    for c in range(n):
        for r in range(max(c - 1, 0), min(c + 2, n)):
            indices.append(r)
            data.append(diag(c) if r == c else off(r, c))
        indptr.append(len(indices))
    return (np.array(indptr, dtype=np.int64), np.array(indices, dtype=np.int64), np.array(data, dtype=np.float64))


def chain_problem(n, diag, off):
    ip, ix, d = chain(n, diag, off)
    coords = np.array([[i, 0.0, 0.0] for i in range(n)], dtype=np.float64)
    return SimpleNamespace(n=n, nn=n, loc=np.arange(n, dtype=np.int64), coords=coords, indptr=ip, indices=ix, data=d)


def dense_problem(a, coords=None, loc=None):
    """Dense matrix a (both triangles) as CSC, one node per DOF unless loc is given."""
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    ip, ix, dd = [0], [], []
    for c in range(n):
        for r in range(n):
            if a[r, c] != 0.0 or r == c:
                ix.append(r)
                dd.append(a[r, c])
        ip.append(len(ix))
    if loc is None:
        loc = np.arange(n, dtype=np.int64)
    nn = int(np.max(loc)) + 1 if n else 0
    if coords is None:
        coords = np.array([[float(i), 0.0, 0.0] for i in range(nn)])
    return SimpleNamespace(n=n, nn=nn, loc=np.asarray(loc, dtype=np.int64), coords=np.asarray(coords, dtype=np.float64),
                           indptr=np.array(ip, dtype=np.int64), indices=np.array(ix, dtype=np.int64),
                           data=np.array(dd, dtype=np.float64))


def laplacian3d(g, sigma=0.0):
    """7-point Laplacian on a g^3 grid shifted by sigma (examples/solver_bench.rs:87-157):
    returns the node graph and a CSC (both triangles) problem with one DOF per node."""
    nn = g ** 3
    idx = lambda i, j, k: i + j * g + k * g * g
    coords = np.zeros((nn, 3))
    rows, cols, vals = [], [], []
    for k in range(g):
        for j in range(g):
            for i in range(g):
                v = idx(i, j, k)
                coords[v] = (i, j, k)
                rows.append(v); cols.append(v); vals.append(6.0 - sigma)
                for (di, dj, dk) in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < g and 0 <= jj < g and 0 <= kk < g:
                        rows.append(idx(ii, jj, kk)); cols.append(v); vals.append(-1.0)
    import scipy.sparse as sp
    a = sp.csc_matrix((vals, (rows, cols)), shape=(nn, nn))
    a.sort_indices()
    p = SimpleNamespace(n=nn, nn=nn, loc=np.arange(nn, dtype=np.int64), coords=coords,
                        indptr=a.indptr.astype(np.int64), indices=a.indices.astype(np.int64), data=a.data.copy())
    return p


def laplacian_sigma(g):
    """sigma midway between two distinct eigenvalues near the 10 % quantile and the
    analytic number of eigenvalues below it (examples/solver_bench.rs:122-141)."""
    c = [2.0 * math.cos(k * math.pi / (g + 1)) for k in range(1, g + 1)]
    lam = sorted(6.0 - a - b - d for a in c for b in c for d in c)
    nn = g ** 3
    q = nn // 10
    while lam[q + 1] - lam[q] < 1e-9:
        q += 1
    sigma = 0.5 * (lam[q] + lam[q + 1])
    below = sum(1 for l in lam if l < sigma)
    return sigma, below


def _with_values(p, a):
    a = a.tocsc()
    a.sort_indices()
    p.data = a.data.copy()
    p.indptr = a.indptr.astype(np.int64)
    p.indices = a.indices.astype(np.int64)
    return p


def bipartite_grid(g, seed, dscale):
    """7-point grid graph with random +-[0.5, 1.5] couplings and a diagonal of
    dscale * N(0, 1): strongly indefinite, exercises 2x2 and delayed pivots
    (zero diagonal on an odd grid is singular)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    p = laplacian3d(g, 0.0)
    a = sp.csc_matrix((p.data, p.indices, p.indptr), shape=(p.n, p.n)).tocoo()
    w = {}
    for r, c in zip(a.row, a.col):
        if r < c:
            w[(r, c)] = rng.uniform(0.5, 1.5) * (1.0 if rng.random() < 0.5 else -1.0)
    vals = [dscale * rng.standard_normal() if r == c else w[(min(r, c), max(r, c))] for r, c in zip(a.row, a.col)]
    return _with_values(p, sp.coo_matrix((vals, (a.row, a.col)), shape=(p.n, p.n)))


def unsym_grid(g, seed, zero_frac=0.0, conv=0.6):
    """Convection-diffusion-like unsymmetric values on the 7-point grid pattern;
    a fraction of zero diagonals forces delayed pivots (all zero: the roots
    take partial-pivoting LU)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    p = laplacian3d(g, 0.0)
    a = sp.csc_matrix((p.data, p.indices, p.indptr), shape=(p.n, p.n)).tocoo()
    vals = []
    for r, c in zip(a.row, a.col):
        if r == c:
            vals.append(0.0 if rng.random() < zero_frac else 6.0 + rng.uniform(-1, 1))
        else:
            vals.append(-1.0 + conv * (1.0 if r > c else -1.0) + 0.2 * rng.standard_normal())
    return _with_values(p, sp.coo_matrix((vals, (a.row, a.col)), shape=(p.n, p.n)))


def csc(p):
    import scipy.sparse as sp
    return sp.csc_matrix((p.data, p.indices, p.indptr), shape=(p.n, p.n))
