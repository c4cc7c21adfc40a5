"""Structure of the oracle's nested dissection (src/sparse.rs:78-253), pinned by
invariants that hold for any correct restatement and by closed forms on
structured grids."""
import numpy as np
import pytest

from rustyfem_amd import meshes


def grid_graph(nx, ny, nz):
    """Node graph of a structured hex grid (27-point node adjacency), box-mesher numbering."""
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coords = np.stack([ii.ravel() / nx, jj.ravel() / ny, kk.ravel() / nz], axis=1)
    i, j, k = ii.ravel(), jj.ravel(), kk.ravel()
    nbrs = []
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                if di == dj == dk == 0:
                    continue
                a, b, c = i + di, j + dj, k + dk
                ok = (a >= 0) & (a <= nx) & (b >= 0) & (b <= ny) & (c >= 0) & (c <= nz)
                q = np.where(ok, (c * (ny + 1) + b) * (nx + 1) + a, -1)
                nbrs.append(q)
    nb = np.stack(nbrs, axis=1)
    nb.sort(axis=1)
    cnt = (nb >= 0).sum(axis=1)
    adj_ptr = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    adj_idx = nb[nb >= 0].astype(np.int64)
    return coords, adj_ptr, adj_idx, nn


def check_invariants(sym, coords, adj_ptr, adj_idx, ndof):
    nn = len(ndof)
    fr = sym.fronts()
    nf = sym.nfronts
    fd = sym.first_dof()
    seen = np.zeros(nn, dtype=int)
    owner = np.zeros(nn, dtype=int)
    for f in range(nf):
        nodes = sym.nodes(f)
        seen[nodes] += 1
        owner[nodes] = f
    assert np.all(seen == 1), "every node sits in exactly one front"
    # postorder and contiguous numbering
    start = 0
    for f in range(nf):
        assert fr["start"][f] == start
        start += fr["k"][f]
        if fr["parent"][f] >= 0:
            assert fr["parent"][f] > f
    assert start == sym.n == int(np.sum(ndof))
    children = [[] for _ in range(nf)]
    for f in range(nf):
        if fr["parent"][f] >= 0:
            children[fr["parent"][f]].append(f)
    assert all(len(c) == fr["nchild"][f] for f, c in enumerate(children))
    # subtree membership
    sub = [None] * nf
    for f in range(nf):
        s = set(sym.nodes(f).tolist())
        for c in children[f]:
            s |= sub[c]
        sub[f] = s
    work = np.zeros(nf)
    for f in range(nf):
        rows = sym.rows(f)
        end = fr["start"][f] + fr["k"][f]
        assert np.all(np.diff(rows) > 0), "rows sorted"
        assert rows.size == 0 or rows[0] >= end, "rows after own pivots"
        # rows = later neighbours of the whole subtree (node-level elimination structure)
        later = set()
        for v in sub[f]:
            for w in adj_idx[adj_ptr[v]:adj_ptr[v + 1]]:
                if w not in sub[f]:
                    later.add(int(w))
        expect = sorted(d for w in later for d in range(fd[w], fd[w] + ndof[w]))
        assert rows.tolist() == expect
        # the separator disconnects the children's subtrees
        if len(children[f]) == 2:
            a, b = sub[children[f][0]], sub[children[f][1]]
            for v in a:
                for w in adj_idx[adj_ptr[v]:adj_ptr[v + 1]]:
                    assert int(w) not in b
        k, u = float(fr["k"][f]), float(len(rows))
        work[f] = k * k * k / 3.0 + u * k * k + u * u * k + sum(work[c] for c in children[f])
    assert sym.work() == work[nf - 1]
    fs = sum(int(fr["k"][f]) * (int(fr["k"][f]) - 1) // 2 + int(fr["k"][f]) * int(fr["u"][f]) for f in range(nf))
    assert sym.factor_size() == fs
    # the pattern holds every lower entry of every node block and is sorted per column
    colptr, rowidx = sym.pattern()
    for c in range(sym.n):
        col = rowidx[colptr[c]:colptr[c + 1]]
        assert col[0] == c and np.all(np.diff(col) > 0)


@pytest.mark.parametrize("leaf", [16, 48])
def test_invariants_cantilever(ora, leaf):
    p = meshes.cantilever_cps4()
    nodes, coords, ndof, ap, ai = meshes.node_graph(p)
    sym = ora.Symbolic(coords, ap, ai, ndof, leaf)
    check_invariants(sym, coords, ap, ai, ndof)
    assert max(len(sym.nodes(f)) for f in range(sym.nfronts) if sym.fronts()["nchild"][f] == 0) <= leaf


@pytest.mark.parametrize("n,leaf", [(6, 16), (8, 16), (7, 48), (5, 1)])
def test_invariants_hex(ora, n, leaf):
    coords, ap, ai, nn = grid_graph(n, n, n)
    ndof = np.full(nn, 3, dtype=np.int64)
    sym = ora.Symbolic(coords, ap, ai, ndof, leaf)
    check_invariants(sym, coords, ap, ai, ndof)


def test_mixed_ndof_and_irregular(ora):
    rng = np.random.default_rng(3)
    nn = 120
    coords = rng.uniform(0, 1, (nn, 3))
    ap, ai = [0], []
    for v in range(nn):
        d = np.linalg.norm(coords - coords[v], axis=1)
        nb = np.nonzero((d < 0.3) & (d > 0))[0]
        ai.extend(nb.tolist())
        ap.append(len(ai))
    ndof = rng.integers(1, 4, nn)
    sym = ora.Symbolic(coords, np.array(ap), np.array(ai), ndof, 8)
    check_invariants(sym, coords, np.array(ap), np.array(ai), ndof)


def test_root_separator_closed_form_64(ora):
    """64^3 C3D8 block clamped at x=0 (the 811,200-DOF configuration): the root
    separator is a full y-plane of 64 x 65 free nodes, k = 12,480, u = 0; the
    y, z and x planes tie at score 8450 and y wins because it is tried first."""
    coords, ap, ai, nn = grid_graph(64, 64, 64)
    free = coords[:, 0] > 0
    fid = np.cumsum(free) - 1
    keep = np.nonzero(free)[0]
    ap2, ai2 = [0], []
    for v in keep:
        nb = ai[ap[v]:ap[v + 1]]
        nb = nb[free[nb]]
        ai2.extend(fid[nb].tolist())
        ap2.append(len(ai2))
    ndof = np.full(len(keep), 3, dtype=np.int64)
    sym = ora.Symbolic(coords[keep], np.array(ap2), np.array(ai2), ndof, 16)
    fr = sym.fronts()
    root = sym.nfronts - 1
    assert sym.n == 811200
    assert fr["k"][root] == 12480 and fr["u"][root] == 0
    ys = coords[keep][sym.nodes(root)][:, 1]
    assert np.all(ys == ys[0])
