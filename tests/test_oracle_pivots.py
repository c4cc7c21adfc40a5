"""The oracle's pivot-sequence export, pinned by determinants and inertia:
the D blocks of an LDL^T of P A P^T multiply to det A and their signs give the
inertia (Sylvester); the U diagonal of an LU with symmetric permutations (and
a row permutation at the roots) multiplies to +-det A."""
import numpy as np
import pytest

from tests._problems import bipartite_grid, csc, laplacian3d, laplacian_sigma, unsym_grid


def _ldl_logdet_inertia(kind, d, e):
    logdet, neg, pos = 0.0, 0, 0
    i = 0
    while i < len(d):
        if kind[i] == 2:
            det = d[i] * d[i + 1] - e[i] * e[i]
            logdet += np.log(abs(det))
            ev = np.linalg.eigvalsh(np.array([[d[i], e[i]], [e[i], d[i + 1]]]))
            neg += int((ev < 0).sum())
            pos += int((ev > 0).sum())
            i += 2
        else:
            assert kind[i] == 1
            logdet += np.log(abs(d[i]))
            neg += int(d[i] < 0)
            pos += int(d[i] > 0)
            i += 1
    return logdet, neg, pos


@pytest.mark.parametrize("case", ["laplace6", "bip5", "bip6"])
def test_indef_pivots_determinant_and_inertia(ora, case):
    if case == "laplace6":
        sigma, _ = laplacian_sigma(6)
        p = laplacian3d(6, sigma)
    else:
        p = bipartite_grid(int(case[3:]), 1, 1e-2)
    h, err, ine = ora.Handle.factor_indef(p)
    assert h is not None
    eq, d, e, kind = h.pivots()
    assert sorted(eq.tolist()) == list(range(p.n))
    logdet, neg, pos = _ldl_logdet_inertia(kind, d, e)
    a = csc(p).toarray()
    sign, ref = np.linalg.slogdet(a)
    assert abs(logdet - ref) < 1e-8 * max(1.0, abs(ref))
    ev = np.linalg.eigvalsh(a)
    assert [neg, pos] == [int((ev < 0).sum()), int((ev > 0).sum())] == [ine[0], ine[2]]
    st = h.stats()
    assert st[4] == int((kind == 2).sum())
    if case != "laplace6":
        assert st[4] > 0


@pytest.mark.parametrize("zf", [0.0, 0.6, 1.0])
def test_lu_pivots_determinant(ora, zf):
    p = unsym_grid(4 if zf == 1.0 else 5, 0, zf)
    h, err = ora.Handle.factor_unsym(p)
    assert h is not None
    eq, d, e, kind = h.pivots()
    assert sorted(eq.tolist()) == list(range(p.n))
    _, ref = np.linalg.slogdet(csc(p).toarray())
    assert abs(np.sum(np.log(np.abs(d))) - ref) < 1e-8 * max(1.0, abs(ref))
    if zf == 1.0:
        assert (kind == 3).any()
        # the root's row permutation is a permutation of its partial-pivoting block
        r = e[kind == 3].astype(int)
        assert sorted(r.tolist()) == list(range(len(r)))
