"""Smoothed-aggregation AMG setup (src/solver/iterative/amg.rs:108-341) pinned by
hand-derived aggregations and by mathematics, for the oracle (oracle/ora_iter.c)
and for the product's host aggregation (rustyfem_amg_aggregate, the code the
GPU handle's setup runs).  The GPU handle's whole hierarchy is checked the same
way in tests/test_gpu_amg.py.

Pass 3 of the aggregation is not exercised: a block that pass 1 leaves free
had an assigned strong neighbour when it was visited, so pass 2 always finds
one in its snapshot; pass 3 cannot fire (in the reference either).
"""
import numpy as np
import pytest

from rustyfem_amd import meshes
from tests import _amg_checks as ac


@pytest.fixture(scope="module")
def rf():
    from rustyfem_amd import ffi
    ffi.lib()
    return ffi


def _cases():
    a, bp, want = ac.chain7()
    yield "chain7", a, bp, want
    a, bp, want = ac.chain_permuted()
    yield "chain_permuted", a, bp, want
    aw, bp, ww, as_, ws = ac.grid3_weak()
    yield "grid_weak", aw, bp, ww
    yield "grid_strong", as_, bp, ws
    for i, (a, bp, want) in enumerate(ac.blocks3_pair()):
        yield f"blocks3_{i}", a, bp, want


@pytest.mark.parametrize("case", [c[0] for c in _cases()])
def test_hand_derived_aggregation(rf, ora, case):
    _, a, bptr, want = next(c for c in _cases() if c[0] == case)
    got_o, n_o = ora.amg_aggregate(a, bptr, 0.08)
    got_p, n_p = rf.amg_aggregate(a, bptr, 0.08)
    assert got_o.tolist() == want.tolist()
    assert got_p.tolist() == want.tolist()
    assert n_o == n_p == want.max() + 1


def test_snapshot_rule_changes_the_answer():
    # the live-map variant of pass 2 (the plausible slip) gives another answer
    # on the permuted chain: this is what the case above pins
    a, bptr, want = ac.chain_permuted()
    strong = [sorted(j for j in a[i].indices if j != i) for i in range(6)]
    agg = [-1] * 6
    n = 0
    for b in range(6):
        if agg[b] == -1 and all(agg[c] == -1 for c in strong[b]):
            agg[b] = n
            for c in strong[b]:
                agg[c] = n
            n += 1
    for b in range(6):
        if agg[b] == -1:
            c = next(c for c in strong[b] if agg[c] != -1)
            agg[b] = agg[c]
    assert agg != want.tolist()


@pytest.mark.parametrize("case", ["hex10", "hetero10"])
def test_oracle_hierarchy_mathematics(ora, case):
    p = meshes.hex_block(10, nrhs=1) if case == "hex10" else meshes.hetero_block(10)
    nodes, coords, ndof, ap, ai = meshes.node_graph(p)
    cp, ri, v = ac.lower_of(p)
    h = ora.AmgHierarchy(p.n, cp, ri, v, coords, ndof)
    assert h.nlev >= 1
    for l in range(h.nlev):
        ac.check_level(h.level(l), h.level(l + 1))


def test_product_aggregation_equals_oracle_on_fine_levels(rf, ora):
    # the first level's aggregation of two real matrices, 3 x 3 node blocks
    for p in (meshes.hex_block(10, nrhs=1), meshes.hetero_block(10)):
        nodes, coords, ndof, ap, ai = meshes.node_graph(p)
        a = p.to_scipy().tocsr()
        bptr = np.concatenate([[0], np.cumsum(ndof)])
        for theta in (0.08, 0.04):
            go, no = ora.amg_aggregate(a, bptr, theta)
            gp, npr = rf.amg_aggregate(a, bptr, theta)
            assert no == npr and np.array_equal(go, gp)
