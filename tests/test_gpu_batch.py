"""Batches of independent plates factored and solved together through
rustyfem_factor_batch, against the oracle problem by problem: each problem's
ordering equals the oracle's for that problem alone, and each solution matches
the oracle's.  Refinement tests every problem's relative residual (not the
concatenated one) against 1e-10; slender plates cannot reach 1e-10 in FP64
(their rounding floor eps ||A| |x|| / ||b|| is about 1e-9, and the FP64 CPU
oracle itself ends at 1e-10 to 4e-10), so each plate must reach
max(1e-10, 2 x the oracle's own residual)."""
import numpy as np
import pytest

from rustyfem_amd import meshes
from tests._problems import chain_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rf():
    from rustyfem_amd import ffi
    ffi.lib()
    return ffi


@pytest.fixture(scope="module")
def plates():
    return meshes.plate_batch(count=24, seed=5, nrhs=3)


@pytest.mark.parametrize("precision", [None, "mixed"])
def test_batch_matches_oracle(rf, ora, plates, precision):
    print(f"case batch precision={precision}", flush=True)
    b = rf.Batch(plates, precision=precision, max_refine=30, refine_tol=1e-10)
    assert b.ok, rf.last_error()
    assert np.all(b.err == -1)
    info = b.info()
    assert info["problems"] == len(plates) and info["n"] == sum(p.n for p in plates)
    xs = b.solve([p.rhs for p in plates])
    work = 0.0
    for i, p in enumerate(plates):
        ho, _, _ = ora.Handle.factor(p)
        assert np.array_equal(b.order(i), ho.order())
        work += ho.info()[0]
        xo = ho.solve(p.rhs)
        a = p.to_scipy()
        rel = np.linalg.norm(p.rhs - a @ xs[i], axis=0) / np.linalg.norm(p.rhs, axis=0)
        rel_o = np.linalg.norm(p.rhs - a @ xo, axis=0) / np.linalg.norm(p.rhs, axis=0)
        assert np.all(rel <= np.maximum(1e-10, 2.0 * rel_o)), (i, rel, rel_o)
        assert np.max(np.abs(xs[i] - xo)) / np.max(np.abs(xo)) <= 1e-9
    assert abs(info["work"] - work) <= 1e-12 * work


def test_batch_device_solve_and_refactor(rf, plates):
    from rustyfem_amd import hip
    b = rf.Batch(plates)
    xs = b.solve([p.rhs for p in plates])
    nrhs = plates[0].rhs.shape[1]
    allb = np.concatenate([p.rhs for p in plates], axis=0)
    d = hip.DeviceArray.from_host(np.asfortranarray(allb).reshape(-1, order="F"))
    b.solve_device(d.ptr, nrhs)
    got = d.download().reshape(allb.shape, order="F")
    d.free()
    assert np.array_equal(got, np.concatenate(xs, axis=0))
    rc, err = b.refactor([2.0 * p.data for p in plates])
    assert rc == 0 and np.all(err == -1)
    x2 = b.solve([p.rhs for p in plates])
    for u, v in zip(x2, xs):
        assert np.allclose(u, 0.5 * v, rtol=1e-12, atol=0)


def test_batch_fused_solves_match_superblock_kernels(rf, monkeypatch, plates):
    """Every level of a batch of small plates is solved by the fused per-front
    kernels; their solutions equal the super-block kernels' bit for bit."""
    monkeypatch.setenv("RUSTYFEM_NO_SMALL_SOLVE", "1")
    b_sb = rf.Batch(plates, max_refine=0)
    monkeypatch.delenv("RUSTYFEM_NO_SMALL_SOLVE")
    b_f = rf.Batch(plates, max_refine=0)
    assert b_sb.stats()["small_levels"] == 0 and b_f.stats()["small_levels"] == b_f.info()["levels"]
    for u, v in zip(b_f.solve([p.rhs for p in plates]), b_sb.solve([p.rhs for p in plates])):
        assert np.array_equal(u, v)


def test_batch_reports_failing_problem(rf, ora, plates):
    bad = chain_problem(30, lambda c: 2.0 - 1.3, lambda r, c: -1.0)
    probs = [plates[0], bad, plates[1]]
    b = rf.Batch(probs)
    assert not b.ok
    _, erro, _ = ora.Handle.factor(bad)
    assert b.err[0] == -1 and b.err[2] == -1 and b.err[1] == erro >= 0
