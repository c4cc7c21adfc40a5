"""The GPU handle's AMG hierarchy (rustyfem_pcg_setup with
RUSTYFEM_PCG_KEEP_HIERARCHY) against mathematics and against the oracle, and
the heterogeneous-stiffness block (the 128^3 workload's generator at 12^3)
solved by the FP32-V-cycle FCG and the FP64 PCG against the oracle's direct
solve (src/solver/iterative/amg.rs:108-341, src/solver/iterative/csr.rs:326-380).

The level checks (tests/_amg_checks.py): aggregates partition the blocks,
P~ couples a DOF only to its aggregate's columns, P~^T P~ = I, P~ B_c = B,
P = P~ - omega D^-1 A P~ with omega = 4 / (3 rho), A_c = P^T A P symmetric.
Against the oracle: the same aggregates on every level (integers), P~ to
1e-12 and A_c to 1e-9 (the spectral radius comes from the same power method
in another summation order on the GPU).
"""
import numpy as np
import pytest

from rustyfem_amd import meshes
from tests import _amg_checks as ac

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rf():
    from rustyfem_amd import ffi
    ffi.lib()
    return ffi


@pytest.mark.parametrize("case", ["hex10", "hetero10"])
def test_gpu_hierarchy_mathematics_and_oracle(rf, ora, case):
    p = meshes.hex_block(10, nrhs=1) if case == "hex10" else meshes.hetero_block(10)
    print(f"case amg hierarchy {case} n={p.n}", flush=True)
    nodes, coords, ndof, ap, ai = meshes.node_graph(p)
    cp, ri, v = ac.lower_of(p)
    s = rf.Pcg(p.n, cp, ri, v, coords=coords, ndof=ndof, flags=rf.PCG_KEEP_HIERARCHY)
    h = ora.AmgHierarchy(p.n, cp, ri, v, coords, ndof)
    nlev = s.stats()["levels"] - 1
    assert nlev == h.nlev >= 1
    for l in range(nlev):
        g, o = s.level(l), h.level(l)
        ac.check_level(g, s.level(l + 1))
        assert np.array_equal(g["agg"], o["agg"])
        assert np.array_equal(g["bptr"], o["bptr"])
        d = (g["Pt"] - o["Pt"]).toarray()
        assert np.max(np.abs(d)) <= 1e-12
        gc, oc = s.level(l + 1)["A"].toarray(), h.level(l + 1)["A"].toarray()
        assert np.max(np.abs(gc - oc)) <= 1e-9 * np.max(np.abs(oc))


@pytest.mark.parametrize("arm", ["fcg_fp32_vcycle", "pcg_fp64"])
def test_heterogeneous_block_against_oracle_direct(rf, ora, arm):
    p = meshes.hetero_block(12)
    print(f"case hetero12 {arm} n={p.n}", flush=True)
    nodes, coords, ndof, ap, ai = meshes.node_graph(p)
    cp, ri, v = ac.lower_of(p)
    kw = dict(krylov="fcg", vcycle="fp32") if arm == "fcg_fp32_vcycle" else dict(krylov="pcg", vcycle="fp64")
    s = rf.Pcg(p.n, cp, ri, v, coords=coords, ndof=ndof, **kw)
    b = p.rhs[:, 0].copy()
    x, it, rel, conv = s.solve(b, tol=1e-8)
    assert conv and rel <= 1e-8
    a = p.to_scipy()
    assert np.linalg.norm(b - a @ x) / np.linalg.norm(b) <= 1e-8
    ho, _, _ = ora.Handle.factor(p)
    xo = ho.solve(b)
    err = np.max(np.abs(x - xo)) / np.max(np.abs(xo))
    print(f"iterations={it} rel={rel:.2e} err_vs_oracle_direct={err:.2e}", flush=True)
    # a 1e-8 residual on this contrast-1e3 block leaves errors of a few 1e-8
    assert err <= 1e-6
    xo2, rep = ora.solve_lower(p.n, cp, ri, v, b, "amg", coords, ndof, tol=1e-8)
    assert rep["converged"]
    assert s.stats()["sizes"] == rep["amg_sizes"][:len(s.stats()["sizes"])]


def _problem(case):
    if case == "hex10":
        return meshes.hex_block(10, nrhs=1)
    if case == "hetero12":
        return meshes.hetero_block(12)
    if case == "tet10":
        return meshes.tet_block(10, seed=5)
    return meshes.plate_cpe4(60, 12, 7e10, 0.3, nrhs=1)


@pytest.mark.parametrize("case", ["hex10", "hetero12", "tet10", "plate60x12"])
def test_device_setup_equals_host_setup(rf, case, monkeypatch):
    """The hierarchy built on the GPU (amg_device.hpp) is the host setup's bit
    for bit: aggregates, block boundaries, near-null spaces, P~, P and every
    level matrix, and the solves on the two handles return the same bits."""
    p = _problem(case)
    print(f"case amg device vs host setup {case} n={p.n}", flush=True)
    nodes, coords, ndof, ap, ai = meshes.node_graph(p)
    cp, ri, v = ac.lower_of(p)
    monkeypatch.delenv("RUSTYFEM_AMG_HOST_SETUP", raising=False)
    d = rf.Pcg(p.n, cp, ri, v, coords=coords, ndof=ndof, krylov="fcg", vcycle="fp32", flags=rf.PCG_KEEP_HIERARCHY)
    monkeypatch.setenv("RUSTYFEM_AMG_HOST_SETUP", "1")
    h = rf.Pcg(p.n, cp, ri, v, coords=coords, ndof=ndof, krylov="fcg", vcycle="fp32", flags=rf.PCG_KEEP_HIERARCHY)
    sd, sh = d.stats(), h.stats()
    assert sd["setup_on_device"] and not sh["setup_on_device"]
    assert sd["sizes"] == sh["sizes"] and sd["blocks"] == sh["blocks"]
    for l in range(sd["levels"]):
        gd, gh = d.level(l), h.level(l)
        for key in ("nblk", "nagg", "nbm"):
            assert gd[key] == gh[key]
        assert np.array_equal(gd["bptr"], gh["bptr"])
        assert np.array_equal(gd["B"], gh["B"])
        for key in ("A", "Pt", "P"):
            if key not in gh:
                continue
            md, mh = gd[key], gh[key]
            assert np.array_equal(md.indptr, mh.indptr) and np.array_equal(md.indices, mh.indices), (l, key)
            assert np.array_equal(md.data, mh.data), (l, key, np.max(np.abs(md.data - mh.data)))
        if "agg" in gh:
            assert np.array_equal(gd["agg"], gh["agg"])
    b = p.rhs[:, 0].copy()
    xd, itd, reld, convd = d.solve(b, tol=1e-8)
    xh, ith, relh, convh = h.solve(b, tol=1e-8)
    assert convd and convh and itd == ith
    assert np.array_equal(xd, xh)
