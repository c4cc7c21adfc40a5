"""Positive definite multifrontal LDL^T on the GPU through the C ABI, against the
oracle (same inputs) and against the true residual.

Integer results (ordering, info, error DOF) must be identical to the oracle's.
Floating point: the FP64 path matches the oracle's pivots to a relative 1e-9
and its solution to the problem's conditioning (relative residual <= 1e-12 on
these well-conditioned blocks, <= 1e-9 on the slender cantilever); the mixed
path (FP32 Schur updates + FP64 refinement) must reach a relative residual
<= 1e-10 and a displacement relative error <= 1e-9 against the oracle.
"""
import numpy as np
import pytest

from rustyfem_amd import meshes
from tests._problems import chain_problem, dense_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rf():
    from rustyfem_amd import ffi
    ffi.lib()
    return ffi


def relres(p, x, b):
    a = p.to_scipy() if hasattr(p, "to_scipy") else None
    if a is None:
        import scipy.sparse as sp
        a = sp.csc_matrix((p.data, p.indices, p.indptr), shape=(p.n, p.n))
    r = b - a @ x
    return np.linalg.norm(r, axis=0) / np.linalg.norm(b, axis=0)


@pytest.mark.parametrize("name", ["cantilever", "hex6", "hex10"])
def test_fp64_matches_oracle(rf, ora, name):
    p = {"cantilever": meshes.cantilever_cps4, "hex6": lambda: meshes.hex_block(6, nrhs=3),
         "hex10": lambda: meshes.hex_block(10, nrhs=2)}[name]()
    print(f"case {name} n={p.n}", flush=True)
    ho, eo, do = ora.Handle.factor(p, want_d=True)
    hg, eg, dg = rf.Handle.factor(p, want_d=True)
    assert hg is not None, rf.last_error()
    assert eg == eo == -1
    assert np.array_equal(hg.order(), ho.order())
    assert np.array_equal(hg.info(), ho.info())
    assert np.max(np.abs(dg - do) / np.abs(do)) < 1e-9
    xo = ho.solve(p.rhs)
    xg = hg.solve(p.rhs)
    tol = 1e-9 if name == "cantilever" else 1e-12
    assert np.all(relres(p, xg, p.rhs) < tol)
    assert np.max(np.abs(xg - xo)) / np.max(np.abs(xo)) < (1e-6 if name == "cantilever" else 1e-10)


@pytest.mark.parametrize("name", ["hex8", "hex12"])
def test_mixed_refined_meets_target(rf, ora, name):
    n = int(name[3:])
    p = meshes.hex_block(n, nrhs=4)
    print(f"case mixed {name} n={p.n}", flush=True)
    hg, eg, _ = rf.Handle.factor(p, precision="mixed")
    assert hg is not None, rf.last_error()
    xg = hg.solve(p.rhs)
    st = hg.stats()
    assert st["rel_residual"] <= 1e-10
    assert np.all(relres(p, xg, p.rhs) <= 1e-10)
    ho, _, _ = ora.Handle.factor(p)
    xo = ho.solve(p.rhs)
    assert np.max(np.abs(xg - xo)) / np.max(np.abs(xo)) <= 1e-9


def test_error_dof_matches_oracle(rf, ora):
    rng = np.random.default_rng(11)
    nn = 60
    coords = rng.uniform(0, 1, (nn, 3))
    a = np.zeros((nn, nn))
    for i in range(nn):
        for j in range(i + 1, nn):
            if np.linalg.norm(coords[i] - coords[j]) < 0.35:
                a[i, j] = a[j, i] = -rng.uniform(0.1, 1.0)
    a += np.diag(-a.sum(axis=1) + 0.05)
    a -= np.eye(nn) * np.linalg.eigvalsh(a)[nn // 4]
    p = dense_problem(a, coords=coords)
    ho, eo, _ = ora.Handle.factor(p)
    hg, eg, _ = rf.Handle.factor(p)
    assert ho is None and hg is None
    assert eg == eo >= 0


def test_chain_refused_and_uncoupled(rf, ora):
    # the tridiagonal K - 1.3 I of the reference ffi test is refused by the PD path
    p = chain_problem(40, lambda c: 2.0 - 1.3, lambda r, c: -1.0)
    h, err, _ = rf.Handle.factor(p)
    ho, erro, _ = ora.Handle.factor(p)
    assert h is None and err == erro >= 0
    a = np.array([[4.0, 1.0, 0.0, 0.0, 0.0], [1.0, 3.0, 0.0, 0.0, 0.0], [0.0, 0.0, 2.0, 0.0, 0.0],
                  [0.0, 0.0, 0.0, 5.0, -1.0], [0.0, 0.0, 0.0, -1.0, 6.0]])
    q = dense_problem(a, coords=np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]], float), loc=[0, 0, 1, 2, 2])
    h, err, d = rf.Handle.factor(q, want_d=True)
    assert h is not None
    ho, _, do = ora.Handle.factor(q, want_d=True)
    assert np.array_equal(h.order(), ho.order())
    assert np.allclose(d, do, rtol=1e-14)
    b = np.arange(1.0, 6.0)
    assert np.allclose(h.solve(b), np.linalg.solve(a, b), rtol=1e-13)


def test_refactor_same_pattern(rf):
    p = meshes.hex_block(6, nrhs=1)
    h, err, _ = rf.Handle.factor(p)
    x1 = h.solve(p.rhs)
    assert h.refactor(p.data * 2.0) == -1
    x2 = h.solve(p.rhs)
    assert np.allclose(x2, 0.5 * x1, rtol=1e-12, atol=0)
    assert h.refactor(None) == -1
    x3 = h.solve(p.rhs)
    assert np.allclose(x3, x2, rtol=1e-14, atol=0)


@pytest.mark.parametrize("precision", [None, "mixed"])
def test_hex24_sixteen_rhs_matches_oracle(rf, ora, precision):
    """The C3D8 block of the headline workload at 24^3 (45,000 DOF; fronts up
    to 1,875 pivots, so several super-blocks, the second stream and fused and
    super-block solve levels all take part) with 16 right-hand sides (the
    pressure load and 15 random nodal loads; the MFMA multi-RHS products):
    both arms against the oracle's FP64 solution.  The mixed arm refined to
    the default 1e-10 meets 1e-9 on the pressure load; the random loads' rough
    solutions amplify a 1e-10 residual to about 2e-9 in the solution, so the
    mixed arm is also refined to 1e-12, which must bring every column within
    5e-11 of the oracle."""
    p = meshes.hex_block(24, nrhs=16, seed=11)
    print(f"case hex24 16 rhs precision={precision} n={p.n}", flush=True)
    ho, _, _ = ora.Handle.factor(p)
    xo = ho.solve(p.rhs)
    scale = np.max(np.abs(xo), axis=0)
    hg, eg, _ = rf.Handle.factor(p, precision=precision)
    assert hg is not None, rf.last_error()
    xg = hg.solve(p.rhs)
    assert np.all(relres(p, xg, p.rhs) <= 1e-10)
    err = np.max(np.abs(xg - xo), axis=0) / scale
    if precision is None:
        assert np.all(err <= 1e-11), err
        return
    assert err[0] <= 1e-9, err
    ht, _, _ = rf.Handle.factor(p, precision="mixed", refine_tol=1e-12)
    xt = ht.solve(p.rhs)
    assert ht.stats()["rel_residual"] <= 1e-12
    err = np.max(np.abs(xt - xo), axis=0) / scale
    assert np.all(err <= 5e-11), err


@pytest.mark.parametrize("nrhs", [1, 5, 17])
def test_many_rhs_match_oracle(rf, ora, nrhs):
    """Multi-RHS solves (passes of 1, 4 and 16 columns) against the oracle."""
    p = meshes.hex_block(7, nrhs=nrhs, seed=5)
    print(f"case nrhs={nrhs} n={p.n}", flush=True)
    hg, _, _ = rf.Handle.factor(p)
    ho, _, _ = ora.Handle.factor(p)
    xg = hg.solve(p.rhs)
    xo = ho.solve(p.rhs)
    assert np.max(np.abs(xg - xo)) / np.max(np.abs(xo)) < 1e-10
    for j in range(nrhs):
        assert np.allclose(hg.solve(p.rhs[:, j]), xg[:, j], rtol=0, atol=1e-12 * np.max(np.abs(xg[:, j])))


def test_graph_replays_match_eager(rf, monkeypatch):
    """The captured solve graph, replayed several times with eager work and
    synchronizations in between, gives exactly the eager solve's result."""
    p = meshes.hex_block(6, nrhs=2)
    monkeypatch.setenv("RUSTYFEM_NO_GRAPH", "1")
    he, _, _ = rf.Handle.factor(p, precision="mixed")
    monkeypatch.delenv("RUSTYFEM_NO_GRAPH")
    hg, _, _ = rf.Handle.factor(p, precision="mixed")
    ref = he.solve(p.rhs)
    for rep in range(4):
        x = hg.solve(p.rhs)
        assert np.array_equal(x, ref), f"replay {rep}"
        assert he.refactor(None) == -1 and hg.refactor(None) == -1
        y = he.solve(p.rhs[:, :1])
        assert np.array_equal(hg.solve(p.rhs[:, :1]), y)


@pytest.mark.parametrize("precision", [None, "mixed"])
def test_superblock_solve_matches_blockwise_solve(rf, monkeypatch, precision):
    """The super-block triangular solves (512-column chains) against the
    block-by-block schedule on the same factor: same solution to rounding."""
    p = meshes.hex_block(9, nrhs=5, seed=3)
    monkeypatch.setenv("RUSTYFEM_LEGACY_SOLVE", "1")
    h_old, _, _ = rf.Handle.factor(p, precision=precision, max_refine=0)
    monkeypatch.delenv("RUSTYFEM_LEGACY_SOLVE")
    h_new, _, _ = rf.Handle.factor(p, precision=precision, max_refine=0)
    x_old = h_old.solve(p.rhs)
    x_new = h_new.solve(p.rhs)
    assert np.max(np.abs(x_new - x_old)) / np.max(np.abs(x_old)) < 1e-12


@pytest.mark.parametrize("precision", [None, "mixed"])
def test_split_panel_matches_fused_panel(rf, monkeypatch, precision):
    """The inner pivot step as two launches (pivot blocks, then row tiles) does
    the fused kernel's arithmetic: identical D and identical solutions."""
    p = meshes.hex_block(10, nrhs=2, seed=4)
    monkeypatch.setenv("RUSTYFEM_LEGACY_PANEL", "1")
    h_old, _, d_old = rf.Handle.factor(p, want_d=True, precision=precision, max_refine=0)
    monkeypatch.delenv("RUSTYFEM_LEGACY_PANEL")
    h_new, _, d_new = rf.Handle.factor(p, want_d=True, precision=precision, max_refine=0)
    assert np.array_equal(d_new, d_old)
    assert np.array_equal(h_new.solve(p.rhs), h_old.solve(p.rhs))


@pytest.mark.parametrize("precision", [None, "mixed"])
@pytest.mark.parametrize("nrhs,block", [(1, 4), (5, 4), (17, 4), (17, 16)])
def test_fused_small_front_solves_match_superblock_kernels(rf, monkeypatch, precision, nrhs, block):
    """Levels whose fronts have at most 384 rows (192 with more than four
    right-hand sides) and 128 pivots are solved by one fused kernel per level
    and direction; it repeats the super-block kernels' arithmetic in their
    order, so the solutions are bit-identical.  The 2D cantilever is fused at
    every level; the 9^3 block at its seven lower levels with at most four
    right-hand sides, and with more only at its third
    level (fronts of at most 192 rows and 36 pivots), between levels of the
    super-block kernels (MFMA products) on both sides.  With 16 right-hand
    sides per pass (RUSTYFEM_SOLVE_RHS_BLOCK=16) a level whose fused kernels
    would need more LDS than they may take runs the super-block kernels, which
    need the super-block inverses of its fronts."""
    monkeypatch.setenv("RUSTYFEM_SOLVE_RHS_BLOCK", str(block))
    for p in (meshes.cantilever_cps4(nx=40, ny=4), meshes.hex_block(9, nrhs=1, seed=3)):
        rng = np.random.default_rng(nrhs)
        b = rng.standard_normal((p.n, nrhs))
        print(f"case fused solve {p.name} nrhs={nrhs} block={block} precision={precision}", flush=True)
        monkeypatch.setenv("RUSTYFEM_NO_SMALL_SOLVE", "1")
        h_sb, _, _ = rf.Handle.factor(p, precision=precision, max_refine=0)
        monkeypatch.delenv("RUSTYFEM_NO_SMALL_SOLVE")
        h_f, _, _ = rf.Handle.factor(p, precision=precision, max_refine=0)
        assert h_sb.stats()["small_levels"] == 0
        nsmall = h_f.stats()["small_levels"]
        assert 0 < nsmall <= h_f.stats()["levels"]
        if p.ndof == 2:
            assert nsmall == h_f.stats()["levels"]
        else:
            assert nsmall < h_f.stats()["levels"]
        assert np.array_equal(h_f.solve(b), h_sb.solve(b))


@pytest.mark.parametrize("precision", [None, "mixed"])
@pytest.mark.parametrize("nrhs", [5, 16, 33])
def test_mfma_multi_rhs_solve_matches_valu_solve(rf, monkeypatch, precision, nrhs):
    """More than four right-hand sides: the FP64-MFMA forward row update and
    backward partial sums (16 right-hand sides per pass) against the VALU
    kernels (4 per pass) on the same factor, at the levels whose fronts do not
    fit the fused small-front kernels: the same solution to FP64 rounding (the
    sums are formed in another order), and the true residual of the FP64
    factor's solve."""
    p = meshes.hex_block(14, nrhs=1, seed=8)
    b = np.random.default_rng(nrhs).standard_normal((p.n, nrhs))
    print(f"case mfma multi-rhs nrhs={nrhs} precision={precision}", flush=True)
    monkeypatch.setenv("RUSTYFEM_NO_MRHS_MFMA", "1")
    h_v, _, _ = rf.Handle.factor(p, precision=precision, max_refine=0)
    monkeypatch.delenv("RUSTYFEM_NO_MRHS_MFMA")
    h_m, _, _ = rf.Handle.factor(p, precision=precision, max_refine=0)
    x_v = h_v.solve(b)
    x_m = h_m.solve(b)
    assert np.max(np.abs(x_m - x_v)) / np.max(np.abs(x_v)) < 1e-12
    if precision is None:
        assert np.all(relres(p, x_m, b) < 1e-12)


@pytest.mark.parametrize("precision", [None, "mixed"])
def test_gemm64_tiles_match_gemm128_tiles(rf, monkeypatch, precision):
    """GEMM tiles that update at most 64 rows and columns run on the one-wave
    64 x 64 kernel (first in their launch), which repeats the 128 x 128
    kernel's first wave: identical D and solutions with and without it
    (RUSTYFEM_NO_GEMM64) on a 2D tree (most of its update blocks are small)
    and on a 3D block (only some are, mixed with larger ones in a launch)."""
    for p in (meshes.cantilever_cps4(), meshes.hex_block(9, nrhs=2, seed=6)):
        print(f"case gemm64 {p.name} precision={precision}", flush=True)
        monkeypatch.setenv("RUSTYFEM_NO_GEMM64", "1")
        h_old, _, d_old = rf.Handle.factor(p, want_d=True, precision=precision, max_refine=0)
        monkeypatch.delenv("RUSTYFEM_NO_GEMM64")
        h_new, _, d_new = rf.Handle.factor(p, want_d=True, precision=precision, max_refine=0)
        assert np.array_equal(d_new, d_old)
        assert np.array_equal(h_new.solve(p.rhs), h_old.solve(p.rhs))


def test_inverse_and_chain_superblock_solves_agree(rf, monkeypatch):
    """Super-block solves with the explicit FP64 inverse X = L_JJ^-1 against the
    in-LDS block chain on the same factor (a front with several super-blocks)."""
    p = meshes.hex_block(14, nrhs=3, seed=8)
    monkeypatch.setenv("RUSTYFEM_SOLVE_CHAIN", "1")
    h_c, _, _ = rf.Handle.factor(p)
    monkeypatch.delenv("RUSTYFEM_SOLVE_CHAIN")
    h_i, _, _ = rf.Handle.factor(p)
    x_c = h_c.solve(p.rhs)
    x_i = h_i.solve(p.rhs)
    assert np.max(np.abs(x_i - x_c)) / np.max(np.abs(x_c)) < 1e-11
    a = p.to_scipy()
    assert np.all(np.linalg.norm(p.rhs - a @ x_i, axis=0) / np.linalg.norm(p.rhs, axis=0) < 1e-12)


@pytest.mark.parametrize("precision", [None, "mixed"])
def test_gemm_register_layouts_agree(rf, monkeypatch, precision):
    """The Schur GEMM kernels (k_gemm, k_gemm_agpr, k_gemm_ilv, k_gemm_k64 and,
    in FP32, k_gemm_q) run the same arithmetic: forcing any one gives
    identical D and solutions, and the autotuner times every GEMM launch of
    the plan."""
    p = meshes.hex_block(12, nrhs=2, seed=5)
    monkeypatch.setenv("RUSTYFEM_AUTOTUNE", "0")
    h_v, _, d_v = rf.Handle.factor(p, want_d=True, precision=precision, max_refine=0)
    monkeypatch.setenv("RUSTYFEM_AUTOTUNE", "agpr")
    h_a, _, d_a = rf.Handle.factor(p, want_d=True, precision=precision, max_refine=0)
    monkeypatch.setenv("RUSTYFEM_AUTOTUNE", "ilv")
    h_i, _, d_i = rf.Handle.factor(p, want_d=True, precision=precision, max_refine=0)
    monkeypatch.setenv("RUSTYFEM_AUTOTUNE", "k64")
    h_k, _, d_k = rf.Handle.factor(p, want_d=True, precision=precision, max_refine=0)
    monkeypatch.setenv("RUSTYFEM_AUTOTUNE", "q")
    h_q, _, d_q = rf.Handle.factor(p, want_d=True, precision=precision, max_refine=0)
    monkeypatch.delenv("RUSTYFEM_AUTOTUNE")
    h_t, _, d_t = rf.Handle.factor(p, want_d=True, precision=precision, max_refine=0)
    assert np.array_equal(d_a, d_v) and np.array_equal(d_i, d_v) and np.array_equal(d_k, d_v)
    assert np.array_equal(d_q, d_v)
    assert np.array_equal(d_t, d_v)
    x = h_v.solve(p.rhs)
    for h in (h_a, h_i, h_k, h_q, h_t):
        assert np.array_equal(h.solve(p.rhs), x)
    st = h_t.stats()
    assert st["tuned_launches"] > 0
    assert 0 <= st["agpr_launches"] + st["ilv_launches"] + st["q_launches"] <= st["tuned_launches"]
    if precision is None:
        assert st["q_launches"] == 0
    assert h_v.stats()["tuned_launches"] == 0 and h_i.stats()["tuned_launches"] == 0


@pytest.mark.parametrize("precision", [None, "mixed"])
def test_lookahead_gives_identical_factor(rf, monkeypatch, precision):
    """The rest of each outer block's update runs on a second stream beside the
    next outer block's pivot-block launches; on a block whose top fronts have
    more than 512 pivots, the factor and the solutions are identical to those
    of the single-stream schedule (RUSTYFEM_NO_LOOKAHEAD=1)."""
    p = meshes.hex_block(20, nrhs=2, seed=9)
    print(f"case lookahead hex20 n={p.n} precision={precision}", flush=True)
    monkeypatch.setenv("RUSTYFEM_NO_LOOKAHEAD", "1")
    h_s, _, d_s = rf.Handle.factor(p, want_d=True, precision=precision, max_refine=0)
    monkeypatch.delenv("RUSTYFEM_NO_LOOKAHEAD")
    h_l, _, d_l = rf.Handle.factor(p, want_d=True, precision=precision, max_refine=0)
    assert np.array_equal(d_l, d_s)
    assert np.array_equal(h_l.solve(p.rhs), h_s.solve(p.rhs))
    # a refactor on the same handle (the lookahead events are reused) as well
    h_l.refactor(p.data)
    assert np.array_equal(h_l.solve(p.rhs), h_s.solve(p.rhs))


def test_refactor_graph_replays_match_eager(rf, monkeypatch):
    """From the second factorization on, the refactorization is one captured
    HIP graph replayed per call: the factors (seen through the solutions and
    a failing pivot) equal those of the eager launches bit for bit."""
    p = meshes.hex_block(10, nrhs=2)
    print(f"case refactor graph vs eager n={p.n}", flush=True)
    out = {}
    for mode in ("graph", "eager"):
        if mode == "eager":
            monkeypatch.setenv("RUSTYFEM_NO_FACTOR_GRAPH", "1")
        else:
            monkeypatch.delenv("RUSTYFEM_NO_FACTOR_GRAPH", raising=False)
        h, err, _ = rf.Handle.factor(p, precision=0)
        assert err == -1
        xs = []
        for k in range(3):
            assert h.refactor(p.data * (1.0 + 0.25 * k)) == -1
            xs.append(h.solve(p.rhs))
        xs.append(h.refactor(-p.data))
        out[mode] = xs
        h.free()
    for a, c in zip(out["graph"][:3], out["eager"][:3]):
        np.testing.assert_array_equal(a, c)
    assert out["graph"][3] == out["eager"][3] >= 0
