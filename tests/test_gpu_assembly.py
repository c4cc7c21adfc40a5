"""GPU element assembly (rustyfem_assembly_*) through the C ABI.

Bar: the assembled values equal the oracle's sequential restatement of the
reference's element loop and scatter (oracle.assemble) bit for bit (compared
with ==, which leaves the sign of exact zeros open), for every element type,
on distorted meshes with per-element materials, on collapsed (repeated-node)
elements, and at the 811,200-DOF C3D8 block's full size.  The pattern equals
the structured assemblies' of rustyfem_amd/meshes.py.  Assembling straight
into a factorization's layout (rustyfem_assembly_refactor) must give the
same solution bits as factoring the host-assembled matrix.
"""
import numpy as np
import pytest

from rustyfem_amd import meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rf():
    from rustyfem_amd import ffi
    ffi.lib()
    return ffi


class _Block:
    """The node and DOF data of a box-mesher block clamped at x = 0."""

    def __init__(self, nx, ny, nz, with_matrix):
        kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
        self.coords = np.stack([ii.ravel() / nx, jj.ravel() / ny, kk.ravel() / nz], axis=1)
        self.nn = self.coords.shape[0]
        self.free = (ii > 0).ravel()
        self.n = 3 * int(self.free.sum())
        self.loc = np.repeat(np.nonzero(self.free)[0], 3).astype(np.int64)
        self.rhs = np.zeros((self.n, 1), order="F")
        self.indptr = self.indices = None
        if with_matrix:
            q = meshes.hex_block(nx, ny, nz)
            assert q.n == self.n and np.array_equal(q.coords, self.coords)
            self.indptr, self.indices = q.indptr, q.indices
        self.inner = ((ii > 0) & (ii < nx) & (jj > 0) & (jj < ny) & (kk > 0) & (kk < nz)).ravel()


def perturbed_hex(nx, ny, nz, seed, amp=0.2, with_matrix=True):
    rng = np.random.default_rng(seed)
    p = _Block(nx, ny, nz, with_matrix)
    coords = p.coords.copy()
    h = np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])
    coords[p.inner] += rng.uniform(-amp, amp, (int(p.inner.sum()), 3)) * h
    conn = meshes.grid_elements((nx, ny, nz))
    return p, conn, coords, meshes.equation_numbers(p.free, 3)


def perturbed_plate(nx, ny, seed, amp=0.2):
    rng = np.random.default_rng(seed)
    p = meshes.plate_cpe4(nx, ny, 7e10, 0.3, nrhs=1)
    coords = p.coords.copy()
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    inner = ((ii > 0) & (ii < nx) & (jj > 0) & (jj < ny)).ravel()
    coords[inner, :2] += rng.uniform(-amp, amp, (int(inner.sum()), 2)) / ny
    free = np.ones((ny + 1, nx + 1), dtype=bool)
    free[:, 0] = False
    return p, meshes.grid_elements((nx, ny)), coords, meshes.equation_numbers(free, 2)


def materials(ne, seed):
    rng = np.random.default_rng(seed)
    return rng.lognormal(np.log(2e11), 1.0, ne), rng.uniform(0.05, 0.45, ne), rng.uniform(0.5, 2.0, ne)


def check_bits(rf, ora, etype, conn, coords, eq, n, young, nu, thick, pattern=None):
    a = rf.Assembly(etype, conn, coords, eq, n)
    if pattern is not None:
        np.testing.assert_array_equal(a.indptr, pattern[0])
        np.testing.assert_array_equal(a.indices, pattern[1])
    got = a.values(young, nu, thick)
    want = ora.assemble(etype, conn, coords, eq, n, young, nu, a.indptr, a.indices, thickness=thick)
    assert not isinstance(want, int)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} of {got.size} entries differ, e.g. {got[bad[:3]]} vs {want[bad[:3]]}"
    return a, got


@pytest.mark.parametrize("etype", ["C3D8", "C3D4", "CPS4", "CPE4"])
def test_bit_exact_against_oracle(rf, ora, etype):
    print(f"case {etype}", flush=True)
    if etype == "C3D8":
        p, conn, coords, eq = perturbed_hex(7, 5, 4, seed=1)
    elif etype == "C3D4":
        p = meshes.tet_block(6, seed=2, elements=True)
        conn, coords, eq = p.meta["conn"], p.coords, p.meta["eq"]
    else:
        p, conn, coords, eq = perturbed_plate(12, 5, seed=3)
    young, nu, thick = materials(conn.shape[0], 4)
    check_bits(rf, ora, etype, conn, coords, eq, p.n, young, nu, thick if etype in ("CPS4", "CPE4") else None,
               pattern=(p.indptr, p.indices))


def test_matches_structured_assembly(rf):
    p = meshes.hex_block(6, 5, 4, e=210e9, nu=0.3)
    free = np.ones((5, 6, 7), dtype=bool)
    free[:, :, 0] = False
    a = rf.Assembly("C3D8", meshes.grid_elements((6, 5, 4)), p.coords, meshes.equation_numbers(free, 3), p.n)
    got = a.values(210e9, 0.3)
    np.testing.assert_allclose(got, p.data, rtol=0, atol=1e-12 * np.abs(p.data).max())


def test_collapsed_hexahedra(rf, ora):
    """Wedges meshed as hexahedra with a repeated node: the reference adds
    every local pair, the repeated ones included, in its loop order."""
    p, conn, coords, eq = perturbed_hex(4, 3, 3, seed=5, amp=0.1)
    conn = conn.copy()
    # collapse the top edge 6-7 of every other element into node 6
    conn[::2, 7] = conn[::2, 6]
    young, nu, _ = materials(conn.shape[0], 6)
    check_bits(rf, ora, "C3D8", conn, coords, eq, p.n, young, nu, None)


def test_two_calls_and_device_output(rf, ora):
    from rustyfem_amd import hip
    p, conn, coords, eq = perturbed_hex(5, 4, 3, seed=7)
    a = rf.Assembly("C3D8", conn, coords, eq, p.n)
    y1, nu1, _ = materials(conn.shape[0], 8)
    y2, nu2, _ = materials(conn.shape[0], 9)
    first = a.values(y1, nu1)
    out = hip.DeviceArray(a.nnz)
    a.values(y2, nu2, out=out.ptr)
    hip.synchronize()
    second = out.download()
    out.free()
    np.testing.assert_array_equal(first, ora.assemble("C3D8", conn, coords, eq, p.n, y1, nu1, a.indptr, a.indices))
    np.testing.assert_array_equal(second, ora.assemble("C3D8", conn, coords, eq, p.n, y2, nu2, a.indptr, a.indices))


def test_distorted_element_is_reported(rf):
    p, conn, coords, eq = perturbed_hex(4, 3, 3, seed=2, amp=0.05)
    conn = conn.copy()
    for e in (20, 9):
        conn[e] = conn[e][[1, 0, 3, 2, 5, 4, 7, 6]]
    a = rf.Assembly("C3D8", conn, coords, eq, p.n)
    with pytest.raises(rf.DistortedElement) as ex:
        a.values(2e11, 0.3)
    assert ex.value.element == 9


def test_invalid_input_is_rejected(rf):
    p, conn, coords, eq = perturbed_hex(3, 2, 2, seed=2)
    bad = conn.copy()
    bad[3, 2] = p.nn
    with pytest.raises(RuntimeError, match="out of range"):
        rf.Assembly("C3D8", bad, coords, eq, p.n)
    with pytest.raises(RuntimeError, match="no DOF"):
        rf.Assembly("C3D8", conn, coords, eq, p.n + 3)


@pytest.mark.parametrize("precision", [0, 1])
def test_refactor_through_the_factor_layout(rf, precision):
    p, conn, coords, eq = perturbed_hex(8, 6, 6, seed=11)
    a = rf.Assembly("C3D8", conn, coords, eq, p.n)
    y1, nu1, _ = materials(conn.shape[0], 12)
    y2, nu2, _ = materials(conn.shape[0], 13)
    prob = meshes.Problem("asm", p.n, p.nn, p.loc, coords, a.indptr, a.indices, a.values(y1, nu1), p.rhs)
    h, err, _ = rf.Handle.factor(prob, precision=precision)
    assert h is not None and err == -1
    assert a.refactor(h, y2, nu2) == -1
    b = np.random.default_rng(0).standard_normal((p.n, 2))
    x1 = h.solve(b)
    prob2 = meshes.Problem("asm2", p.n, p.nn, p.loc, coords, a.indptr, a.indices, a.values(y2, nu2), p.rhs)
    h2, err2, _ = rf.Handle.factor(prob2, precision=precision)
    assert err2 == -1
    x2 = h2.solve(b)
    np.testing.assert_array_equal(x1, x2)
    # a second refactor on the bound layout, back to the first material
    assert a.refactor(h, y1, nu1) == -1
    h3, _, _ = rf.Handle.factor(prob, precision=precision)
    np.testing.assert_array_equal(h.solve(b), h3.solve(b))
    assert a.stats()["bound_gather_ms"] > 0


def test_refactor_rejects_another_pattern(rf):
    p, conn, coords, eq = perturbed_hex(4, 3, 3, seed=3)
    a = rf.Assembly("C3D8", conn, coords, eq, p.n)
    h, err, _ = rf.Handle.factor(meshes.hex_block(3, 3, 3))
    assert err == -1
    with pytest.raises(RuntimeError, match="pattern"):
        a.refactor(h, 2e11, 0.3)


def test_full_size_c3d8_block(rf, ora):
    """The 811,200-DOF C3D8 block (64^3 elements, x = 0 clamped), perturbed,
    with a lognormal Young's modulus per element: bit-exact at full size."""
    n = 64
    p, conn, coords, eq = perturbed_hex(n, n, n, seed=21, with_matrix=False)
    assert p.n == 811200
    young, nu, _ = materials(conn.shape[0], 22)
    print("case full-size C3D8", flush=True)
    a, _ = check_bits(rf, ora, "C3D8", conn, coords, eq, p.n, young, nu, None)
    # the pattern: every column holds the rows of its node's neighbours (27 nodes inside)
    assert a.nnz == 63695790
    st = a.stats()
    nfree = (eq[conn] >= 0).reshape(conn.shape[0], -1).sum(axis=1)
    assert st["contributions"] == int((nfree.astype(np.int64) ** 2).sum())
    print(st, flush=True)
