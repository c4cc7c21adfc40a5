"""Synthetic test problems: structured meshes, linear-elastic stiffness and loads.

These build the inputs that a caller of the solver ABI would pass: a symmetric
matrix in CSC form (both triangles) in the caller's DOF numbering, the node of
every DOF (``loc``) and the node coordinates.  They restate the element and
mesher rules of the reference only as far as needed to produce matrices of the
same structure:

* ``plate`` mesher node numbering and bilinear corner interpolation:
  src/model.rs:403-451;
* ``box`` mesher node numbering: src/model.rs:454-497;
* CPS4 (full 2x2 Gauss, plane stress) and CPE4 (planar B-bar):
  src/elements/continuum2d.rs:24-27, 35-80;
* C3D8 (2x2x2 Gauss, volumetric B-bar): src/elements/continuum3d.rs:1-6,
  src/elements/common.rs:142-202;
* C3D4 (linear tetrahedron, one point of weight 1/6): src/elements/solid_higher.rs:109,
  the isoparametric map of src/elements/solid_higher.rs:144-190 and the B
  matrix of src/elements/common.rs:142-174;
* isotropic D matrices: src/materials/elastic.rs:47-70;
* the symmetric part 0.5(ke + ke^T) scattered into the lower triangle:
  src/fem/assembly.rs:356-401.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

G2 = 1.0 / math.sqrt(3.0)


@dataclasses.dataclass
class Problem:
    """A linear system in the form the C ABI takes (caller numbering)."""

    name: str
    n: int                  # DOFs
    nn: int                 # nodes (all, including nodes without DOFs)
    loc: np.ndarray         # int64 (n,)  node of each DOF
    coords: np.ndarray      # float64 (nn, 3)
    indptr: np.ndarray      # int64 (n+1,) CSC, both triangles
    indices: np.ndarray     # int64 (nnz,)
    data: np.ndarray        # float64 (nnz,)
    rhs: np.ndarray         # float64 (n, nrhs), Fortran order (column per load case)
    ndof: int = 3           # DOFs per free node
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.indptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.data, self.indices, self.indptr), shape=(self.n, self.n))


# ---------------------------------------------------------------------------
# material matrices (src/materials/elastic.rs:47-70)

def dmat_plane_stress(e: float, nu: float) -> np.ndarray:
    c = e / (1.0 - nu * nu)
    return np.array([[c, c * nu, 0.0], [c * nu, c, 0.0], [0.0, 0.0, c * (1.0 - nu) / 2.0]])


def dmat_3d(e: float, nu: float, ncomp: int = 6) -> np.ndarray:
    lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = e / (2.0 * (1.0 + nu))
    d = np.zeros((ncomp, ncomp))
    d[:3, :3] = lam
    for i in range(3):
        d[i, i] = lam + 2.0 * mu
    for i in range(3, ncomp):
        d[i, i] = mu
    return d


# ---------------------------------------------------------------------------
# element stiffness matrices

def quad4_stiffness(x: np.ndarray, d: np.ndarray, thickness: float, planar_bbar: bool) -> np.ndarray:
    """CPS4 (plane stress, d 3x3) or CPE4 (planar B-bar, d 4x4) stiffness, 8x8.

    src/elements/continuum2d.rs:35-80 (integration points), common.rs:142-202 (B, B-bar)."""
    nat = np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]])
    ns = d.shape[0]
    bs, ws = [], []
    for xi, eta in ((-G2, -G2), (G2, -G2), (-G2, G2), (G2, G2)):
        dxi = 0.25 * nat[:, 0] * (1.0 + eta * nat[:, 1])
        deta = 0.25 * nat[:, 1] * (1.0 + xi * nat[:, 0])
        j = np.array([[dxi @ x[:, 0], dxi @ x[:, 1]], [deta @ x[:, 0], deta @ x[:, 1]]])
        det = j[0, 0] * j[1, 1] - j[0, 1] * j[1, 0]
        assert det > 0.0
        dndx = np.stack([(j[1, 1] * dxi - j[0, 1] * deta) / det, (-j[1, 0] * dxi + j[0, 0] * deta) / det], axis=1)
        b = np.zeros((ns, 8))
        shear = ns - 1
        b[0, 0::2] = dndx[:, 0]
        b[1, 1::2] = dndx[:, 1]
        b[shear, 0::2] = dndx[:, 1]
        b[shear, 1::2] = dndx[:, 0]
        bs.append(b)
        ws.append(det * thickness)
    if planar_bbar:
        wsum = sum(ws)
        vol = [b[0] + b[1] + (b[2] if ns == 4 else 0.0) for b in bs]
        avg = sum(w * v for w, v in zip(ws, vol)) / wsum
        for b, v in zip(bs, vol):
            corr = (avg - v) / 2.0
            b[0] += corr
            b[1] += corr
    ke = np.zeros((8, 8))
    for b, w in zip(bs, ws):
        ke += w * (b.T @ d @ b)
    return 0.5 * (ke + ke.T)


HEX8_NAT = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                     [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)


def hex8_stiffness(x: np.ndarray, d: np.ndarray) -> np.ndarray:
    """C3D8 with volumetric B-bar, 2x2x2 Gauss (src/elements/continuum3d.rs:39-80,
    src/elements/common.rs:142-202).  Returns a symmetric 24x24 matrix."""
    bs, ws = [], []
    for p in [(a, b, c) for c in (-G2, G2) for b in (-G2, G2) for a in (-G2, G2)]:
        f = 1.0 + np.asarray(p)[None, :] * HEX8_NAT
        dn = np.stack([0.125 * HEX8_NAT[:, 0] * f[:, 1] * f[:, 2],
                       0.125 * HEX8_NAT[:, 1] * f[:, 0] * f[:, 2],
                       0.125 * HEX8_NAT[:, 2] * f[:, 0] * f[:, 1]])
        j = dn @ x
        det = np.linalg.det(j)
        assert det > 0.0
        dndx = np.linalg.solve(j, dn).T  # 8 x 3
        b = np.zeros((6, 24))
        for i in range(8):
            c = 3 * i
            b[0, c] = dndx[i, 0]
            b[1, c + 1] = dndx[i, 1]
            b[2, c + 2] = dndx[i, 2]
            b[3, c] = dndx[i, 1]
            b[3, c + 1] = dndx[i, 0]
            b[4, c] = dndx[i, 2]
            b[4, c + 2] = dndx[i, 0]
            b[5, c + 1] = dndx[i, 2]
            b[5, c + 2] = dndx[i, 1]
        bs.append(b)
        ws.append(det)
    wsum = sum(ws)
    vol = [b[0] + b[1] + b[2] for b in bs]
    avg = sum(w * v for w, v in zip(ws, vol)) / wsum
    for b, v in zip(bs, vol):
        corr = (avg - v) / 3.0
        b[0] += corr
        b[1] += corr
        b[2] += corr
    ke = np.zeros((24, 24))
    for b, w in zip(bs, ws):
        ke += w * (b.T @ d @ b)
    return 0.5 * (ke + ke.T)


# ---------------------------------------------------------------------------
# structured-grid assembly

def _assemble_structured(shape_e, ke_unit, escale, free_mask, nd):
    """Assemble a structured grid of identical (up to a scale) elements.

    shape_e: element counts per axis (2 or 3 axes, x fastest in node numbering);
    ke_unit: (nen*nd, nen*nd) element matrix with local node a at offsets in {0,1}^dim
    ordered as the mesher orders them; escale: per-element scale (elements in
    x-fastest order) or None; free_mask: bool per node (x-fastest).

    Returns CSC arrays over free DOFs numbered node-major (node order, then component).
    Each column's entries come from the column node's own stencil, which makes the
    matrix exactly symmetric because the element matrix is exactly symmetric and the
    element sums run in the same order for (p, q) and (q, p).
    """
    dim = len(shape_e)
    shape_n = tuple(s + 1 for s in shape_e)
    if dim == 2:
        local = [(0, 0), (1, 0), (1, 1), (0, 1)]
    else:
        local = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    offs = [o for o in np.ndindex(*([3] * dim))]  # offsets -1..1 stored as 0..2, C order over axes
    # stencil indexed [axis-reversed node grid][offset][nd][nd]; arrays use (z, y, x) order
    grid_n = tuple(reversed(shape_n))
    grid_e = tuple(reversed(shape_e))
    sten = np.zeros(grid_n + (3 ** dim, nd, nd))
    es = np.ones(grid_e) if escale is None else np.asarray(escale, dtype=np.float64).reshape(grid_e)
    for ai, a in enumerate(local):
        for bi, b in enumerate(local):
            blk = ke_unit[ai * nd:(ai + 1) * nd, bi * nd:(bi + 1) * nd]
            dvec = tuple(bb - aa for aa, bb in zip(a, b))
            oi = offs.index(tuple(x + 1 for x in dvec))
            sl = tuple(slice(aa, aa + se) for aa, se in zip(reversed(a), grid_e))
            sten[sl + (oi,)] += es[..., None, None] * blk
    nnodes = int(np.prod(shape_n))
    sten = sten.reshape(nnodes, 3 ** dim, nd, nd)
    free = np.asarray(free_mask, dtype=bool).reshape(nnodes)
    fid = np.full(nnodes, -1, dtype=np.int64)
    fid[free] = np.arange(int(free.sum()), dtype=np.int64)
    nf = int(free.sum())
    n = nf * nd
    # neighbour node for every (node, offset)
    coords_idx = np.stack(np.unravel_index(np.arange(nnodes), grid_n), axis=1)  # (z,y,x)
    nbr = np.full((nnodes, 3 ** dim), -1, dtype=np.int64)
    for oi, o in enumerate(offs):
        dv = np.array(list(reversed([x - 1 for x in o])))  # to (z,y,x)
        q = coords_idx + dv[None, :]
        ok = np.all((q >= 0) & (q < np.array(grid_n)[None, :]), axis=1)
        qi = np.full(nnodes, -1, dtype=np.int64)
        qi[ok] = np.ravel_multi_index(tuple(q[ok].T), grid_n)
        nbr[:, oi] = qi
    # order offsets by neighbour node index (node numbering is x fastest; offsets ordered
    # z-major then y then x give increasing node ids)
    order = sorted(range(3 ** dim), key=lambda oi: tuple(offs[oi][::-1]))
    nbr = nbr[:, order]
    sten = sten[:, order]
    # exact symmetry: a block below the diagonal (neighbour q < p) is taken as the
    # transpose of q's block towards p; the sorted offsets pair index j with 3^dim-1-j
    nst = 3 ** dim
    pidx = np.arange(nnodes)[:, None]
    low = (nbr >= 0) & (nbr < pidx)
    rr_, jj_ = np.nonzero(low)
    sten[rr_, jj_] = np.swapaxes(sten[nbr[rr_, jj_], nst - 1 - jj_], -1, -2)
    nbr_f = np.where(nbr >= 0, fid[np.maximum(nbr, 0)], -1)
    valid = (nbr_f >= 0) & free[:, None]
    cnt_node = valid.sum(axis=1)  # neighbour free nodes per node
    col_cnt = np.repeat(cnt_node[free] * nd, nd)
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(col_cnt, out=indptr[1:])
    nnz = int(indptr[-1])
    indices = np.empty(nnz, dtype=np.int64)
    data = np.empty(nnz, dtype=np.float64)
    fnodes = np.nonzero(free)[0]
    # per free node: for c in comps: for neighbours (sorted): rows nd*q + r
    vmask = valid[fnodes]                      # (nf, 27)
    qf = nbr_f[fnodes]                         # (nf, 27)
    st = sten[fnodes]                          # (nf, 27, nd, nd) [p-comp c][q-comp r]
    k = int(vmask.sum(axis=1).max()) if nf else 0
    # compact the valid neighbours to the front of each row, keeping order
    rank = np.cumsum(vmask, axis=1) - 1
    qc = np.full((nf, k), -1, dtype=np.int64)
    sc = np.zeros((nf, k, nd, nd))
    rr, cc = np.nonzero(vmask)
    qc[rr, rank[rr, cc]] = qf[rr, cc]
    sc[rr, rank[rr, cc]] = st[rr, cc]
    # entries for column (node p, comp c): for each valid neighbour j, rows nd*q_j + r: value K[(q,r),(p,c)] = S[p,d][c,r]
    pos = indptr[:-1].reshape(nf, nd)
    for c in range(nd):
        base = pos[:, c]
        for j in range(k):
            ok = qc[:, j] >= 0
            for r in range(nd):
                idx = base[ok] + j * nd + r
                indices[idx] = nd * qc[ok, j] + r
                data[idx] = sc[ok, j, c, r]
    return n, fid, indptr, indices, data


def cantilever_cps4(nx=80, ny=8, length=2.0, height=0.2, thickness=0.01, e=210e9, nu=0.3, fy_total=-1000.0):
    """The 2D plane-stress cantilever: plate x=0..2, y=0..0.2, h=0.025 (80x8 CPS4),
    clamped at x=0, total fy at x=2 (examples/cantilever_plate.rusty)."""
    c = np.array([[0.0, 0.0], [length, 0.0], [length, height], [0.0, height]])
    nn = (nx + 1) * (ny + 1)
    coords = np.zeros((nn, 3))
    for j in range(ny + 1):
        for i in range(nx + 1):
            s = i / nx
            t = j / ny
            w = [(1.0 - s) * (1.0 - t), s * (1.0 - t), s * t, (1.0 - s) * t]
            x = -0.0
            y = -0.0
            for kk in range(4):
                x += w[kk] * c[kk, 0]
            for kk in range(4):
                y += w[kk] * c[kk, 1]
            coords[j * (nx + 1) + i] = (x, y, 0.0)
    hx, hy = length / nx, height / ny
    xe = np.array([[0, 0], [hx, 0], [hx, hy], [0, hy]], dtype=np.float64)
    ke = quad4_stiffness(xe, dmat_plane_stress(e, nu), thickness, planar_bbar=False)
    free = np.ones((ny + 1, nx + 1), dtype=bool)
    free[:, 0] = False
    n, fid, indptr, indices, data = _assemble_structured((nx, ny), ke, None, free.ravel(), 2)
    loc = np.repeat(np.nonzero(free.ravel())[0], 2).astype(np.int64)
    # consistent edge loads on x = length: end nodes half weight
    rhs = np.zeros((n, 1))
    wts = np.ones(ny + 1)
    wts[0] = wts[-1] = 0.5
    wts /= wts.sum()
    for j in range(ny + 1):
        node = j * (nx + 1) + nx
        rhs[2 * fid[node] + 1, 0] = fy_total * wts[j]
    return Problem("cantilever_cps4_80x8", n, nn, loc, coords, indptr, indices, data, np.asfortranarray(rhs), ndof=2,
                   meta={"nx": nx, "ny": ny, "e": e, "nu": nu})


def hex_block(nx, ny=None, nz=None, e=210e9, nu=0.3, escale=None, clamp="x0", load="pressure_x1",
              pressure=1e6, body=1e4, nrhs=1, seed=2, size=1.0):
    """C3D8 B-bar block on [0,size]^3 with nx*ny*nz elements (node numbering of the
    box mesher, src/model.rs:454-497).  clamp: 'x0' or 'z0'.  load: 'pressure_x1'
    (uniform pressure on x=size, consistent face loads) or 'body_z' (body force along -z).
    Extra right-hand sides 2..nrhs are seeded random nodal loads N(0,1)*1e3."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    hx, hy, hz = size / nx, size / ny, size / nz
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coords = np.stack([0.0 + (size - 0.0) * ii.ravel() / nx, 0.0 + (size - 0.0) * jj.ravel() / ny,
                       0.0 + (size - 0.0) * kk.ravel() / nz], axis=1)
    xe = HEX8_NAT.copy()
    xe = (xe + 1.0) / 2.0 * np.array([hx, hy, hz])
    ke = hex8_stiffness(xe, dmat_3d(1.0, nu)) * e
    free = np.ones((nz + 1, ny + 1, nx + 1), dtype=bool)
    if clamp == "x0":
        free[:, :, 0] = False
    elif clamp == "z0":
        free[0, :, :] = False
    else:
        raise ValueError(clamp)
    n, fid, indptr, indices, data = _assemble_structured((nx, ny, nz), ke, escale, free.ravel(), 3)
    loc = np.repeat(np.nonzero(free.ravel())[0], 3).astype(np.int64)
    rhs = np.zeros((n, nrhs), order="F")
    f = np.zeros((nn, 3))
    if load == "pressure_x1":
        # each face quad of area hy*hz on x=size gives p*area/4 to its 4 nodes, along -x
        w = np.zeros((nz + 1, ny + 1))
        q = pressure * hy * hz / 4.0
        w[:-1, :-1] += q
        w[1:, :-1] += q
        w[:-1, 1:] += q
        w[1:, 1:] += q
        nodes = (np.arange(nz + 1)[:, None] * (ny + 1) + np.arange(ny + 1)[None, :]) * (nx + 1) + nx
        f[nodes.ravel(), 0] -= w.ravel()
    elif load == "body_z":
        vol = hx * hy * hz
        w = np.zeros((nz + 1, ny + 1, nx + 1))
        q = body * vol / 8.0
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    w[dz:dz + nz, dy:dy + ny, dx:dx + nx] += q
        f[:, 2] -= w.ravel()
    else:
        raise ValueError(load)
    fn = np.nonzero(free.ravel())[0]
    rhs[:, 0] = f[fn].ravel()
    if nrhs > 1:
        rng = np.random.default_rng(seed)
        rhs[:, 1:] = rng.standard_normal((n, nrhs - 1)) * 1e3
    return Problem(f"hex_c3d8_{nx}x{ny}x{nz}", n, nn, loc, coords, indptr, indices, data, rhs, ndof=3,
                   meta={"nx": nx, "ny": ny, "nz": nz, "e": e, "nu": nu, "clamp": clamp, "load": load})


def grid_elements(shape_e):
    """Connectivity of the structured grid's elements (the box and plate
    meshers' node numbering, x fastest): (prod(shape_e), 8) for three axes in
    the C3D8 corner order (HEX8_NAT), (prod(shape_e), 4) for two axes in the
    CPS4/CPE4 order (counter-clockwise from (0, 0)); elements x fastest."""
    shape_n = [s + 1 for s in shape_e]
    if len(shape_e) == 2:
        nx, ny = shape_e
        jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        base = (ii + shape_n[0] * jj).ravel()
        corner = np.array([0, 1, 1 + shape_n[0], shape_n[0]])
    else:
        nx, ny, nz = shape_e
        kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        base = (ii + shape_n[0] * (jj + shape_n[1] * kk)).ravel()
        nat = (HEX8_NAT.astype(np.int64) + 1) // 2
        corner = nat[:, 0] + shape_n[0] * (nat[:, 1] + shape_n[1] * nat[:, 2])
    return (base[:, None] + corner[None, :]).astype(np.int64)


def equation_numbers(free, dim):
    """(nn, dim) equation numbers of the free nodes, numbered node-major
    (node order, then component) as the problems here number their DOFs; -1
    for the DOFs of fixed nodes."""
    free = np.asarray(free, dtype=bool).ravel()
    eq = np.full((free.size, dim), -1, dtype=np.int64)
    nf = int(free.sum())
    eq[free] = np.arange(nf * dim, dtype=np.int64).reshape(nf, dim)
    return eq


def plate_cpe4(nx, ny, e, nu, nrhs=8, seed=3, thickness=1.0):
    """A CPE4 (planar B-bar) plate of unit height: h = 1/ny, length nx*h, clamped at
    x=0, with `nrhs` random edge-traction load cases on the other three edges."""
    h = 1.0 / ny
    length = nx * h
    c = np.array([[0.0, 0.0], [length, 0.0], [length, 1.0], [0.0, 1.0]])
    nn = (nx + 1) * (ny + 1)
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    s = ii.ravel() / nx
    t = jj.ravel() / ny
    coords = np.zeros((nn, 3))
    coords[:, 0] = s * (1 - t) * c[1, 0] + s * t * c[2, 0]
    coords[:, 1] = s * t * c[2, 1] + (1 - s) * t * c[3, 1]
    xe = np.array([[0, 0], [h, 0], [h, h], [0, h]], dtype=np.float64)
    ke = quad4_stiffness(xe, dmat_3d(e, nu, ncomp=4), thickness, planar_bbar=True)
    free = np.ones((ny + 1, nx + 1), dtype=bool)
    free[:, 0] = False
    n, fid, indptr, indices, data = _assemble_structured((nx, ny), ke, None, free.ravel(), 2)
    loc = np.repeat(np.nonzero(free.ravel())[0], 2).astype(np.int64)
    rng = np.random.default_rng(seed)
    rhs = np.zeros((n, nrhs), order="F")
    edges = {
        "bottom": [j * 0 * (nx + 1) + i for i in range(nx + 1) for j in [0]],
        "top": [ny * (nx + 1) + i for i in range(nx + 1)],
        "right": [j * (nx + 1) + nx for j in range(ny + 1)],
    }
    for lc in range(nrhs):
        for name, nodes in edges.items():
            tx, ty = rng.standard_normal(2) * 1e6
            wts = np.ones(len(nodes)) * h
            wts[0] *= 0.5
            wts[-1] *= 0.5
            for nd_, w in zip(nodes, wts):
                if fid[nd_] >= 0:
                    rhs[2 * fid[nd_], lc] += tx * w
                    rhs[2 * fid[nd_] + 1, lc] += ty * w
    return Problem(f"plate_cpe4_{nx}x{ny}", n, nn, loc, coords, indptr, indices, data, rhs, ndof=2,
                   meta={"nx": nx, "ny": ny, "e": e, "nu": nu})


def plate_batch(count=512, seed=3, nrhs=8):
    """Batch of independent CPE4 plates with nx ~ U{20..200}, ny = nx // 10,
    E ~ U[70e9, 210e9], nu ~ U[0.25, 0.35], each with `nrhs` load cases."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nx = int(rng.integers(20, 201))
        ny = max(nx // 10, 1)
        e = float(rng.uniform(70e9, 210e9))
        nu = float(rng.uniform(0.25, 0.35))
        out.append(plate_cpe4(nx, ny, e, nu, nrhs=nrhs, seed=int(rng.integers(0, 2**31 - 1))))
    return out


def lognormal_field(shape, log_std, corr_len, seed, clip_contrast=1e3):
    """Unit-variance Gaussian random field on a cell grid (FFT filtering with a
    Gaussian kernel of correlation length `corr_len` in unit-cube coordinates),
    exponentiated with log-std `log_std` and clipped to `clip_contrast` around
    the median.  Returns multipliers (median 1) in (z, y, x) order."""
    rng = np.random.default_rng(seed)
    g = rng.standard_normal(shape)
    freqs = [np.fft.fftfreq(s) * s for s in shape]  # cycles per unit length
    kz, ky, kx = np.meshgrid(*freqs, indexing="ij")
    k2 = kz ** 2 + ky ** 2 + kx ** 2
    filt = np.exp(-0.5 * (2 * np.pi) ** 2 * k2 * (corr_len ** 2) / 2.0)
    f = np.real(np.fft.ifftn(np.fft.fftn(g) * filt))
    f = (f - f.mean()) / max(f.std(), 1e-300)
    m = np.exp(log_std * f)
    lim = math.sqrt(clip_contrast)
    return np.clip(m, 1.0 / lim, lim)


def hetero_block(nx, log_std=1.5, corr_len=0.1, seed=4, e=210e9, nu=0.3, body=1e4):
    """The heterogeneous-stiffness block: C3D8 B-bar on nx^3 cells of the unit cube,
    lognormal Young's modulus per element, clamped base z=0, body load along -z."""
    field = lognormal_field((nx, nx, nx), log_std, corr_len, seed)
    return hex_block(nx, e=e, nu=nu, escale=field.ravel(), clamp="z0", load="body_z", body=body)


# Kuhn split of a cube into six tetrahedra around the diagonal from corner 0 to
# corner 7 (corner bits: x = 1, y = 2, z = 4); every tetrahedron follows one
# monotone path 0 -> 7, so all cubes of a grid split compatibly and a node's
# neighbours are the 14 offsets d and -d with d in {0,1}^3 \ {0}.
KUHN = ((0, 1, 3, 7), (0, 1, 5, 7), (0, 2, 3, 7), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 6, 7))
TET_DN = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def tet4_stiffness(x: np.ndarray, e: np.ndarray, nu: float):
    """C3D4 element matrices for a batch of tetrahedra x (ne, 4, 3) with Young's
    moduli e (ne,): K = V B^T D B with the constant B of the linear tetrahedron
    (one integration point, weight 1/6 det J = V), symmetrized as the assembly
    does.  Returns (K (ne, 12, 12), V (ne,)); raises on a non-positive det J
    (src/elements/solid_higher.rs:190)."""
    j = x[:, 1:, :] - x[:, :1, :]                     # J[b][c] = d x_c / d xi_b
    det = np.linalg.det(j)
    if not np.all(det > 0):
        raise ValueError("distorted tetrahedron (det J <= 0)")
    dndx = np.matmul(TET_DN, np.swapaxes(np.linalg.inv(j), 1, 2))  # dN_a/dx_c = sum_r Jinv[c][r] dN_a/dxi_r
    ne = x.shape[0]
    b = np.zeros((ne, 6, 12))
    for a in range(4):
        c = 3 * a
        d0, d1, d2 = dndx[:, a, 0], dndx[:, a, 1], dndx[:, a, 2]
        b[:, 0, c] = d0
        b[:, 1, c + 1] = d1
        b[:, 2, c + 2] = d2
        b[:, 3, c] = d1
        b[:, 3, c + 1] = d0
        b[:, 4, c] = d2
        b[:, 4, c + 2] = d0
        b[:, 5, c + 1] = d2
        b[:, 5, c + 2] = d1
    vol = det / 6.0
    d = dmat_3d(1.0, nu)
    k = np.matmul(np.swapaxes(b, 1, 2), np.matmul(d, b)) * (vol * e)[:, None, None]
    return 0.5 * (k + np.swapaxes(k, 1, 2)), vol


def tet_block(nx, log_std=None, corr_len=0.1, seed=42, e=210e9, nu=0.3, body=1e4, perturb=0.2, chunk=400000,
              clamp="z0", min_quality=0.1, elements=False):
    """C3D4 block on nx^3 cubes of the unit cube, each split into six tetrahedra
    (KUHN), interior vertices moved by U[-perturb h, perturb h]^3 (the moves of
    the vertices of a tetrahedron left with less than min_quality of its
    undistorted volume are halved until none is left), a lognormal
    Young's modulus per cube (log-std drawn from U[0.5, 2] when not given,
    correlation length corr_len), clamped base z = 0 (clamp=None: no support),
    body load along -z.
    Nodes are numbered as the box mesher numbers them (x fastest).  With
    elements=True the connectivity and per-element Young's moduli are kept in
    meta["conn"], meta["young"] and the equation numbers in meta["eq"]."""
    rng = np.random.default_rng(seed)
    if log_std is None:
        log_std = float(rng.uniform(0.5, 2.0))
    h = 1.0 / nx
    n1 = nx + 1
    nn = n1 ** 3
    kk, jj, ii = np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1), indexing="ij")
    coords = np.stack([ii.ravel() * h, jj.ravel() * h, kk.ravel() * h], axis=1)
    inner = ((ii > 0) & (ii < nx) & (jj > 0) & (jj < nx) & (kk > 0) & (kk < nx)).ravel()
    disp = np.zeros_like(coords)
    disp[inner] = rng.uniform(-perturb * h, perturb * h, size=(int(inner.sum()), 3))
    field = lognormal_field((nx, nx, nx), log_std, corr_len, int(rng.integers(1 << 31)))
    # corner c of cube (i, j, k) is node (i + c&1) + n1 (j + (c>>1)&1) + n1^2 (k + (c>>2)&1)
    ck, cj, ci = np.meshgrid(np.arange(nx), np.arange(nx), np.arange(nx), indexing="ij")
    base = (ci + n1 * (cj + n1 * ck)).ravel()
    corner = np.array([(c & 1) + n1 * ((c >> 1) & 1) + n1 * n1 * ((c >> 2) & 1) for c in range(8)])
    tets = np.array(KUHN)
    # orient every tetrahedron positively on the undistorted cube (a corner swap)
    cube = np.array([[c & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], dtype=np.float64)
    for t in range(6):
        if np.linalg.det(cube[tets[t, 1:]] - cube[tets[t, 0]]) < 0:
            tets[t, [2, 3]] = tets[t, [3, 2]]
    conn = (base[:, None, None] + corner[tets][None, :, :]).reshape(-1, 4)   # (6 nx^3, 4), cube-major
    # every tetrahedron keeps at least min_quality of its undistorted volume:
    # the vertices of one that does not are pulled halfway back, until none is left
    for _ in range(40):
        x = coords + disp
        bad = []
        for s0 in range(0, conn.shape[0], chunk):
            v = x[conn[s0:s0 + chunk]]
            det = np.linalg.det(v[:, 1:] - v[:, :1])
            bad.append(s0 + np.nonzero(det < min_quality * h ** 3)[0])
        bad = np.concatenate(bad)
        if bad.size == 0:
            break
        disp[np.unique(conn[bad])] *= 0.5
    else:
        raise ValueError("tet_block: distorted tetrahedra remain")
    coords = coords + disp
    ecube = np.repeat(np.arange(nx ** 3), 6)
    estiff = e * field.ravel()[ecube]
    # stencil offsets sorted by node index: d in {-1,0,1}^3 with all components of one sign
    offs = sorted({(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
                   if min(dz, dy, dx) >= 0 or max(dz, dy, dx) <= 0})
    assert len(offs) == 15
    oid = {o: i for i, o in enumerate(offs)}
    # offset index of every ordered pair (a, b) of local nodes of each of the 6 tetrahedra
    pair_off = np.zeros((6, 4, 4), dtype=np.int64)
    cvec = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)])
    for t in range(6):
        for a in range(4):
            for b_ in range(4):
                dv = tuple(int(v) for v in cvec[tets[t, b_]] - cvec[tets[t, a]])
                pair_off[t, a, b_] = oid[dv]
    sten = np.zeros((nn * 15, 9))
    f = np.zeros((nn, 3))
    ne = conn.shape[0]
    for s0 in range(0, ne, chunk):
        s1 = min(ne, s0 + chunk)
        cn = conn[s0:s1]
        ke, vol = tet4_stiffness(coords[cn], estiff[s0:s1], nu)
        tloc = np.arange(s0, s1) % 6
        for a in range(4):
            f[:, 2] -= np.bincount(cn[:, a], weights=body * vol / 4.0, minlength=nn)
        idx = np.concatenate([cn[:, a] * 15 + pair_off[tloc, a, b_] for a in range(4) for b_ in range(4)])
        blk = np.concatenate([ke[:, 3 * a:3 * a + 3, 3 * b_:3 * b_ + 3].reshape(-1, 9)
                              for a in range(4) for b_ in range(4)])
        for q in range(9):
            sten[:, q] += np.bincount(idx, weights=blk[:, q], minlength=nn * 15)
    sten = sten.reshape(nn, 15, 3, 3)
    # exact symmetry: the block towards a lower-numbered neighbour is the transpose
    # of that neighbour's block towards this node (offsets o and 14 - o are
    # opposite), and diagonal blocks are symmetrized
    zi, yi, xi = np.unravel_index(np.arange(nn), (n1, n1, n1))
    for o in range(8, 15):
        dz, dy, dx = offs[o]
        z, y, x = zi + dz, yi + dy, xi + dx
        ok = (z < n1) & (y < n1) & (x < n1)
        src = np.nonzero(ok)[0]
        dst = (x + n1 * (y + n1 * z))[ok]
        sten[dst, 14 - o] = np.swapaxes(sten[src, o], -1, -2)
    sten[:, 7] = 0.5 * (sten[:, 7] + np.swapaxes(sten[:, 7], -1, -2))
    free = np.ones(nn, dtype=bool)
    if clamp == "z0":
        free[(kk == 0).ravel()] = False
    elif clamp is not None:
        raise ValueError(clamp)
    n, indptr, indices, data = _stencil_csc(sten, offs, (n1, n1, n1), free)
    loc = np.repeat(np.nonzero(free)[0], 3).astype(np.int64)
    rhs = np.zeros((n, 1), order="F")
    rhs[:, 0] = f[free].ravel()
    meta = {"nx": nx, "log_std": log_std, "seed": seed, "e": e, "nu": nu}
    if elements:
        meta.update(conn=conn.astype(np.int64), young=estiff, eq=equation_numbers(free, 3))
    return Problem(f"tet_c3d4_{nx}", n, nn, loc, coords, indptr, indices, data, rhs, ndof=3, meta=meta)


def _stencil_csc(sten, offs, shape_n, free):
    """CSC (both triangles) over free DOFs, node-major, from per-node 3x3 stencil
    blocks sten[p, o] = K(p, p + offs[o]) with offsets sorted by neighbour index:
    entry (row (q, r), column (p, c)) = sten[p, o(q - p)][c, r] (the matrix is
    symmetric block-wise, sten[q, o(p - q)] = sten[p, o(q - p)]^T)."""
    nz_, ny_, nx_ = shape_n
    nn = nz_ * ny_ * nx_
    zi, yi, xi = np.unravel_index(np.arange(nn), shape_n)
    nbr = np.full((nn, len(offs)), -1, dtype=np.int64)
    for o, (dz, dy, dx) in enumerate(offs):
        z, y, x = zi + dz, yi + dy, xi + dx
        ok = (z >= 0) & (z < nz_) & (y >= 0) & (y < ny_) & (x >= 0) & (x < nx_)
        nbr[ok, o] = (x + nx_ * (y + ny_ * z))[ok]
    fid = np.full(nn, -1, dtype=np.int64)
    fid[free] = np.arange(int(free.sum()))
    nbr_f = np.where(nbr >= 0, fid[np.maximum(nbr, 0)], -1)
    valid = (nbr_f >= 0) & free[:, None]
    fnodes = np.nonzero(free)[0]
    nf = len(fnodes)
    n = 3 * nf
    cnt = valid[fnodes].sum(axis=1)
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.repeat(3 * cnt, 3), out=indptr[1:])
    nnz = int(indptr[-1])
    indices = np.empty(nnz, dtype=np.int64)
    data = np.empty(nnz)
    vm = valid[fnodes]
    rank = np.cumsum(vm, axis=1) - 1
    k = int(cnt.max()) if nf else 0
    qc = np.full((nf, k), -1, dtype=np.int64)
    sc = np.zeros((nf, k, 3, 3))
    rr, cc = np.nonzero(vm)
    qc[rr, rank[rr, cc]] = nbr_f[fnodes][rr, cc]
    sc[rr, rank[rr, cc]] = sten[fnodes][rr, cc]
    pos = indptr[:-1].reshape(nf, 3)
    for c in range(3):
        for j in range(k):
            ok = qc[:, j] >= 0
            for r in range(3):
                idx = pos[ok, c] + 3 * j + r
                indices[idx] = 3 * qc[ok, j] + r
                data[idx] = sc[ok, j, c, r]
    return n, indptr, indices, data


def node_graph(p: Problem):
    """Node-level view of a problem: compact nodes with DOFs (in DOF order), their
    coordinates, DOF counts and adjacency (CSR).  Used by the iterative path and tests."""
    nodes, first = np.unique(p.loc, return_index=True)
    order = np.argsort(first, kind="stable")
    nodes = nodes[order]
    cid = np.full(p.nn, -1, dtype=np.int64)
    cid[nodes] = np.arange(len(nodes))
    ndof = np.bincount(cid[p.loc], minlength=len(nodes)).astype(np.int64)
    col_node = np.repeat(cid[p.loc], np.diff(p.indptr))
    row_node = cid[p.loc[p.indices]]
    m = col_node != row_node
    a, b = col_node[m], row_node[m]
    key = np.unique(a * len(nodes) + b)
    aa, bb = key // len(nodes), key % len(nodes)
    adj_ptr = np.zeros(len(nodes) + 1, dtype=np.int64)
    np.add.at(adj_ptr, aa + 1, 1)
    adj_ptr = np.cumsum(adj_ptr)
    return nodes, p.coords[nodes].copy(), ndof, adj_ptr, bb.astype(np.int64)


# ---------------------------------------------------------------------------
# shifted 3D Laplacian (examples/solver_bench.rs:87-157): the reference's
# indefinite benchmark problem


def laplacian_sigma(g: int):
    """sigma midway between two distinct eigenvalues of the g^3 7-point
    Laplacian near its 10 % quantile, and the number of eigenvalues below it
    (examples/solver_bench.rs:122-141: eigenvalues 6 - sum_a 2 cos(k_a pi/(g+1)))."""
    c = 2.0 * np.cos(np.arange(1, g + 1) * math.pi / (g + 1))
    lam = np.sort((((6.0 - c[:, None, None]) - c[None, :, None]) - c[None, None, :]).ravel())
    q = g ** 3 // 10
    while lam[q + 1] - lam[q] < 1e-9:
        q += 1
    sigma = 0.5 * (lam[q] + lam[q + 1])
    return float(sigma), int(np.count_nonzero(lam < sigma))


def laplacian3d(g: int, sigma: float = 0.0, nrhs: int = 1, seed: int = 5) -> Problem:
    """The 7-point Laplacian on a g^3 grid (unit spacing, node i + g j + g^2 k at
    (i, j, k), one DOF per node) shifted by sigma: diagonal 6 - sigma, -1 to
    each grid neighbour (examples/solver_bench.rs:87-157); both triangles in
    CSC, sorted rows.  Right-hand sides: seeded N(0, 1)."""
    nn = g ** 3
    k, j, i = np.meshgrid(np.arange(g), np.arange(g), np.arange(g), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    v = i + g * (j + g * k)
    rows, cols, vals = [v], [v], [np.full(nn, 6.0 - sigma)]
    for di, dj, dk in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        ii, jj, kk = i + di, j + dj, k + dk
        ok = (ii >= 0) & (ii < g) & (jj >= 0) & (jj < g) & (kk >= 0) & (kk < g)
        rows.append(ii[ok] + g * (jj[ok] + g * kk[ok]))
        cols.append(v[ok])
        vals.append(np.full(int(ok.sum()), -1.0))
    import scipy.sparse as sp
    a = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nn, nn))
    a.sort_indices()
    coords = np.stack([i, j, k], axis=1).astype(np.float64)
    rhs = np.asfortranarray(np.random.default_rng(seed).standard_normal((nn, nrhs)))
    return Problem(name=f"laplace3d_{g}", n=nn, nn=nn, loc=np.arange(nn, dtype=np.int64), coords=coords,
                   indptr=a.indptr.astype(np.int64), indices=a.indices.astype(np.int64), data=a.data.copy(), rhs=rhs,
                   ndof=1, meta={"sigma": sigma})
