"""Eigenvalue test problems with known answers.

* The 7-point Laplacian on a g^3 grid with B = I: eigenvalues
  6 - 2 sum_a cos(k_a pi / (g + 1)) (examples/solver_bench.rs:122-131), with
  a triple eigenvalue at k = (2, 1, 1) and its permutations.
* The C3D8 block with a lumped mass matrix (rho * element volume / 8 per node
  and direction), against scipy's shift-invert Lanczos.
"""
import math

import numpy as np
import scipy.sparse as sp

from rustyfem_amd import meshes


def laplacian_lowest(g, count):
    c = 2.0 * np.cos(np.arange(1, g + 1) * math.pi / (g + 1))
    lam = np.sort((((6.0 - c[:, None, None]) - c[None, :, None]) - c[None, None, :]).ravel())
    return lam[:count]


def laplacian_case(g):
    p = meshes.laplacian3d(g, 0.0)
    return p, sp.identity(p.n, format="csc")


def lumped_mass(p, rho=7850.0):
    """rho * (volume per node) on every DOF of a unit-cube block with h = 1/nx:
    each hexahedron gives h^3 / 8 to its 8 nodes."""
    coords = np.asarray(p.coords)
    nx = int(round(1.0 / np.min(np.diff(np.unique(coords[:, 0])))))
    h = 1.0 / nx
    ijk = np.rint(coords / h).astype(np.int64)
    # elements touching a node: product over axes of (1 if on the boundary else 2)
    touch = np.prod(np.where((ijk == 0) | (ijk == nx), 1, 2), axis=1)
    m_node = rho * touch * h ** 3 / 8.0
    return sp.diags(m_node[p.loc], format="csc")


def elastic_case(nx):
    p = meshes.hex_block(nx, nrhs=1)
    return p, lumped_mass(p)
