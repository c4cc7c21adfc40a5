"""Trains the learned smoother (csrc/mpnn.hpp) on C3D4 tetrahedral blocks.

    python -m rustyfem_amd.train_smoother --systems 2000 --nx 32 --epochs 1 --out rustyfem_amd/data/smoother_c3d4.npz

Every system is a ``meshes.tet_block(nx, seed=s)`` (s = first_seed, first_seed
+ 1, ...: perturbed tetrahedra, lognormal stiffness with a log-std drawn per
system from U[0.5, 2]).  Worker processes generate the systems and the error
vectors the network is trained on, and stream them: by default (--errors
solver) the errors the solver itself leaves, e_k = x* - x_k for the iterates
x_k (k = 1..8) of the AMG-FCG that the learned smoother serves (FP32
V-cycle, Chebyshev smoothing) on a random load, x* its converged solution,
each passed through the solver's own coarse-grid correction (its smoothed
prolongator and Galerkin operator, exact coarse solve): the component of the
solver's error that its smoother has to remove.  (--errors gaussian trains
on random Gaussian vectors.)
Adam steps (steps_per_system per system and epoch, one load's errors each)
minimize the smoothing loss mean ||(I - M A)^nu e||_A^2 / ||e||_A^2, computed
with its gradient on the GPU (rustyfem_mpnn_loss).  The network sees
dimensionless block features (s_i s_j A_ij), so systems of a few thousand
unknowns train it for any mesh size of the same element family.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

H = 32  # hidden units (csrc/mpnn.hpp)
NPARAMS_LAYOUT = (H * 10, H, H * 3 * H, H, 9 * H, 9)


def lower_of(p):
    import scipy.sparse as sp
    a = p.to_scipy().tocoo()
    k = a.row >= a.col
    low = sp.csc_matrix((a.data[k], (a.row[k], a.col[k])), shape=a.shape)
    low.sort_indices()
    return low.indptr.astype(np.int64), low.indices.astype(np.int64), low.data.copy()


def initial_params(seed=0, scale=0.1, out_scale=1e-3):
    """Random hidden layers and a small output layer: the smoother starts at
    scaled damped Jacobi."""
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal(NPARAMS_LAYOUT[0]) * scale, rng.standard_normal(NPARAMS_LAYOUT[1]) * scale,
             rng.standard_normal(NPARAMS_LAYOUT[2]) * scale, rng.standard_normal(NPARAMS_LAYOUT[3]) * scale,
             rng.standard_normal(NPARAMS_LAYOUT[4]) * out_scale, np.zeros(NPARAMS_LAYOUT[5])]
    return np.concatenate(parts)


ITERATES = tuple(range(1, 9))  # solver iterates whose errors train the network


def solver_errors(solver, n, rng, loads=1, ks=ITERATES, tol=1e-10, a=None):
    """Errors the solver leaves: for each of `loads` random loads b, the
    iterates x_k (k in ks, from x0 = 0) of `solver` against its solution x*
    at tol, e_k = x* - x_k.  With the fine matrix `a` (scipy), each e_k is
    then passed through the solver's own coarse-grid correction with an exact
    coarse solve, e_k - P A_c^-1 P^T A e_k (P the finest level's smoothed
    prolongator, A_c = P^T A P the Galerkin operator of the hierarchy,
    RUSTYFEM_PCG_KEEP_HIERARCHY): what is left is the part of the solver's
    error that its smoother has to remove.  Returns n x (loads len(ks))."""
    out = []
    for _ in range(loads):
        b = rng.standard_normal(n)
        xs, _, _, _ = solver.solve(b, tol=tol, maxit=400)
        for k in ks:
            xk, _, _, _ = solver.solve(b, tol=1e-30, maxit=k)
            out.append(xs - xk)
    e = np.stack(out, axis=1)
    if a is not None:
        import scipy.sparse.linalg as sla
        l0, l1 = solver.level(0), solver.level(1)
        pm = l0["P"].tocsr()
        lu = sla.splu(l1["A"].tocsc())
        e = e - pm @ lu.solve(np.asarray(pm.T @ (a @ e)))
    return e


def _system(args):
    """(n, lower CSC, training errors) of tet_block(nx, seed), run in a worker
    process: the errors are the solver's (see solver_errors), or None for
    Gaussian vectors drawn by the trainer."""
    nx, s, mode = args
    from rustyfem_amd import meshes
    p = meshes.tet_block(nx, seed=s)
    low = lower_of(p)
    if mode != "solver":
        return p.n, low, None
    from rustyfem_amd import ffi
    _, coords, ndof, _, _ = meshes.node_graph(p)
    solver = ffi.Pcg(p.n, *low, coords=coords, ndof=ndof, krylov="fcg", vcycle="fp32",
                     flags=ffi.PCG_NO_DIRECT_FALLBACK | ffi.PCG_KEEP_HIERARCHY)
    e = solver_errors(solver, p.n, np.random.default_rng(s), a=p.to_scipy().tocsr())
    del solver
    return p.n, low, e.astype(np.float32)


def train(systems=256, nx=10, first_seed=1000, epochs=4, nvec=8, nu=2, lr=3e-3, seed=0, workers=12,
          steps_per_system=2, errors="solver", log=print, checkpoint=None, every=250):
    """Adam over the systems, streamed from a process pool: every system is
    generated once per epoch and gets steps_per_system steps (on its solver
    errors, or each on fresh Gaussian vectors).  With `checkpoint`, the state
    (parameters, Adam moments, position) is saved there every `every`
    systems, and a run started with an existing checkpoint continues from it
    (the system order of an epoch depends on the seed only)."""
    import multiprocessing as mp
    # the worker processes start before this process touches the GPU
    pool = mp.get_context("spawn").Pool(max(1, workers))
    from rustyfem_amd import ffi
    ffi.lib()
    assert sum(NPARAMS_LAYOUT) == ffi.mpnn_nparams()
    rng = np.random.default_rng(seed)
    w = initial_params(seed)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    b1, b2, eps = 0.9, 0.999, 1e-8
    step = 0
    hist = []
    ep0, k0 = 0, 0
    if checkpoint and os.path.exists(checkpoint):
        with np.load(checkpoint, allow_pickle=False) as z:
            w, m, v = z["w"].copy(), z["m"].copy(), z["v"].copy()
            step, ep0, k0 = int(z["step"]), int(z["epoch"]), int(z["k"])
            hist = [float(x) for x in z["hist"]]
        log(f"resuming at epoch {ep0}, system {k0} (step {step})")
    t0 = time.perf_counter()
    with pool:
        for ep in range(epochs):
            losses = []
            order = [(nx, int(s), errors) for s in first_seed + rng.permutation(systems)]
            if ep < ep0:
                continue
            skip = k0 if ep == ep0 else 0
            for k, (n, low, es) in enumerate(pool.imap(_system, order[skip:], chunksize=1), start=skip):
                h = ffi.Mpnn(n, *low)
                for st in range(steps_per_system):
                    if es is None:
                        e0 = rng.standard_normal((n, nvec))
                    else:
                        c0 = (st * nvec) % es.shape[1]
                        e0 = np.ascontiguousarray(es[:, c0:c0 + nvec], dtype=np.float64)
                    # one error set per system: the steps of a system share it
                    loss, g = h.loss(w, e0, nu=nu, grad=True)
                    gn = np.linalg.norm(g)
                    if gn > 1.0:
                        g = g / gn
                    step += 1
                    m = b1 * m + (1 - b1) * g
                    v = b2 * v + (1 - b2) * g * g
                    w = w - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
                    losses.append(loss)
                del h
                if checkpoint and (k + 1) % every == 0:
                    np.savez(checkpoint, w=w, m=m, v=v, step=step, epoch=ep, k=k + 1, hist=np.array(hist))
                if (k + 1) % 100 == 0:
                    log(f"epoch {ep}: {k + 1} systems, mean loss of the last 100 {np.mean(losses[-100 * steps_per_system:]):.5f} "
                        f"({time.perf_counter() - t0:.1f}s)")
            hist.append(float(np.mean(losses)))
            log(f"epoch {ep}: mean smoothing loss {hist[-1]:.5f} ({time.perf_counter() - t0:.1f}s)")
    return w, hist


def evaluate(params, systems=16, nx=10, first_seed=900, nvec=8, nu=2, seed=1, errors="solver"):
    """Mean smoothing loss of params and of the starting point on held-out
    systems, on the solver's errors (or Gaussian vectors)."""
    from rustyfem_amd import ffi, meshes
    rng = np.random.default_rng(seed)
    start = initial_params(0)
    start[-(NPARAMS_LAYOUT[4] + NPARAMS_LAYOUT[5]):] = 0.0
    lt, l0 = [], []
    for s in range(first_seed, first_seed + systems):
        p = meshes.tet_block(nx, seed=s)
        low = lower_of(p)
        h = ffi.Mpnn(p.n, *low)
        if errors == "solver":
            _, coords, ndof, _, _ = meshes.node_graph(p)
            solver = ffi.Pcg(p.n, *low, coords=coords, ndof=ndof, krylov="fcg", vcycle="fp32",
                             flags=ffi.PCG_NO_DIRECT_FALLBACK | ffi.PCG_KEEP_HIERARCHY)
            e0 = solver_errors(solver, p.n, np.random.default_rng(s), a=p.to_scipy().tocsr())[:, :nvec]
            del solver
        else:
            e0 = rng.standard_normal((h.n, nvec))
        lt.append(h.loss(params, e0, nu=nu))
        l0.append(h.loss(start, e0, nu=nu))
    return float(np.mean(lt)), float(np.mean(l0))


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--systems", type=int, default=256)
    ap.add_argument("--nx", type=int, default=10)
    ap.add_argument("--first-seed", type=int, default=1000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--nvec", type=int, default=8)
    ap.add_argument("--nu", type=int, default=2)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--workers", type=int, default=12, help="processes generating systems")
    ap.add_argument("--steps-per-system", type=int, default=2)
    ap.add_argument("--checkpoint", default=None, help="state file, saved every 250 systems and resumed from")
    ap.add_argument("--errors", choices=["solver", "gaussian"], default="solver",
                    help="training vectors: the solver's own iterate errors (default) or Gaussian vectors")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "data", "smoother_c3d4.npz"))
    args = ap.parse_args()

    def log(msg):
        print(f"[train] {msg}", file=sys.stderr, flush=True)

    w, hist = train(args.systems, args.nx, args.first_seed, args.epochs, args.nvec, args.nu, args.lr,
                    workers=args.workers, steps_per_system=args.steps_per_system, errors=args.errors, log=log,
                    checkpoint=args.checkpoint)
    lt, l0 = evaluate(w, nx=args.nx, nvec=args.nvec, nu=args.nu, errors=args.errors)
    log(f"held-out systems: trained loss {lt:.5f}, scaled damped Jacobi {l0:.5f}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.savez(args.out, params=w, history=np.array(hist),
             meta=json.dumps({"systems": args.systems, "nx": args.nx, "first_seed": args.first_seed,
                              "epochs": args.epochs, "nvec": args.nvec, "nu": args.nu, "lr": args.lr,
                              "errors": args.errors, "iterates": list(ITERATES) if args.errors == "solver" else None,
                              "heldout_loss": lt, "heldout_jacobi_loss": l0}))
    print(json.dumps({"heldout_loss": lt, "heldout_jacobi_loss": l0, "history": hist}), flush=True)


if __name__ == "__main__":
    main()
