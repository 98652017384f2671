"""Eigenvalue adjoint through torch.autograd (a thin test wrapper).

For A(theta, w) = L_w + diag(theta), L_w the weighted Laplacian of a graph
((L_w)_ii = sum of incident weights, (L_w)_ij = -w_ij), and the loss
loss = sum_k c_k lambda_k over the nev smallest eigenpairs, the
Hellmann-Feynman theorem gives, for M-normalised eigenvectors v_k,

    d lambda_k / d A_ij = v_ik v_jk    (- lambda_k m-term for K x = lambda M x)

so dloss/dA restricted to A's pattern is the SDDMM
G_ij = sum_k c_k v_ik v_jk, which trl_hip_sddmm_adjoint evaluates on the
device.  The chain rule to the parameters is then

    dloss/dtheta_i = G_ii,    dloss/dw_e = G_ii + G_jj - G_ij - G_ji  (e = (i, j)).

The forward solve is the FP64 thick-restart block Lanczos of
libtrl_hip_solver.so; the reference has no adjoint (it stops at the
eigenpairs, include/trl/eigensolvers/lanczos.hh:22-312).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from . import problems as P
from .solver import DeviceCSR, EigensolverParams, solve


def diag_slots(csr: P.CSR) -> np.ndarray:
    """CSR slot of each diagonal entry (every row must hold one)."""
    rows = np.repeat(np.arange(csr.n), np.diff(csr.rowptr))
    on_diag = rows == csr.col
    slot = np.full(csr.n, -1, dtype=np.int64)
    slot[rows[on_diag]] = np.nonzero(on_diag)[0]
    if (slot < 0).any():
        raise ValueError("every row needs a diagonal entry")
    return slot


def grid_values(grid: P.WeightedGrid, dslot: np.ndarray, theta: np.ndarray, w: np.ndarray) -> np.ndarray:
    """CSR values of L_w + diag(theta) on the grid's fixed pattern."""
    n, e = grid.nx * grid.ny, grid.edges
    deg = np.bincount(e[:, 0], weights=w, minlength=n) + np.bincount(e[:, 1], weights=w, minlength=n)
    val = np.empty(grid.csr.nnz)
    val[dslot] = deg + theta
    val[grid.edge_slot[:, 0]] = -w
    val[grid.edge_slot[:, 1]] = -w
    return val


def param_grads(grid: P.WeightedGrid, dslot: np.ndarray, G: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Chain rule from dloss/dA on the pattern to (dloss/dtheta, dloss/dw)."""
    gd = G[dslot]
    e = grid.edges
    return gd, gd[e[:, 0]] + gd[e[:, 1]] - G[grid.edge_slot[:, 0]] - G[grid.edge_slot[:, 1]]


class GridSpectralLoss:
    """loss(theta, w) = sum_k c_k lambda_k(L_w + diag(theta)) on a fixed graph.

    The graph topology (edges, CSR pattern, slot maps) is fixed at
    construction; theta and w are the differentiable inputs (float64 torch
    tensors on the CPU).  Every forward runs one FP64 GPU solve from the same
    seeded start block."""

    def __init__(self, ctx: L.Context, grid: P.WeightedGrid, c: np.ndarray, b: int = 8, ncv: int = 128, tol: float = 1e-10, max_restarts: int = 5000,
                 seed: int = 42):
        self.ctx = ctx
        self.g = grid
        self.n = grid.nx * grid.ny
        self.c = np.asarray(c, dtype=np.float64)
        self.nev = self.c.size
        self.b = b
        self.params = EigensolverParams(self.nev, ncv, max_restarts=max_restarts, tolerance=tol)
        self.dslot = diag_slots(grid.csr)
        self.dA = DeviceCSR(ctx, grid.csr, precisions=("f64",))
        self.V0 = ctx.to_device(P.initial_block(self.n, b, seed))
        self.last = None  # (evals, X, resid, result) of the latest forward

    def eigs(self, theta: np.ndarray, w: np.ndarray):
        self.dA.val["f64"].upload(grid_values(self.g, self.dslot, theta, w))
        out = solve(self.dA, self.b, self.params, self.V0, dtype="f64", order="ascending", reorth="cgs2")
        if not out.result["converged"]:
            raise RuntimeError(f"forward eigensolve did not converge: {out.result}")
        return out

    def sddmm(self, X: L.DeviceArray, weights: np.ndarray) -> np.ndarray:
        """G on A's pattern: G_ij = sum_k weights_k x_ik x_jk (device), downloaded."""
        nblk = X.shape[0]
        wk = np.zeros(nblk * self.b)
        wk[: weights.size] = weights
        wd = self.ctx.to_device(wk)
        G = self.ctx.empty((self.g.csr.nnz,), np.float64)
        L.check(L.hip().trl_hip_sddmm_adjoint(self.ctx.h, L.F64, self.n, self.dA.rowptr.ptr, self.dA.col.ptr, self.b, nblk, X.ptr, wd.ptr, None, G.ptr,
                                               None), "trl_hip_sddmm_adjoint")
        return G.download()

    def __call__(self, theta: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
        return _SpectralLossFn.apply(theta, w, self)


class _SpectralLossFn(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta, w, mod: GridSpectralLoss):
        out = mod.eigs(theta.detach().cpu().numpy().astype(np.float64), w.detach().cpu().numpy().astype(np.float64))
        mod.last = out
        fctx.mod = mod
        fctx.X = out.X
        return torch.tensor(float(np.dot(mod.c, out.evals)), dtype=torch.float64, device=theta.device)

    @staticmethod
    def backward(fctx, grad_out):
        mod = fctx.mod
        G = mod.sddmm(fctx.X, float(grad_out) * mod.c)
        gt, gw = param_grads(mod.g, mod.dslot, G)
        dev = grad_out.device
        return torch.from_numpy(gt).to(dev), torch.from_numpy(gw).to(dev), None


def config_c5(side: int = 1000, nev: int = 16, seed: int = 42) -> tuple[P.WeightedGrid, np.ndarray]:
    """The weighted grid side x side (w ~ U[0.5, 2], theta ~ U[0, 1]) and the
    loss weights c ~ U[0, 1] (stream 6)."""
    return P.weighted_grid(side, side, seed), P.uniform(seed, 6, nev)


def fd_coordinates(grid: P.WeightedGrid, count: int = 100, seed: int = 42) -> list[tuple[str, int]]:
    """count coordinates (stream 7), split evenly between theta and edge weights."""
    n, ne = grid.nx * grid.ny, grid.edges.shape[0]
    u = P.uniform(seed, 7, count)
    out = []
    for i, x in enumerate(u):
        out.append(("theta", int(x * n)) if i % 2 == 0 else ("w", int(x * ne)))
    return out


def top_coordinates(gt: np.ndarray, gw: np.ndarray, k: int) -> list[tuple[str, int]]:
    """The k theta and the k edge-weight coordinates of largest |gradient|."""
    return [("theta", int(i)) for i in np.argsort(-np.abs(gt))[:k]] + [("w", int(i)) for i in np.argsort(-np.abs(gw))[:k]]


def central_fd(mod: GridSpectralLoss, theta: np.ndarray, w: np.ndarray, kind: str, i: int, h: float) -> float:
    """(loss(p + h e_i) - loss(p - h e_i)) / 2h from two fresh forward solves."""
    th, ww = theta.copy(), w.copy()
    arr = th if kind == "theta" else ww
    arr[i] += h
    lp = float(mod.c @ mod.eigs(th, ww).evals)
    arr[i] -= 2 * h
    lm = float(mod.c @ mod.eigs(th, ww).evals)
    return (lp - lm) / (2 * h)


def main(argv=None) -> int:
    """Central-FD check of the adjoint on the C5 workload: prints one line per
    coordinate (adjoint, FD, absolute and relative difference) and a summary.
    --top K checks the K theta and K edge-weight coordinates of largest
    |gradient| (relative error bar 1e-4); otherwise --count coordinates of
    stream 7 (absolute bar 1e-6: at most of them the gradient is below the
    FD noise floor)."""
    import argparse
    import time

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--side", type=int, default=1000)
    ap.add_argument("--count", type=int, default=100)
    ap.add_argument("--top", type=int, default=0)
    ap.add_argument("--h", type=float, default=1e-6)
    ap.add_argument("--ncv", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-10)
    a = ap.parse_args(argv)
    g, c = config_c5(a.side)
    ctx = L.Context(0)
    mod = GridSpectralLoss(ctx, g, c, b=8, ncv=a.ncv, tol=a.tol)
    th = torch.tensor(g.theta, requires_grad=True)
    w = torch.tensor(g.weights, requires_grad=True)
    t0 = time.time()
    loss = mod(th, w)
    t1 = time.time()
    loss.backward()
    t2 = time.time()
    gt, gw = th.grad.numpy(), w.grad.numpy()
    print(f"C5 {a.side}x{a.side}: n={g.csr.n} nnz={g.csr.nnz} loss={loss.item():.15e} forward {t1 - t0:.2f}s backward {t2 - t1:.3f}s "
          f"cycles {mod.last.result['iterations']} max resid {mod.last.resid.max():.2e}", flush=True)
    coords = top_coordinates(gt, gw, a.top) if a.top else fd_coordinates(g, a.count)
    worst, worst_rel = 0.0, 0.0
    for k, (kind, i) in enumerate(coords):
        fd = central_fd(mod, g.theta, g.weights, kind, i, a.h)
        adj = float(gt[i] if kind == "theta" else gw[i])
        d = abs(fd - adj)
        r = d / abs(adj) if adj != 0.0 else float("inf")
        worst = max(worst, d)
        worst_rel = max(worst_rel, r)
        print(f"{k:3d} {kind:5s}[{i:7d}] adjoint {adj: .12e} fd {fd: .12e} |diff| {d:.2e} rel {r:.2e}", flush=True)
    if a.top:
        print(f"max relative |adjoint - fd| / |adjoint| = {worst_rel:.3e} over the {a.top} largest-|gradient| theta and w coordinates "
              f"(max |diff| {worst:.3e})", flush=True)
        return 0 if worst_rel <= 1e-4 else 1
    print(f"max |adjoint - fd| = {worst:.3e} over {a.count} coordinates", flush=True)
    return 0 if worst <= 1e-6 else 1


if __name__ == "__main__":
    raise SystemExit(main())
