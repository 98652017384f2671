"""Eigenvalue adjoint (SDDMM on A's pattern) through torch.autograd.

loss = sum_k c_k lambda_k(L_w + diag(theta)) over the 16 smallest
eigenpairs; the backward is trl_hip_sddmm_adjoint on the device plus the
chain rule to (theta, w).  Checked against the FP64 closed form
dloss/dtheta_i = sum_k c_k v_ik^2, dloss/dw_e = sum_k c_k (v_ik - v_jk)^2
evaluated on the host from the same eigenvectors, and against central finite
differences (h = 1e-6) of fresh forward solves.  Bars: on the 10 theta and
10 edge-weight coordinates of largest |gradient| at full size,
|adjoint - FD| / |adjoint| <= 1e-4; on 100 random coordinates of a 64 x 64
instance (most of whose gradients sit below the FD noise floor)
|adjoint - FD| <= 1e-6.
"""

import numpy as np
import pytest
import torch

from trl_amd.adjoint import GridSpectralLoss, config_c5, fd_coordinates

pytestmark = pytest.mark.gpu

H = 1e-6


def closed_form(mod):
    X = mod.last.X.download()  # (nblk, n, b)
    V = np.concatenate(list(X), axis=1)[:, : mod.nev]
    e = mod.g.edges
    return (V**2) @ mod.c, ((V[e[:, 0]] - V[e[:, 1]]) ** 2) @ mod.c


def central_fd(mod, theta, w, kind, i):
    th, ww = theta.copy(), w.copy()
    arr = th if kind == "theta" else ww
    arr[i] += H
    lp = float(mod.c @ mod.eigs(th, ww).evals)
    arr[i] -= 2 * H
    lm = float(mod.c @ mod.eigs(th, ww).evals)
    return (lp - lm) / (2 * H)


def run(mod, g):
    th = torch.tensor(g.theta, requires_grad=True)
    w = torch.tensor(g.weights, requires_grad=True)
    loss = mod(th, w)
    loss.backward()
    return loss, th.grad.numpy().copy(), w.grad.numpy().copy()


def test_adjoint_fd_100_coordinates(ctx):
    # the C5 builder on a 64 x 64 grid (n = 4096): 100 coordinates (stream 7),
    # half theta, half edge weights, each FD from two fresh solves
    print("adjoint fd 64x64", flush=True)
    g, c = config_c5(side=64)
    mod = GridSpectralLoss(ctx, g, c, b=8, ncv=128, tol=1e-12)
    loss, gt, gw = run(mod, g)
    assert abs(loss.item() - float(c @ mod.last.evals)) < 1e-15
    ct, cw = closed_form(mod)
    assert np.abs(gt - ct).max() <= 1e-12 * np.abs(ct).max()
    assert np.abs(gw - cw).max() <= 1e-12 * np.abs(cw).max()
    err = []
    for kind, i in fd_coordinates(g, 100):
        fd = central_fd(mod, g.theta, g.weights, kind, i)
        adj = gt[i] if kind == "theta" else gw[i]
        err.append(abs(fd - adj))
    err = np.array(err)
    print(f"max |adj - fd| {err.max():.3e}  (max |grad| {max(np.abs(gt).max(), np.abs(gw).max()):.3e})")
    assert err.max() <= 1e-6


@pytest.fixture(scope="module")
def c5(ctx):
    # C5 at full size: 1000 x 1000 grid (n = 1M), 16 smallest, FP64
    print("adjoint c5 1000x1000", flush=True)
    g, c = config_c5()
    mod = GridSpectralLoss(ctx, g, c, b=8, ncv=256, tol=1e-10)
    _, gt, gw = run(mod, g)
    return mod, g, gt, gw


def test_adjoint_c5_full_size(c5):
    mod, g, gt, gw = c5
    assert mod.last.resid.max() <= 1e-10
    ct, cw = closed_form(mod)
    rt = np.abs(gt - ct).max() / np.abs(ct).max()
    rw = np.abs(gw - cw).max() / np.abs(cw).max()
    print(f"sddmm vs closed form: theta {rt:.2e}, w {rw:.2e}")
    assert rt <= 1e-12 and rw <= 1e-12


# the 10 theta and the 10 edge-weight coordinates of largest |gradient|, in
# chunks of 5 (each FD value is two fresh full-size solves)
@pytest.mark.parametrize("kind,chunk", [("theta", 0), ("theta", 1), ("w", 0), ("w", 1)])
def test_adjoint_c5_top_gradients_relative(c5, kind, chunk):
    mod, g, gt, gw = c5
    grad = gt if kind == "theta" else gw
    idx = np.argsort(-np.abs(grad))[5 * chunk : 5 * chunk + 5]
    worst = 0.0
    for i in idx:
        fd = central_fd(mod, g.theta, g.weights, kind, int(i))
        r = abs(fd - grad[i]) / abs(grad[i])
        print(f"{kind}[{i}]: adjoint {grad[i]:.12e} fd {fd:.12e} rel {r:.2e}", flush=True)
        worst = max(worst, r)
    assert worst <= 1e-4
