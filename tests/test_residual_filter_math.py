"""The recurrence behind the residual-driven filter of the mixed solve's FP64
sweeps, checked in numpy on small dense problems (host only).

For the scaled Chebyshev filter Y_i = alpha_i A Y_{i-1} + beta_i Y_{i-1} +
gamma_i Y_{i-2} (Y_0 = x) and a Ritz pair (theta, x) with residual
r = A x - theta x, E_i = Y_i - p_i(theta) x obeys the same recurrence plus
the forcing term alpha_i p_{i-1}(theta) r, with E_0 = 0 and E_1 = alpha_1 r.
The solver steps E in FP32 and forms p_d(theta) x + E_d in FP64; these
tests check the identity, the coefficients the solver derives, and that
FP32 rounding of E stays at ~1e-7 of |E| (far below the error removed).
"""

import numpy as np
import pytest


def coefficients(lo, hi, a0, degree):
    # the scaled recurrence of the solver's filter (Zhou & Saad 2007, Alg. 3.2)
    e, c = 0.5 * (hi - lo), 0.5 * (hi + lo)
    s1 = e / (a0 - c)
    al, be, ga = np.zeros(degree + 1), np.zeros(degree + 1), np.zeros(degree + 1)
    al[1], be[1] = s1 / e, -s1 * c / e
    sg = s1
    for i in range(2, degree + 1):
        sn = 1.0 / (2.0 / s1 - sg)
        al[i], be[i], ga[i] = 2.0 * sn / e, -2.0 * sn * c / e, -sg * sn
        sg = sn
    return al, be, ga


def problem(rng, n=60):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.sort(np.concatenate([rng.uniform(0.0, 0.2, 6), rng.uniform(0.5, 4.0, n - 6)]))
    A = (Q * lam) @ Q.T
    return A, lam, Q


def plain_filter(A, X, al, be, ga, degree):
    Y0, Y1 = X.copy(), al[1] * (A @ X) + be[1] * X
    for i in range(2, degree + 1):
        Y0, Y1 = Y1, al[i] * (A @ Y1) + be[i] * Y1 + ga[i] * Y0
    return Y1


def forced_filter(A, X, theta, al, be, ga, degree, dtype=np.float64):
    # the solver's form: E_1 = alpha_1 R, then E_i with the forcing term
    # alpha_i p_{i-1}(theta) R; p_i(theta) by the scalar recurrence
    R = A @ X - X * theta
    E0, E1 = np.zeros_like(X, dtype=dtype), (al[1] * R).astype(dtype)
    Ad = A.astype(dtype)
    pm, pc = np.ones_like(theta), al[1] * theta + be[1]
    for i in range(2, degree + 1):
        g = al[i] * pc
        E0, E1 = E1, (al[i] * (Ad @ E1).astype(np.float64) + be[i] * E1 + ga[i] * E0 + g * R).astype(dtype)
        pm, pc = pc, (al[i] * theta + be[i]) * pc + ga[i] * pm
    return pc * X + E1.astype(np.float64), E1


@pytest.mark.parametrize("degree", [1, 2, 7, 40])
def test_forced_recurrence_equals_the_filter(rng, degree):
    A, lam, Q = problem(rng)
    # rough Ritz pairs of a 4-dimensional subspace near the lowest eigenvectors
    V, _ = np.linalg.qr(Q[:, :4] + 1e-3 * rng.standard_normal((A.shape[0], 4)))
    w, Z = np.linalg.eigh(V.T @ A @ V)
    X, theta = V @ Z, w
    al, be, ga = coefficients(lam[6], lam[-1] * 1.01, theta[0], degree)
    Y = plain_filter(A, X, al, be, ga, degree)
    got, _ = forced_filter(A, X, theta, al, be, ga, degree)
    assert np.abs(got - Y).max() <= 1e-12 * np.abs(Y).max()


def test_e1_is_alpha1_times_the_residual(rng):
    A, lam, Q = problem(rng)
    x = Q[:, :1] + 1e-2 * rng.standard_normal((A.shape[0], 1))
    x /= np.linalg.norm(x)
    theta = (x.T @ A @ x).ravel()
    al, be, ga = coefficients(lam[6], lam[-1], theta[0], 1)
    _, E1 = forced_filter(A, x, theta, al, be, ga, 1)
    assert np.allclose(E1, al[1] * (A @ x - x * theta), rtol=0, atol=1e-15)


def test_fp32_e_keeps_fp64_accuracy(rng):
    # the Ritz vectors are accurate to ~1e-6; stepping E in FP32 (and the
    # operator in FP32) leaves the filtered vectors within ~1e-7 of the size of
    # E of the FP64 result, while filtering x itself in FP32 errs by ~1e-7 of x
    A, lam, Q = problem(rng)
    V, _ = np.linalg.qr(Q[:, :4] + 1e-6 * rng.standard_normal((A.shape[0], 4)))
    w, Z = np.linalg.eigh(V.T @ A @ V)
    X, theta = V @ Z, w
    degree = 40
    al, be, ga = coefficients(lam[6], lam[-1] * 1.01, theta[0], degree)
    Y = plain_filter(A, X, al, be, ga, degree)
    got32, E = forced_filter(A, X, theta, al, be, ga, degree, dtype=np.float32)
    err_forced = np.abs(got32 - Y).max()
    Y32 = plain_filter(A.astype(np.float32), X.astype(np.float32), al, be, ga, degree).astype(np.float64)
    err_plain32 = np.abs(Y32 - Y).max()
    assert err_forced <= 1e-5 * np.abs(E).max() + 1e-15
    assert err_forced < 1e-3 * err_plain32
