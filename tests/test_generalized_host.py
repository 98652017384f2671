"""Host checks of the generalized test pairs (no GPU): the closed form of the
bilinear finite-element Laplacian pair against LAPACK's dense generalized
eigensolver, and the symmetry and positive definiteness of the sparse mass
matrix built on the stiffness pattern."""

import numpy as np
import pytest
import scipy.linalg as sla

from trl_amd import problems as P


@pytest.mark.parametrize("nx,ny", [(3, 3), (9, 6), (17, 12)])
def test_fe_pair_closed_form_matches_lapack(nx, ny):
    K, M = P.fe_laplacian_2d(nx, ny)
    ev = sla.eigh(K.to_dense(), M.to_dense(), eigvals_only=True)
    cf = P.fe_laplacian_2d_eigenvalues(nx, ny)
    assert np.all(np.abs(ev - cf) <= 1e-12 * np.abs(cf))


def test_fe_pair_structure():
    K, M = P.fe_laplacian_2d(7, 5)
    Kd, Md = K.to_dense(), M.to_dense()
    assert np.array_equal(K.rowptr, M.rowptr) and np.array_equal(K.col, M.col)
    assert np.allclose(Kd, Kd.T, rtol=0, atol=0) and np.allclose(Md, Md.T, rtol=0, atol=0)
    # the 1D mass row h/6 (1, 4, 1) summed over the 9-point stencil: interior rows of M sum to h^2 (the element areas)
    h = (1 / 8) * (1 / 6)
    interior = 2 * 7 + 2  # node (2, 2), away from the boundary
    assert abs(Md[interior].sum() - h) < 1e-15
    with pytest.raises(ValueError):
        P.fe_laplacian_2d(2, 5)


def test_sparse_mass_pair_is_spd():
    K, B = P.stiffness_sparse_mass(5, 4, 3, dof=3)
    Bd = B.to_dense()
    assert np.array_equal(Bd, Bd.T)
    assert np.linalg.eigvalsh(Bd).min() > 0
    # strictly diagonally dominant rows, off-diagonal entries only where K has them
    off = Bd - np.diag(np.diag(Bd))
    assert np.all(np.diag(Bd) > np.abs(off).sum(axis=1))
    assert np.all((K.to_dense() != 0) | (Bd == 0))
    assert np.count_nonzero(off) > 0
