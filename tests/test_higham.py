"""Higham's HPL-AI matrix: the coefficient computation (host code of the
library, no GPU) pinned by the definition of the condition number.

A = L U, L unit lower with a below the diagonal, U unit upper with b above.
With Fasi and Higham's signs (a, b) = (-alpha, -beta) the published closed
form gives cond(A, inf) = kappa exactly; numpy's cond(A, inf) of the dense
FP64 matrix must reproduce kappa to 1e-8.  The reference's solver builds the
same L U with (+alpha, +beta) (pmatgen(HMGen), matgen.hpp:36-73), a much
better conditioned matrix; that is checked too, so that a sign mistake in
either direction fails.
"""
import numpy as np
import pytest


def dense(n, a, b):
    L = np.eye(n) + a * np.tril(np.ones((n, n)), -1)
    U = np.eye(n) + b * np.triu(np.ones((n, n)), 1)
    return L @ U


@pytest.mark.parametrize("n", [20, 50, 100, 400, 1000])
@pytest.mark.parametrize("kappa", [100.0, 1000.0])
def test_coefficients_give_kappa(n, kappa):
    import hplai
    alpha, beta = hplai.higham_coeffs(n, kappa, 0.5)
    assert alpha > 0 and beta > 0 and abs(alpha - 0.5 * beta) <= 1e-15 * beta
    c = np.linalg.cond(dense(n, -alpha, -beta), np.inf)
    assert abs(c - kappa) / kappa < 1e-8, (n, kappa, c)
    # the reference's signs: the same factors with positive off-diagonals
    c_ref = np.linalg.cond(dense(n, alpha, beta), np.inf)
    assert c_ref < 0.1 * kappa, c_ref


def test_closed_form_entries():
    """a_ij = b + a b i (i < j), 1 + a b i (i = j), a + a b j (i > j)."""
    n, a, b = 37, 0.03, -0.07
    A = dense(n, a, b)
    i, j = np.indices((n, n))
    want = np.where(i < j, b + a * b * i, np.where(i == j, 1 + a * b * i, a + a * b * j))
    assert np.allclose(A, want, rtol=0, atol=1e-15)


def test_rejects_bad_arguments():
    import hplai
    with pytest.raises(hplai.HplaiError):
        hplai.higham_coeffs(1, 1000.0, 0.5)
    with pytest.raises(hplai.HplaiError):
        hplai.higham_coeffs(100, 0.5, 0.5)
