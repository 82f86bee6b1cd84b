"""Pins the oracle's SU(2)/SU(3) element arithmetic against dense complex
matrices (an independent representation), restating the reference's JUnit
properties: SU2MatrixTest (mult, adj, act = M X M^+, Pauli basis), SU2FieldTest
(square, getLink/proj round trips), SU3EverythingTest (generator convention,
det = 1, U U^+ = 1, exp vs a 100-term power series, act, hermiticity)."""
import numpy as np
import pytest
from scipy.linalg import expm

import oracle as O

from .dense import (coeffs_to_herm, gens, group_to_matrix, herm_to_coeffs, matrix_to_group, power_series_exp,
                    proj_matrix)

RNG = np.random.default_rng(1234)


def random_link(nc, scale=1.0):
    return O.alg_get_link(nc, O.alg_from_coeffs(nc, RNG.normal(0, scale, nc * nc - 1)))


def random_matrix_group_storage(nc):
    """A generic (non-unitary) element in group storage, as the staple sums are."""
    return RNG.normal(size=O.gsize(nc))


@pytest.mark.parametrize("nc", [2, 3])
def test_mult_adj_match_dense(nc):
    for _ in range(20):
        a, b = random_matrix_group_storage(nc), random_matrix_group_storage(nc)
        if nc == 2:  # SU(2) storage only spans e0 + i e.sigma
            pass
        A, B = group_to_matrix(nc, a), group_to_matrix(nc, b)
        assert np.allclose(group_to_matrix(nc, O.group_mult(nc, a, b)), A @ B, atol=1e-13)
        assert np.allclose(group_to_matrix(nc, O.group_adj(nc, a)), A.conj().T, atol=1e-15)


@pytest.mark.parametrize("nc", [2, 3])
def test_proj_is_traceless_antihermitian_part(nc):
    for _ in range(20):
        m = random_matrix_group_storage(nc)
        got = O.alg_coeffs(nc, O.group_proj(nc, m))
        assert np.allclose(got, proj_matrix(nc, group_to_matrix(nc, m)), atol=1e-13)


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("scale", [0.05, 0.7, 2.5])
def test_get_link_is_exp_i_x(nc, scale):
    """getLink(X) = exp(+i X); scale 2.5 takes the SU(3) spectral branch (tr X^2 > 1)."""
    for _ in range(10):
        c = RNG.normal(0, scale, nc * nc - 1)
        U = group_to_matrix(nc, O.alg_get_link(nc, O.alg_from_coeffs(nc, c)))
        X = coeffs_to_herm(nc, c)
        assert np.allclose(U, expm(1j * X), atol=1e-13)
        assert np.allclose(U, power_series_exp(1j * X), atol=1e-12)
        assert np.allclose(U @ U.conj().T, np.eye(nc), atol=1e-13)
        assert abs(np.linalg.det(U) - 1) < 1e-13


def test_su3_degenerate_spectrum_branch():
    """tr X^2 > 1 with two equal eigenvalues (lambda_8 direction) takes the Taylor-25 fallback."""
    c = np.zeros(8)
    c[7] = 2.0
    U = group_to_matrix(3, O.alg_get_link(3, O.alg_from_coeffs(3, c)))
    assert np.allclose(U, expm(1j * coeffs_to_herm(3, c)), atol=1e-7)


@pytest.mark.parametrize("nc", [2, 3])
def test_act_is_conjugation(nc):
    for _ in range(20):
        c = RNG.normal(0, 1, nc * nc - 1)
        g = random_link(nc, 1.0)
        G = group_to_matrix(nc, g)
        got = O.alg_coeffs(nc, O.alg_act(nc, O.alg_from_coeffs(nc, c), g))
        want = herm_to_coeffs(nc, G @ coeffs_to_herm(nc, c) @ G.conj().T)
        assert np.allclose(got, want, atol=1e-13)


@pytest.mark.parametrize("nc", [2, 3])
def test_square_is_sum_of_coefficients_squared(nc):
    for _ in range(10):
        c = RNG.normal(0, 1, nc * nc - 1)
        v = O.alg_from_coeffs(nc, c)
        assert abs(O.alg_square(nc, v) - np.sum(c * c)) < 1e-13
        # = 2 tr X^2
        X = coeffs_to_herm(nc, c)
        assert abs(O.alg_square(nc, v) - 2 * np.real(np.trace(X @ X))) < 1e-12


@pytest.mark.parametrize("nc", [2, 3])
def test_proj_of_get_link_inverts_for_small_fields(nc):
    """SU2FieldTest: getLink().proj() returns the input to first order; exactly
    proj(exp(iX)) = coefficients of sin-part, so compare to the dense value."""
    for _ in range(10):
        c = RNG.normal(0, 1e-3, nc * nc - 1)
        u = O.alg_get_link(nc, O.alg_from_coeffs(nc, c))
        got = O.alg_coeffs(nc, O.group_proj(nc, u))
        assert np.allclose(got, c, rtol=0, atol=1e-8)
        assert np.allclose(got, proj_matrix(nc, group_to_matrix(nc, u)), atol=1e-15)


def test_su3_storage_convention():
    """SU3EverythingTest:40-42: the 9-real storage is the Hermitian matrix of the coefficients."""
    for _ in range(10):
        c = RNG.normal(0, 1, 8)
        v = O.alg_from_coeffs(3, c)
        H = np.array([[v[0], v[1] - 1j * v[3], v[2] - 1j * v[6]],
                      [v[1] + 1j * v[3], v[4], v[5] - 1j * v[7]],
                      [v[2] + 1j * v[6], v[5] + 1j * v[7], v[8]]])
        assert np.allclose(H, coeffs_to_herm(3, c), atol=1e-15)
        assert np.allclose(O.alg_coeffs(3, v), c, atol=1e-14)
        assert abs(np.trace(H)) < 1e-15


def test_su2_pauli_basis():
    """SU2MatrixTest:143-228: U = e0 1 + i e_k sigma_k."""
    for _ in range(10):
        e = RNG.normal(size=4)
        M = group_to_matrix(2, e)
        assert np.allclose(matrix_to_group(2, M), e)
        assert np.allclose(M, e[0] * np.eye(2) + 1j * sum(e[k + 1] * g * 2 for k, g in enumerate(gens(2))))


def test_plausible_mistakes_are_caught():
    """A transposed product or a flipped exponent sign differs from the dense values
    by far more than the tolerances above."""
    a, b = random_link(3), random_link(3)
    A, B = group_to_matrix(3, a), group_to_matrix(3, b)
    assert not np.allclose(A @ B, B @ A, atol=1e-6)
    c = RNG.normal(0, 0.5, 8)
    X = coeffs_to_herm(3, c)
    assert not np.allclose(expm(1j * X), expm(-1j * X), atol=1e-6)
