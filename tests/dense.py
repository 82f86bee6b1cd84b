"""Independent dense-matrix restatement of the group conventions (numpy complex
matrices), used to check the oracle's parameterised arithmetic.

SU(2): U = e0 1 + i e_k sigma_k, algebra X = v_a sigma_a / 2.
SU(3): U = 3x3 complex (18 reals: real parts row-major, then imaginary parts),
       algebra X = c_a lambda_a / 2 with the standard Gell-Mann matrices.
"""
import numpy as np

SIGMA = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]], complex),
         np.array([[1, 0], [0, -1]], complex)]


def gell_mann():
    lam = np.zeros((8, 3, 3), complex)
    lam[0][0, 1] = lam[0][1, 0] = 1
    lam[1][0, 1], lam[1][1, 0] = -1j, 1j
    lam[2][0, 0], lam[2][1, 1] = 1, -1
    lam[3][0, 2] = lam[3][2, 0] = 1
    lam[4][0, 2], lam[4][2, 0] = -1j, 1j
    lam[5][1, 2] = lam[5][2, 1] = 1
    lam[6][1, 2], lam[6][2, 1] = -1j, 1j
    lam[7] = np.diag([1, 1, -2]) / np.sqrt(3)
    return lam


LAMBDA = gell_mann()


def gens(nc):
    return [s / 2 for s in SIGMA] if nc == 2 else [l / 2 for l in LAMBDA]


def group_to_matrix(nc, u):
    u = np.asarray(u, float)
    if nc == 2:
        return u[0] * np.eye(2) + 1j * (u[1] * SIGMA[0] + u[2] * SIGMA[1] + u[3] * SIGMA[2])
    return u[:9].reshape(3, 3) + 1j * u[9:].reshape(3, 3)


def matrix_to_group(nc, m):
    if nc == 2:
        # m = e0 + i e.sigma  ->  e0 = Re tr m / 2, e_k = Im tr(sigma_k m) / 2
        return np.array([np.real(np.trace(m)) / 2] + [np.imag(np.trace(s @ m)) / 2 for s in SIGMA])
    return np.concatenate([m.real.reshape(-1), m.imag.reshape(-1)])


def coeffs_to_herm(nc, c):
    return sum(c[a] * g for a, g in enumerate(gens(nc)))


def herm_to_coeffs(nc, h):
    """c_a = 2 tr(t_a H) for the traceless part of H."""
    h = h - np.trace(h) / nc * np.eye(nc)
    return np.array([2 * np.real(np.trace(g @ h)) for g in gens(nc)])


def proj_matrix(nc, m):
    """Coefficients of the traceless part of (M - M^dagger)/(2i)."""
    return herm_to_coeffs(nc, (m - m.conj().T) / 2j)


def power_series_exp(m, terms=100):
    out = np.eye(m.shape[0], dtype=complex)
    t = np.eye(m.shape[0], dtype=complex)
    for k in range(1, terms + 1):
        t = t @ m / k
        out = out + t
    return out
