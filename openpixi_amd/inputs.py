"""Seeded synthetic inputs for the leapfrog workloads.

Every random number comes from numpy's counter-based Philox-4x32 generator
keyed by (seed, stream), so the same configuration is reproduced exactly on any
host. "Gaussian(0, s)" means mean 0 and standard deviation s.
"""
from __future__ import annotations

import numpy as np

STREAM_LINKS = 1
STREAM_EFIELD = 2
STREAM_CHARGES = 3
STREAM_SIZES = 4


def rng(seed: int, stream: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=[int(seed) & 0xFFFFFFFFFFFFFFFF, int(stream)]))


def gaussian_coefficients(num_cells: int, num_colors: int, sigma: float, seed: int, stream: int) -> np.ndarray:
    """(cells, 3, nc^2-1) float64 Gaussian(0, sigma) generator coefficients."""
    d = num_colors * num_colors - 1
    return rng(seed, stream).normal(0.0, sigma, size=(num_cells, 3, d))


def random_links_coefficients(num_cells, num_colors, sigma=0.1, seed=0):
    """Coefficients A with U = exp(i a A_b t_b) (a = 1)."""
    return gaussian_coefficients(num_cells, num_colors, sigma, seed, STREAM_LINKS)


def random_efield_coefficients(num_cells, num_colors, sigma=0.05, seed=0):
    return gaussian_coefficients(num_cells, num_colors, sigma, seed, STREAM_EFIELD)


def su2_coeffs_to_storage(c: np.ndarray) -> np.ndarray:
    """SU(2) algebra storage is the coefficient vector itself."""
    return np.ascontiguousarray(c, np.float64)


def su3_coeffs_to_storage(c: np.ndarray) -> np.ndarray:
    """Gell-Mann coefficients (..., 8) -> 9-real Hermitian storage of
    math/SU3AlgebraElement.java:3-11 (H = sum_a c_a lambda_a / 2)."""
    c = np.asarray(c, np.float64)
    v = np.zeros(c.shape[:-1] + (9,))
    s3 = np.sqrt(3.0)
    v[..., 1] = c[..., 0] / 2
    v[..., 3] = c[..., 1] / 2
    v[..., 0] = c[..., 2] / 2 + c[..., 7] / (2 * s3)
    v[..., 4] = -c[..., 2] / 2 + c[..., 7] / (2 * s3)
    v[..., 8] = -c[..., 7] / s3
    v[..., 2] = c[..., 3] / 2
    v[..., 6] = c[..., 4] / 2
    v[..., 5] = c[..., 5] / 2
    v[..., 7] = c[..., 6] / 2
    return v


def coeffs_to_storage(nc: int, c: np.ndarray) -> np.ndarray:
    return su2_coeffs_to_storage(c) if nc == 2 else su3_coeffs_to_storage(c)
