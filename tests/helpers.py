"""Shared set-up for parity tests: identical seeded inputs for the oracle and the device."""
from __future__ import annotations

import numpy as np

import oracle as O
from openpixi_amd import inputs


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def make_state(nc, dims, sigma_a=0.1, sigma_e=0.05, seed=0):
    """(U, E) in reference layout: U = getLink(A) computed by the oracle."""
    n = int(np.prod(dims))
    A = inputs.random_links_coefficients(n, nc, sigma_a, seed)
    U = O.links_from_coeffs(nc, A)
    if sigma_e > 0:
        Ec = inputs.random_efield_coefficients(n, nc, sigma_e, seed)
        E = O.algs_from_coeffs(nc, Ec)
    else:
        E = np.zeros((n, 3, O.asize(nc)))
    return U, E


def oracle_grid(nc, dims, U, E, at=0.25, spacing=(1.0, 1.0, 1.0), g=1.0, J=None):
    og = O.OracleGrid(nc, dims, spacing, at, g)
    og.U[:] = U
    og.E[:] = E
    if J is not None:
        og.J[:] = J
    return og


def device_grid(nc, dims, U, E, at=0.25, spacing=(1.0, 1.0, 1.0), g=1.0, precision="fp64", J=None):
    from openpixi_amd import Grid

    dg = Grid(nc, dims, list(spacing), at, g, precision)
    dg.setU(U)
    dg.setE(E)
    if J is not None:
        dg.setJ(J)
    return dg
