"""Pins the oracle's SU(2) adjoint (opx_oracle_adjoint.c).

The reference has no derivative code, so every transposed Jacobian is checked
against central finite differences of the oracle's forward functions, which are
themselves pinned against the reference's formulas elsewhere:

* exp-map VJP vs finite differences of SU2AlgebraElement.getLink, on both sides
  of the small-|w| series branch;
* force VJP (transposed staple/plaquette Jacobian) vs finite differences of the
  E update of the fused step (FastTYMSolver.java:40-62), on anisotropic lattices,
  for arbitrary (not only unitary) link perturbations;
* magnetic loss term and its gradient vs finite differences;
* the loss value vs the lattice average of ProjectedEnergyDensity's
  e_T - e_L (ProjectedEnergyDensity.java:132-160);
* the whole reverse sweep vs finite differences of the N-step loss.
"""
import numpy as np
import pytest

import oracle as O

from .helpers import make_state

SPACINGS = [(1.0, 1.0, 1.0), (1.0, 0.8, 1.3)]


def test_exp_vjp_matches_finite_differences():
    rng = np.random.default_rng(0)
    for scale in (1e-4, 0.05, 0.099, 0.101, 0.7, 2.5):
        for _ in range(5):
            w = rng.normal(size=3)
            w *= scale / np.linalg.norm(w)
            Mbar = rng.normal(size=4)
            d = rng.normal(size=3)
            h = 1e-6
            fd = (O.alg_get_link(2, w + h * d) - O.alg_get_link(2, w - h * d)) / (2 * h)
            assert abs(Mbar @ fd - O.adj_exp_vjp(w, Mbar) @ d) < 1e-8 * (1 + abs(Mbar @ fd)), (scale, w)


def force(og, U):
    """F(U) = proj(U S(U)) per link through the oracle's fused step (E = 0, J = 0)."""
    og.U[:] = U
    og.E[:] = 0.0
    og.step(1.0)
    return og.E.copy()


@pytest.mark.parametrize("spacing", SPACINGS)
def test_force_vjp_matches_finite_differences(spacing):
    dims = (4, 3, 5)
    rng = np.random.default_rng(1)
    U, _ = make_state(2, dims, 0.5, 0.0, seed=3)
    og = O.OracleGrid(2, dims, spacing, 0.25, 1.0)
    n = og.n
    Ebar = rng.normal(size=(n, 3, 3))
    vjp = og.adj_force_vjp(U, Ebar)
    for _ in range(4):
        dU = rng.normal(size=(n, 3, 4))   # any direction in R^4 per link, not only tangent ones
        h = 1e-5
        fd = (force(og, U + h * dU) - force(og, U - h * dU)) / (2 * h)
        lhs, rhs = np.sum(Ebar * fd), np.sum(vjp * dU)
        assert abs(lhs - rhs) < 1e-8 * abs(lhs), (lhs, rhs)


def test_force_vjp_single_link_perturbations():
    """Each of the 12 neighbouring links of one link, and the link itself, perturbed alone."""
    dims = (3, 4, 3)
    rng = np.random.default_rng(2)
    U, _ = make_state(2, dims, 0.4, 0.0, seed=4)
    og = O.OracleGrid(2, dims, (1.0, 1.0, 1.0), 0.25, 1.0)
    Ebar = rng.normal(size=(og.n, 3, 3))
    vjp = og.adj_force_vjp(U, Ebar)
    for cell in (0, 7, 35):
        for d in range(3):
            for comp in range(4):
                dU = np.zeros_like(U)
                dU[cell, d, comp] = 1.0
                h = 1e-5
                fd = (force(og, U + h * dU) - force(og, U - h * dU)) / (2 * h)
                assert abs(np.sum(Ebar * fd) - vjp[cell, d, comp]) < 1e-8 * (1 + abs(vjp[cell, d, comp]))


@pytest.mark.parametrize("spacing", SPACINGS)
@pytest.mark.parametrize("direction", [0, 2])
def test_magnetic_loss_gradient(spacing, direction):
    dims = (4, 5, 3)
    rng = np.random.default_rng(3)
    U, _ = make_state(2, dims, 0.6, 0.0, seed=5)
    og = O.OracleGrid(2, dims, spacing, 0.25, 1.3)
    v, grad = og.adj_bsq(U, direction)
    assert abs(v) > 1e-4
    for _ in range(3):
        dU = rng.normal(size=U.shape)
        h = 1e-6   # |B|^2 is of degree 8 in the link components: keep the h^2 term below 1e-9
        fd = (og.adj_bsq(U + h * dU, direction, grad=False) - og.adj_bsq(U - h * dU, direction, grad=False)) / (2 * h)
        assert abs(fd - np.sum(grad * dU)) < 1e-8 * abs(fd)


@pytest.mark.parametrize("spacing", SPACINGS)
@pytest.mark.parametrize("direction", [0, 1])
def test_loss_is_projected_energy_difference(spacing, direction):
    """L = (1/N_dir) sum_k [e_T,el + e_T,mag - e_L,el - e_L,mag](k) with the reference's measurement."""
    dims = (5, 4, 3)
    U, E = make_state(2, dims, 0.4, 0.3, seed=6)
    og = O.OracleGrid(2, dims, spacing, 0.25, 1.3)
    og.U[:] = U
    og.E[:] = E
    og.step_links()
    for _ in range(3):
        og.store_fields()
        og.step()
    tel, tmag, lel, lmag = og.projected_energy_density(direction)
    expected = np.mean(tel + tmag - lel - lmag)
    assert abs(og.adj_loss(direction) - expected) < 1e-13 * abs(expected) + 1e-16


def loss_after(dims, spacing, g, U, E0, steps, direction):
    og = O.OracleGrid(2, dims, spacing, 0.25, g)
    og.U[:] = U
    og.E[:] = E0
    og.step_links()
    for _ in range(steps):
        og.store_fields()
        og.step()
    return og.adj_loss(direction)


@pytest.mark.parametrize("spacing,g,steps", [((1.0, 1.0, 1.0), 1.0, 1), ((1.0, 1.0, 1.0), 1.0, 6),
                                             ((1.0, 0.8, 1.3), 1.3, 5)])
def test_adjoint_gradient_matches_finite_differences(spacing, g, steps):
    dims = (4, 3, 5)
    rng = np.random.default_rng(4)
    U, E0 = make_state(2, dims, 0.4, 0.2, seed=7)
    og = O.OracleGrid(2, dims, spacing, 0.25, g)
    og.U[:] = U
    og.E[:] = E0
    loss, grad = og.adjoint_gradient(steps, 0)
    assert abs(loss - loss_after(dims, spacing, g, U, E0, steps, 0)) < 1e-14 * abs(loss)
    for _ in range(6):
        v = rng.normal(size=E0.shape)
        h = 1e-5
        fd = (loss_after(dims, spacing, g, U, E0 + h * v, steps, 0) -
              loss_after(dims, spacing, g, U, E0 - h * v, steps, 0)) / (2 * h)
        assert abs(fd - np.sum(grad * v)) < 1e-7 * abs(fd), (fd, np.sum(grad * v))


def test_adjoint_gradient_is_gauge_covariant():
    """A gauge transformation of the initial state rotates dL/dE0 cell by cell (L is gauge invariant)."""
    dims = (3, 4, 3)
    U, E0 = make_state(2, dims, 0.4, 0.2, seed=8)
    og = O.OracleGrid(2, dims, (1.0, 1.0, 1.0), 0.25, 1.0)
    og.U[:] = U
    og.E[:] = E0
    _, g0 = og.adjoint_gradient(4, 2)
    rng = np.random.default_rng(5)
    gcell = O.links_from_coeffs(2, rng.normal(0, 1.0, size=(og.n, 3)))
    og2 = O.OracleGrid(2, dims, (1.0, 1.0, 1.0), 0.25, 1.0)
    og2.U[:] = U
    og2.Unext[:] = U
    og2.E[:] = E0
    og2.gauge_transform(gcell)
    _, g1 = og2.adjoint_gradient(4, 2)
    for c in range(og.n):
        for d in range(3):
            assert np.allclose(g1[c, d], O.alg_act(2, g0[c, d], gcell[c]), rtol=1e-9, atol=1e-14)
