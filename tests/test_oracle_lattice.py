"""Pins the oracle's lattice functions (staples, leapfrog step, Gauss law,
measurements) with checks independent of its own formulas:

* GridTest (T/physics/grid/GridTest.java:21-166): link orientation, plaquette
  inverse and trace, index <-> position, periodic shift.
* a dense-matrix brute force of the textbook staple update on small lattices;
* gauge covariance of the step (GaussConstraintTest uses RandomGauge the same way);
* exact Gauss-law conservation in pure Yang-Mills;
* the GaussConstraintTest closed form |G|^2 = |B|^2 (A.n)^2
  (T/physics/grid/GaussConstraintTest.java:20-128);
* exact solutions: the trivial vacuum, a uniform single-direction E field, and
  the discrete dispersion relation of an Abelian plane wave (the SU2PlaneWave
  input of Wave_Test.yaml), including anisotropic spacings.
"""
import itertools

import numpy as np
import pytest
from scipy.linalg import expm

import oracle as O

from .dense import coeffs_to_herm, group_to_matrix, herm_to_coeffs, matrix_to_group, proj_matrix
from .helpers import make_state, oracle_grid

RNG = np.random.default_rng(99)


def pos_to_index(dims, p):
    return ((p[0] % dims[0]) * dims[1] + (p[1] % dims[1])) * dims[2] + (p[2] % dims[2])


def index_to_pos(dims, i):
    return (i // (dims[1] * dims[2]), (i // dims[2]) % dims[1], i % dims[2])


def unit(d):
    e = [0, 0, 0]
    e[d] = 1
    return e


# ------------------------------------------------------------------ GridTest

def test_shift_and_index():
    dims = (16, 17, 18)
    og = O.OracleGrid(2, dims)
    for _ in range(300):
        i = int(RNG.integers(og.n))
        p = index_to_pos(dims, i)
        for d in range(3):
            for o in (-1, 1):
                q = list(p)
                q[d] += o
                assert og.shift(i, d, o) == pos_to_index(dims, q)


@pytest.mark.parametrize("nc", [2, 3])
def test_plaquette_orientation_and_trace(nc):
    dims = (4, 5, 3)
    U, E = make_state(nc, dims, 0.8, 0.0, seed=3)
    og = oracle_grid(nc, dims, U, E)
    og.Unext[:] = U
    for _ in range(40):
        i = int(RNG.integers(og.n))
        d1, d2 = RNG.choice(3, 2, replace=False)
        P12 = group_to_matrix(nc, og.plaquette(i, d1, d2, 1, 1, 0))
        P21 = group_to_matrix(nc, og.plaquette(i, d2, d1, 1, 1, 0))
        assert np.allclose(P12 @ P21, np.eye(nc), atol=1e-13)
        # backward-oriented plaquettes are closed loops too: equal trace to the
        # forward plaquette they are conjugate to (GridTest:60-82)
        j = og.shift(og.shift(i, d1, 1), d2, 1)
        Pb = group_to_matrix(nc, og.plaquette(j, d1, d2, -1, -1, 0))
        assert abs(np.trace(Pb) - np.trace(P12)) < 1e-13


# ------------------------------------------------------------------ brute force

def dense_step(nc, dims, U, E, J, at, spacing, g):
    """Textbook temporal-gauge leapfrog update with dense matrices (coordinates, not the oracle's shift)."""
    n = int(np.prod(dims))
    Um = {(i, d): group_to_matrix(nc, U[i, d]) for i in range(n) for d in range(3)}
    Eout = np.zeros((n, 3, nc * nc - 1))
    Un = np.zeros_like(U)
    for i in range(n):
        p = index_to_pos(dims, i)
        for d in range(3):
            S = np.zeros((nc, nc), complex)
            for j in range(3):
                if j == d:
                    continue
                pd = [p[k] + unit(d)[k] for k in range(3)]
                pj = [p[k] + unit(j)[k] for k in range(3)]
                mj = [p[k] - unit(j)[k] for k in range(3)]
                pdmj = [p[k] + unit(d)[k] - unit(j)[k] for k in range(3)]
                I = lambda q: pos_to_index(dims, q)  # noqa: E731
                fwd = Um[I(pd), j] @ Um[I(pj), d].conj().T @ Um[i, j].conj().T
                bwd = Um[I(pdmj), j].conj().T @ Um[I(mj), d].conj().T @ Um[I(mj), j]
                S += (fwd + bwd) / spacing[j] ** 2
            c = O.alg_coeffs(nc, E[i, d]) + at * proj_matrix(nc, Um[i, d] @ S)
            c = c - g * spacing[d] * at * O.alg_coeffs(nc, J[i, d])
            Eout[i, d] = c
            Un[i, d] = matrix_to_group(nc, expm(-1j * at * coeffs_to_herm(nc, c)) @ Um[i, d])
    return Eout, Un


@pytest.mark.parametrize("nc", [2, 3])
def test_step_matches_dense_brute_force(nc):
    dims = (3, 4, 5)
    spacing = (1.0, 0.7, 1.4)
    at, g = 0.2, 1.3
    U, E = make_state(nc, dims, 0.6, 0.3, seed=8)
    n = int(np.prod(dims))
    J = O.algs_from_coeffs(nc, RNG.normal(0, 0.2, (n, 3, nc * nc - 1)))
    og = oracle_grid(nc, dims, U, E, at, spacing, g, J=J)
    og.step(at)
    Ed, Ud = dense_step(nc, dims, U, E, J, at, spacing, g)
    Eo = np.array([[O.alg_coeffs(nc, og.E[i, d]) for d in range(3)] for i in range(n)])
    assert np.max(np.abs(Eo - Ed)) < 1e-13
    assert np.max(np.abs(og.Unext - Ud)) < 1e-13
    # a dropped backward staple or a sign flip on the current would be visible
    assert np.max(np.abs(Eo - np.array([[O.alg_coeffs(nc, E[i, d]) for d in range(3)] for i in range(n)]))) > 1e-3


# ------------------------------------------------------------------ gauge covariance

def random_gauge(nc, n, scale=1.0):
    return np.array([O.alg_get_link(nc, O.alg_from_coeffs(nc, RNG.normal(0, scale, nc * nc - 1)))
                     for _ in range(n)])


@pytest.mark.parametrize("nc", [2, 3])
def test_step_is_gauge_covariant(nc):
    """Transforming then stepping equals stepping then transforming (links exactly,
    E through the exact adjoint action)."""
    dims = (4, 3, 5)
    U, E = make_state(nc, dims, 0.5, 0.2, seed=12)
    n = int(np.prod(dims))
    gt = random_gauge(nc, n)
    a = oracle_grid(nc, dims, U, E)
    # transformed copy: U -> g(x) U g(x+i)^+, E -> g E g^+
    Ut = np.zeros_like(U)
    Et = np.zeros_like(E)
    for i in range(n):
        for d in range(3):
            j = a.shift(i, d, 1)
            Gi, Gj = group_to_matrix(nc, gt[i]), group_to_matrix(nc, gt[j])
            Ut[i, d] = matrix_to_group(nc, Gi @ group_to_matrix(nc, U[i, d]) @ Gj.conj().T)
            Et[i, d] = O.alg_act(nc, E[i, d], gt[i])
    b = oracle_grid(nc, dims, Ut, Et)
    a.step_links()
    b.step_links()
    for _ in range(3):
        a.store_fields(); a.step()
        b.store_fields(); b.step()
    for i in range(n):
        for d in range(3):
            j = a.shift(i, d, 1)
            Gi, Gj = group_to_matrix(nc, gt[i]), group_to_matrix(nc, gt[j])
            want = Gi @ group_to_matrix(nc, a.U[i, d]) @ Gj.conj().T
            assert np.allclose(group_to_matrix(nc, b.U[i, d]), want, atol=1e-12)
            assert np.allclose(O.alg_coeffs(nc, b.E[i, d]), O.alg_coeffs(nc, O.alg_act(nc, a.E[i, d], gt[i])),
                               atol=1e-12)
    # gauge-invariant observables agree
    assert np.allclose(a.measure()[:7], b.measure()[:7], rtol=1e-11)


# ------------------------------------------------------------------ Gauss law

@pytest.mark.parametrize("nc", [2, 3])
def test_gauss_law_is_conserved(nc):
    dims = (5, 4, 6)
    U, E = make_state(nc, dims, 0.5, 0.3, seed=21)
    og = oracle_grid(nc, dims, U, E, 0.25, (1.0, 1.1, 0.9))
    og.step_links()
    G0 = np.array([O.alg_coeffs(nc, og.gauss_constraint(i)) for i in range(og.n)])
    assert np.max(np.abs(G0)) > 1e-2  # random E violates Gauss's law
    for _ in range(30):
        og.store_fields()
        og.step()
    G1 = np.array([O.alg_coeffs(nc, og.gauss_constraint(i)) for i in range(og.n)])
    # G(x) is covariant; in temporal gauge it is constant up to rounding
    assert np.max(np.abs(G1 - G0)) < 1e-12


def test_gauss_constraint_closed_form():
    """GaussConstraintTest: E_{x,i}^a = A_i B^a (n.x) on U = 1 has |G|^2 = |B|^2 (A.n)^2,
    also after random gauge transformations and 200 steps (16 x 16 x 1, a = 0.05)."""
    nc, dims, a, at = 2, (16, 16, 1), 0.05, 0.00025
    A = np.array([0.345345, 0.0, 0.0])
    B = np.array([0.2345245, 0.02343, -1.234])
    nvec = np.array([1.567, 0.534, 0.0])
    og = O.OracleGrid(nc, dims, (a, a, a), at, 1.0)
    for i in range(og.n):
        p = index_to_pos(dims, i)
        nx = sum(p[j] * a * nvec[j] for j in range(3))
        for j in range(3):
            og.E[i, j] = A[j] * B * a * nx
    og.step_links(at)
    expected = np.sum(B * B) * np.dot(A, nvec) ** 2
    interior = [pos_to_index(dims, (x, y, 0)) for x in range(4, 12) for y in range(4, 12)]
    for step in range(200):
        if step % 50 == 0:
            gt = random_gauge(nc, og.n)
            og.gauge_transform(gt)
            for i in interior[::7]:
                assert abs(og.gauss_constraint_squared(i) - expected) < 1e-4
        og.store_fields()
        og.step(at)
    for i in interior:
        assert abs(og.gauss_constraint_squared(i) - expected) < 1e-4


# ------------------------------------------------------------------ exact solutions

@pytest.mark.parametrize("nc", [2, 3])
def test_vacuum_is_a_fixed_point(nc):
    dims = (3, 4, 3)
    og = O.OracleGrid(nc, dims)
    for _ in range(5):
        og.store_fields()
        og.step()
    ident = O.links_from_coeffs(nc, np.zeros((og.n, 3, nc * nc - 1)))
    assert np.max(np.abs(og.U - ident)) == 0.0
    assert np.max(np.abs(og.E)) == 0.0


@pytest.mark.parametrize("nc", [2, 3])
def test_uniform_field_in_one_direction(nc):
    """U = 1 and a uniform E along x0 feel no force: after k steps U_0 = exp(-i k dt E)."""
    dims = (3, 3, 4)
    c = RNG.normal(0, 0.3, nc * nc - 1)
    og = O.OracleGrid(nc, dims, (1.0, 1.0, 1.0), 0.2)
    for i in range(og.n):
        og.E[i, 0] = O.alg_from_coeffs(nc, c)
    og.step_links(0.2)
    for _ in range(7):
        og.store_fields()
        og.step(0.2)
    want = expm(-1j * 8 * 0.2 * coeffs_to_herm(nc, c))
    for i in range(og.n):
        assert np.allclose(group_to_matrix(nc, og.Unext[i, 0]), want, atol=1e-13)
        assert np.allclose(O.alg_coeffs(nc, og.E[i, 0]), c, atol=1e-14)


@pytest.mark.parametrize("a0,at", [(1.0, 0.25), (0.8, 0.1)])
def test_abelian_wave_dispersion(a0, at):
    """A small single-colour wave A_1(x0) = eps cos(k x0) obeys the leapfrog recurrence
    c_{m+1} + c_{m-1} = (2 - dt^2 K^2) c_m, K = (2/a0) sin(k a0 / 2)."""
    N = 32
    dims = (N, 1, 1)
    k = 2 * np.pi * 3 / (N * a0)
    eps = 1e-4
    og = O.OracleGrid(2, dims, (a0, 1.0, 1.0), at)
    xs = np.arange(N) * a0
    for x in range(N):
        og.U[x, 1] = O.alg_get_link(2, np.array([0.0, 0.0, eps * np.cos(k * xs[x])]))
    og.Unext[:] = og.U
    cs = []
    for _ in range(40):
        theta = np.array([2 * np.arctan2(og.Unext[x, 1, 3], og.Unext[x, 1, 0]) for x in range(N)])
        cs.append(np.dot(theta, np.cos(k * xs)) / np.dot(np.cos(k * xs), np.cos(k * xs)))
        og.store_fields()
        og.step(at)
    cs = np.array(cs)
    K = 2 / a0 * np.sin(k * a0 / 2)
    lhs = cs[2:] + cs[:-2]
    rhs = (2 - at * at * K * K) * cs[1:-1]
    assert np.max(np.abs(lhs - rhs)) < 1e-8 * eps
    # a wrong area factor (1/a instead of 1/a^2) would change K by 25 % at a0 = 0.8
    if a0 != 1.0:
        K_wrong = 2 / np.sqrt(a0) * np.sin(k * a0 / 2) / np.sqrt(a0) * np.sqrt(a0)
        assert np.max(np.abs(lhs - (2 - at * at * K_wrong ** 2) * cs[1:-1])) > 1e-3 * eps


@pytest.mark.parametrize("nc", [2, 3])
def test_leapfrog_is_second_order(nc):
    """At fixed physical time t = 4 the observables converge as dt^2: halving dt
    shrinks the change of the measured energy about fourfold. (E^2 + B^2 of
    BulkQuantitiesInTime is conserved only up to O(field^4), so it is compared
    across step sizes rather than in time.)"""
    dims = (6, 6, 6)
    U, E = make_state(nc, dims, 0.2, 0.1, seed=4)
    finals = []
    for at, steps in ((0.1, 40), (0.05, 80), (0.025, 160)):
        og = oracle_grid(nc, dims, U, E, at)
        og.step_links(at)
        for _ in range(steps):
            og.store_fields()
            og.step(at)
        finals.append(og.measure()[:6].sum())
    ratio = (finals[0] - finals[1]) / (finals[1] - finals[2])
    assert 3.0 < ratio < 5.5, finals


def test_weak_field_energy_is_conserved():
    """In the weak-field limit the measured energy is the conserved one: its
    oscillation is bounded, shrinks as dt^2, and does not drift."""
    nc, dims = 2, (6, 6, 6)
    U, E = make_state(nc, dims, 0.01, 0.005, seed=4)
    spreads = []
    for at, steps in ((0.1, 200), (0.05, 400)):
        og = oracle_grid(nc, dims, U, E, at)
        og.step_links(at)
        es = []
        for _ in range(steps):
            og.store_fields()
            og.step(at)
            es.append(og.measure()[:6].sum())
        es = np.array(es)
        spreads.append((es.max() - es.min()) / es.mean())
        assert abs(es[-steps // 4:].mean() - es[:steps // 4].mean()) / es.mean() < 5e-3
    assert spreads[0] < 5e-2
    assert 3.0 < spreads[0] / spreads[1] < 5.5, spreads


def test_measurements_brute_force():
    """E^2, B^2 and |G|^2 means against dense sums (FieldMeasurements units: divide by (g a_i)^2)."""
    nc, dims, spacing, g = 2, (3, 4, 3), (1.0, 0.9, 1.2), 1.5
    U, E = make_state(nc, dims, 0.5, 0.3, seed=5)
    og = oracle_grid(nc, dims, U, E, 0.25, spacing, g)
    og.Unext[:] = O.links_from_coeffs(nc, RNG.normal(0, 0.4, (og.n, 3, 3)))
    m = og.measure()
    n = og.n
    e2 = [sum(np.sum(O.alg_coeffs(nc, E[i, d]) ** 2) for i in range(n)) / (g * spacing[d]) ** 2 / n for d in range(3)]
    assert np.allclose(m[:3], e2, rtol=1e-13)
    b2 = []
    for d in range(3):
        j, k = (d + 1) % 3, (d + 2) % 3
        tot = 0.0
        for links in (og.U, og.Unext):
            for i in range(n):
                p = index_to_pos(dims, i)
                pj = pos_to_index(dims, [p[q] + unit(j)[q] for q in range(3)])
                pk = pos_to_index(dims, [p[q] + unit(k)[q] for q in range(3)])
                P = (group_to_matrix(nc, links[i, j]) @ group_to_matrix(nc, links[pj, k]) @
                     group_to_matrix(nc, links[pk, j]).conj().T @ group_to_matrix(nc, links[i, k]).conj().T)
                Bc = proj_matrix(nc, P) * spacing[d] / (spacing[j] * spacing[k])
                tot += 0.5 * np.sum(Bc ** 2)
        b2.append(tot / (g * spacing[d]) ** 2 / n)
    assert np.allclose(m[3:6], b2, rtol=1e-12)
    gs = 0.0
    for i in range(n):
        p = index_to_pos(dims, i)
        G = np.zeros(3)
        for d in range(3):
            im = pos_to_index(dims, [p[q] - unit(d)[q] for q in range(3)])
            V = group_to_matrix(nc, og.U[im, d])
            Em = V.conj().T @ coeffs_to_herm(nc, O.alg_coeffs(nc, E[im, d])) @ V
            G += (O.alg_coeffs(nc, E[i, d]) - herm_to_coeffs(nc, Em)) / (spacing[d] * g * spacing[d])
        gs += np.sum(G * G)
    assert abs(m[6] - gs / n) < 1e-12 * gs / n


def test_plausible_mistake_in_a_staple_is_caught():
    """Replacing one staple by its forward-only half changes the step visibly:
    the brute force above would flag such a restatement."""
    nc, dims = 2, (3, 3, 3)
    U, E = make_state(nc, dims, 0.6, 0.0, seed=2)
    og = oracle_grid(nc, dims, U, E)
    S_full = og.staple_sum(0, 0)
    # forward-only staple of direction 0 at cell 0
    fwd_only = np.zeros(4)
    for j in (1, 2):
        x1 = og.shift(0, 0, 1)
        x2 = og.shift(0, j, 1)
        m = group_to_matrix(nc, U[x1, j]) @ group_to_matrix(nc, U[x2, 0]).conj().T @ group_to_matrix(nc, U[0, j]).conj().T
        fwd_only = fwd_only + matrix_to_group(nc, m)
    assert np.max(np.abs(S_full - fwd_only)) > 1e-2
