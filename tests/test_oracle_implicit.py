"""Pins the oracle's ImplicitTYMSolver restatement (oracle/opx_oracle_implicit.c) and
the group logarithm / power its damping uses, with checks independent of its code:

* the logarithm against scipy's matrix logarithm and exponential, on every branch the
  reference takes (SU(2) asin form on both sides of e0 = 0; SU(3) eigen decomposition,
  the near-identity series and the degenerate-eigenvalue series);
* with no iterations the implicit step is the explicit FastTYMSolver step, bit for bit;
* one full implicit step (begin, two iterations with damping, end) against a dense
  numpy brute force with the beam-direction time levels written out as a table;
* gauge covariance of the undamped implicit step.
"""
import numpy as np
import pytest
from scipy.linalg import expm, logm

import oracle as O

from .dense import coeffs_to_herm, group_to_matrix, herm_to_coeffs, matrix_to_group, proj_matrix
from .helpers import make_state, oracle_grid

RNG = np.random.default_rng(31)


def random_su(nc, sigma):
    return O.links_from_coeffs(nc, RNG.normal(0, sigma, (nc * nc - 1,)))


def herm_of(nc, v):
    return coeffs_to_herm(nc, O.alg_coeffs(nc, v))


# ------------------------------------------------------------------ logarithm and power

@pytest.mark.parametrize("nc,sigma", [(2, 0.3), (2, 1.0), (3, 0.3), (3, 0.8), (3, 1.5)])
def test_log_inverts_exp(nc, sigma):
    for _ in range(20):
        u = random_su(nc, sigma)
        if nc == 2 and u[0] < 0:
            continue
        H = herm_of(nc, O.group_log(nc, u))
        assert abs(np.trace(H)) < 1e-12
        assert np.allclose(expm(1j * H), group_to_matrix(nc, u), atol=1e-12)


@pytest.mark.parametrize("nc", [2, 3])
def test_log_is_the_principal_logarithm_for_small_angles(nc):
    for _ in range(10):
        u = random_su(nc, 0.4)
        want = herm_to_coeffs(nc, -1j * logm(group_to_matrix(nc, u)))
        assert np.allclose(O.alg_coeffs(nc, O.group_log(nc, u)), want, atol=1e-12)


def test_su2_log_asin_branch_beyond_a_quarter_turn():
    """For e0 < 0, 2 asin|e| / |e| e is the logarithm of the element with e0 -> -e0."""
    for _ in range(10):
        u = random_su(2, 2.5)
        if u[0] >= 0:
            continue
        want = u.copy()
        want[0] = -u[0]
        assert np.allclose(O.group_pow(2, u, 1.0), want, atol=1e-12)


def test_su3_log_near_identity_and_degenerate_series():
    # Re tr U >= 3 - 1e-2: the 15-term series
    u = random_su(3, 0.03)
    assert np.trace(group_to_matrix(3, u)).real >= 2.99
    assert np.allclose(expm(1j * herm_of(3, O.group_log(3, u))), group_to_matrix(3, u), atol=1e-12)
    # a doubly degenerate spectrum diag(e^{it}, e^{it}, e^{-2it}) in a random basis: the 50-term series
    t = 0.2
    V = group_to_matrix(3, random_su(3, 1.0))
    M = V @ np.diag(np.exp(1j * np.array([t, t, -2 * t]))) @ V.conj().T
    u = matrix_to_group(3, M)
    assert np.trace(M).real < 2.99
    assert np.allclose(expm(1j * herm_of(3, O.group_log(3, u))), M, atol=1e-12)


@pytest.mark.parametrize("nc", [2, 3])
def test_pow(nc):
    ident = matrix_to_group(nc, np.eye(nc, dtype=complex))
    for _ in range(10):
        u = random_su(nc, 0.5)
        assert np.array_equal(O.group_pow(nc, u, 0.0), ident)
        M = group_to_matrix(nc, u)
        for x in (0.3, 0.5, 1.0, 1.7):
            assert np.allclose(group_to_matrix(nc, O.group_pow(nc, u, x)), expm(x * logm(M)), atol=1e-12)


# ------------------------------------------------------------------ implicit step

def pair(nc, dims, seed, J=True, active=None, sigma_a=0.4, sigma_e=0.2):
    U, E = make_state(nc, dims, sigma_a, sigma_e, seed)
    n = int(np.prod(dims))
    og = oracle_grid(nc, dims, U, E, 0.25, (1.0, 0.9, 1.2), 1.3)
    rng = np.random.default_rng(seed + 7)
    og.Unext[:] = O.links_from_coeffs(nc, rng.normal(0, sigma_a, (n, 3, nc * nc - 1)))
    if J:
        og.J[:] = O.algs_from_coeffs(nc, rng.normal(0, 0.1, (n, 3, nc * nc - 1)))
    if active is not None:
        og.set_active_region(*active)
    return og


@pytest.mark.parametrize("nc", [2, 3])
def test_no_iterations_is_the_explicit_step(nc):
    dims = (4, 5, 3)
    a = pair(nc, dims, 3, active=((1, 0, 0), (3, 3, 2)))
    b = pair(nc, dims, 3, active=((1, 0, 0), (3, 3, 2)))
    a.step()
    b.implicit_step(iterations=0)
    assert np.array_equal(a.E, b.E)
    assert np.array_equal(a.Unext, b.Unext)
    assert np.array_equal(a.U, b.U)


def idx(dims, p):
    return ((p[0] % dims[0]) * dims[1] + (p[1] % dims[1])) * dims[2] + (p[2] % dims[2])


def pos(dims, i):
    return [i // (dims[1] * dims[2]), (i // dims[2]) % dims[1], i % dims[2]]


def mv(p, d, o):
    q = list(p)
    q[d] += o
    return q


# time levels (of the d-link, of the i-link) for the plaquettes of link d in the
# (d, i) plane, beam direction 0: implicit in the transverse plane, and only the
# beam-direction link implicit in planes containing the beam
def time_levels(d, i, t):
    if d != 0 and i != 0:
        return t, t
    if d == 0:
        return t, 0
    return 0, t


def brute_force_implicit(og, dims, spacing, g, at, iterations, damping):
    nc, n = og.nc, og.n
    U = [[group_to_matrix(nc, og.U[c, d]) for d in range(3)] for c in range(n)]
    Uold = [[group_to_matrix(nc, og.Unext[c, d]) for d in range(3)] for c in range(n)]
    E0 = [[herm_of(nc, og.E[c, d]) for d in range(3)] for c in range(n)]
    J = [[herm_of(nc, og.J[c, d]) for d in range(3)] for c in range(n)]
    act = np.ones(n, bool)   # every cell active
    uf = [-(g * spacing[d]) * at for d in range(3)]

    def link_exp(Hm):
        # getLink(E * (-at)) with E as a Hermitian matrix: exp(i (-at) E)
        return expm(1j * (-at) * Hm)

    def staple_force(c, d):
        p = pos(dims, c)
        S = np.zeros((nc, nc), complex)
        for i in range(3):
            if i == d:
                continue
            u1 = U[idx(dims, mv(p, d, 1))][i] @ U[idx(dims, mv(p, i, 1))][d].conj().T @ U[c][i].conj().T
            xm = idx(dims, mv(p, i, -1))
            u2 = (U[xm][d] @ U[idx(dims, mv(mv(p, d, 1), i, -1))][i]).conj().T @ U[xm][i]
            S += (u1 + u2) / spacing[i] ** 2
        return U[c][d] @ S

    iE = [[E0[c][d].copy() for d in range(3)] for c in range(n)]
    iU = [[U[c][d].copy() for d in range(3)] for c in range(n)]
    iUn = [[Uold[c][d].copy() for d in range(3)] for c in range(n)]
    for c in range(n):
        if not act[c]:
            continue
        for d in range(3):
            iE[c][d] = iE[c][d] + coeffs_to_herm(nc, proj_matrix(nc, staple_force(c, d))) * at + J[c][d] * uf[d]
            iUn[c][d] = link_exp(iE[c][d]) @ U[c][d]

    def L(t, cell, dir_, cur_iU):
        return {-1: Uold, 0: U, 1: cur_iU}[t][cell][dir_]

    for _ in range(iterations):
        iU, iUn = iUn, iU
        new_iUn = [[iUn[c][d] for d in range(3)] for c in range(n)]
        for c in range(n):
            if not act[c]:
                continue
            p = pos(dims, c)
            for d in range(3):
                e = E0[c][d].copy()
                for t in (+1, -1):
                    S = np.zeros((nc, nc), complex)
                    for i in range(3):
                        if i == d:
                            continue
                        td, ti = time_levels(d, i, t)
                        xpd, xpi, xmi = idx(dims, mv(p, d, 1)), idx(dims, mv(p, i, 1)), idx(dims, mv(p, i, -1))
                        xpdmi = idx(dims, mv(mv(p, d, 1), i, -1))
                        u1 = L(ti, xpd, i, iU) @ L(td, xpi, d, iU).conj().T @ L(ti, c, i, iU).conj().T
                        u2 = (L(td, xmi, d, iU) @ L(ti, xpdmi, i, iU)).conj().T @ L(ti, xmi, i, iU)
                        S += L(td, c, d, iU) @ (u1 + u2) / spacing[i] ** 2
                    e = e + coeffs_to_herm(nc, proj_matrix(nc, S)) * (at * 0.5)
                e = e + J[c][d] * uf[d]
                iE[c][d] = e
                new = link_exp(e) @ U[c][d]
                # damping: U^damping Unext^(1 - damping) with principal logarithms
                new_iUn[c][d] = expm(damping * logm(iU[c][d])) @ expm((1 - damping) * logm(new))
        iUn = new_iUn
    Eout = [[iE[c][d] if act[c] else E0[c][d] for d in range(3)] for c in range(n)]
    Unext_out = [[iUn[c][d] if act[c] else Uold[c][d] for d in range(3)] for c in range(n)]
    return Eout, Unext_out


@pytest.mark.parametrize("nc", [2, 3])
def test_implicit_step_matches_dense_brute_force(nc):
    dims, spacing, g, at = (3, 4, 3), (1.0, 0.9, 1.2), 1.3, 0.25
    og = pair(nc, dims, 11, sigma_a=0.3, sigma_e=0.15)
    Eb, Ub = brute_force_implicit(og, dims, spacing, g, at, iterations=2, damping=0.3)
    og.implicit_step(iterations=2, damping=0.3)
    for c in range(og.n):
        for d in range(3):
            assert np.allclose(herm_of(nc, og.E[c, d]), Eb[c][d], atol=1e-11), (c, d)
            assert np.allclose(group_to_matrix(nc, og.Unext[c, d]), Ub[c][d], atol=1e-11), (c, d)


@pytest.mark.parametrize("nc,sigma_g", [(2, 0.2), (3, 1.0)])
def test_undamped_implicit_step_is_gauge_covariant(nc, sigma_g):
    """Undamped damping is pow(Unext, 1) = exp(log Unext). For SU(3) that is Unext on every
    branch. The SU(2) logarithm 2 asin|e| / |e| e reflects links with e0 < 0 (see above), so
    the SU(2) check keeps every link within a quarter turn: a gauge transformation near 1."""
    dims = (3, 4, 5)
    a = pair(nc, dims, 5)
    n = a.n
    gt = O.links_from_coeffs(nc, np.random.default_rng(2).normal(0, sigma_g, (n, 1, nc * nc - 1))).reshape(n, -1)
    b = pair(nc, dims, 5)
    for c in range(n):
        for d in range(3):
            j = b.shift(c, d, 1)
            Gi, Gj = group_to_matrix(nc, gt[c]), group_to_matrix(nc, gt[j])
            for arr in (b.U, b.Unext):
                arr[c, d] = matrix_to_group(nc, Gi @ group_to_matrix(nc, arr[c, d]) @ Gj.conj().T)
            b.E[c, d] = O.alg_act(nc, b.E[c, d], gt[c])
            b.J[c, d] = O.alg_act(nc, b.J[c, d], gt[c])
    a.implicit_step(iterations=3, damping=0.0)
    b.implicit_step(iterations=3, damping=0.0)
    for c in range(n):
        for d in range(3):
            j = a.shift(c, d, 1)
            Gi, Gj = group_to_matrix(nc, gt[c]), group_to_matrix(nc, gt[j])
            want = Gi @ group_to_matrix(nc, a.Unext[c, d]) @ Gj.conj().T
            assert np.allclose(group_to_matrix(nc, b.Unext[c, d]), want, atol=1e-12)
            assert np.allclose(O.alg_coeffs(nc, b.E[c, d]), O.alg_coeffs(nc, O.alg_act(nc, a.E[c, d], gt[c])),
                               atol=1e-12)


def test_iterations_converge():
    """Undamped iterations are a contraction for a small time step: successive E differ
    by geometrically decreasing amounts."""
    diffs = []
    prev = None
    for k in range(1, 7):
        og = pair(2, (4, 4, 4), 9, J=False)
        og.implicit_step(iterations=k, damping=0.0)
        if prev is not None:
            diffs.append(np.max(np.abs(og.E - prev)))
        prev = og.E.copy()
    assert all(b < 0.5 * a for a, b in zip(diffs, diffs[1:])), diffs
