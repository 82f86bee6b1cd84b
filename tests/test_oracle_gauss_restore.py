"""The oracle's restatement of GaussConstraintRestoration and UnitarityTester
(diagnostics/methods/GaussConstraintRestoration.java:79-163, UnitarityTester.java:31-46),
pinned by what the mathematics fixes rather than by retyping its formulas:

- the violation is exactly zero when Unext = U, and for a small time step it tends to
  (a g)^2 times the lattice Gauss residual (Grid.getGaussConstraint) with an O(dt^2) error,
  which fixes the signs, the link orientations and the 1/(a_t a) normalisation;
- it is gauge invariant and the correction is gauge covariant;
- one correction decreases the violation for a small gamma (a descent step), the kept
  corrections decrease it monotonically, and a correction that makes it worse is undone
  bit for bit;
- the unitarity flags equal a numpy evaluation of |Re tr U| < N + tol on U and Unext.
"""
import numpy as np
import pytest

import oracle as O

from .helpers import make_state, oracle_grid
from .test_oracle_lattice import random_gauge


def grid_after_step_links(nc, dims, at, spacing=1.0, g=1.0, sigma_a=0.4, sigma_e=0.3, seed=0):
    U, E = make_state(nc, dims, sigma_a, sigma_e, seed)
    og = oracle_grid(nc, dims, U, E, at, (spacing,) * 3, g)
    og.step_links(at)
    return og


@pytest.mark.parametrize("nc", [2, 3])
def test_violation_zero_when_links_do_not_move(nc):
    dims = (3, 4, 3)
    U, E = make_state(nc, dims, 0.5, 0.0, seed=1)
    og = oracle_grid(nc, dims, U, E)
    og.Unext[:] = og.U
    v, gv = og.restore_violation()
    assert v == 0.0 and np.max(np.abs(gv)) < 1e-14   # U U^+ - U^+ U: the identity up to rounding
    it, reason, b, a, hist = og.gauss_restore(0.25, 10, 1e-12)
    assert (it, reason, b, a) == (0, 0, 0.0, 0.0)


@pytest.mark.parametrize("spacing,g", [(1.0, 1.0), (0.5, 1.3)])
@pytest.mark.parametrize("nc", [2, 3])
def test_violation_tends_to_the_gauss_residual(nc, spacing, g):
    """Unext = exp(-i dt E) U: proj(gv) = -a g G + O(dt^2) with G the Gauss constraint, so the mean
    violation over (a g)^2 times the Gauss residual tends to 1, the error falling fourfold when dt
    halves."""
    dims = (4, 3, 5)
    errs = []
    for at in (0.02, 0.01, 0.005):
        og = grid_after_step_links(nc, dims, at, spacing, g, seed=2)
        v, _ = og.restore_violation()
        gauss = og.measure()[6]
        assert gauss > 1e-2
        errs.append(abs(v / ((spacing * g) ** 2 * gauss) - 1.0))
    assert errs[0] < 1e-3
    assert 3.0 < errs[0] / errs[1] < 5.0 and 3.0 < errs[1] / errs[2] < 5.0


@pytest.mark.parametrize("nc", [2, 3])
def test_gauge_invariance_and_covariance(nc):
    dims = (3, 4, 4)
    a = grid_after_step_links(nc, dims, 0.25, seed=3)
    b = grid_after_step_links(nc, dims, 0.25, seed=3)
    gt = random_gauge(nc, a.n)
    b.gauge_transform(gt)
    va, gva = a.restore_violation()
    vb, gvb = b.restore_violation()
    assert abs(va - vb) < 1e-12 * va
    a.restore_correct(gva, 0.2)
    b.restore_correct(gvb, 0.2)
    # transforming the corrected field of `a` gives the corrected field of `b`
    a.gauge_transform(gt)
    assert np.max(np.abs(a.E - b.E)) < 1e-12 * np.max(np.abs(b.E))
    assert np.max(np.abs(a.Unext - b.Unext)) < 1e-12


@pytest.mark.parametrize("nc", [2, 3])
def test_restoration_descends_monotonically(nc):
    dims = (4, 4, 3)
    og = grid_after_step_links(nc, dims, 0.25, seed=4)
    v0, _ = og.restore_violation()
    it, reason, before, after, hist = og.gauss_restore(0.1, 40, 1e-30)
    assert before == v0 and it >= 10
    assert np.all(np.diff(hist) <= 0.0) and hist[-1] == after
    assert after < 1e-2 * before
    v_end, _ = og.restore_violation()
    assert v_end == after      # the state left behind is the last kept one
    assert reason in (1, 2)


@pytest.mark.parametrize("nc", [2, 3])
def test_worsening_correction_is_undone(nc):
    """gamma far beyond the stable range: the first correction overshoots, the violation grows,
    and the grid returns to its input bit for bit (restoreGrid)."""
    dims = (3, 3, 4)
    og = grid_after_step_links(nc, dims, 0.25, seed=5)
    E0, Un0, U0 = og.E.copy(), og.Unext.copy(), og.U.copy()
    it, reason, before, after, _ = og.gauss_restore(40.0, 10, 1e-30)
    assert (it, reason) == (0, 1) and after == before
    assert np.array_equal(og.E, E0) and np.array_equal(og.Unext, Un0) and np.array_equal(og.U, U0)


def test_absolute_value_and_iteration_limit():
    og = grid_after_step_links(2, (3, 3, 3), 0.25, seed=6)
    v0, _ = og.restore_violation()
    assert og.gauss_restore(0.1, 5, 10 * v0)[:2] == (0, 0)          # already below the goal
    it, reason, _, _, hist = og.gauss_restore(0.1, 3, 1e-30)
    assert (it, reason) == (3, 2) and len(hist) == 4
    assert og.gauss_restore(0.1, 0, 1e-30)[:2] == (0, 3)


def test_uneven_spacings_refused():
    U, E = make_state(2, (3, 3, 3), 0.2, 0.1, seed=7)
    og = oracle_grid(2, (3, 3, 3), U, E, 0.25, (1.0, 1.0, 0.5))
    with pytest.raises(RuntimeError):
        og.gauss_restore(0.1, 3, 1e-9)


@pytest.mark.parametrize("nc", [2, 3])
def test_unitarity_flags(nc):
    dims = (3, 2, 4)
    U, E = make_state(nc, dims, 0.6, 0.0, seed=8)
    og = oracle_grid(nc, dims, U, E)
    og.Unext[:] = og.U
    og.U[0, 0] = O.links_from_coeffs(nc, np.zeros(nc * nc - 1))   # identity: |tr| = N exactly, not flagged
    og.U[2, 1] *= 1.0 + 1e-6
    og.Unext[5, 2] *= 1.0 + 1e-6
    og.U[7, 0, 0] = np.nan
    cnt, flags = og.unitarity_flags(1e-8)

    def real_trace(u):
        return 2.0 * u[..., 0] if nc == 2 else u[..., 0] + u[..., 4] + u[..., 8]

    with np.errstate(invalid="ignore"):
        ok = (np.abs(real_trace(og.U)) < nc + 1e-8) & (np.abs(real_trace(og.Unext)) < nc + 1e-8)
    assert np.array_equal(flags.astype(bool), ~ok)
    assert cnt == int((~ok).sum())
    # the scaled links are flagged only where their trace leaves the range
    assert flags[7, 0] == 1 and flags[0, 0] == 0
