"""GaussConstraintRestoration and UnitarityTester on the device (opx_gauss_restore,
opx_unitarity_test) against the oracle's restatement of
diagnostics/methods/GaussConstraintRestoration.java:79-163 and UnitarityTester.java:31-46:
FP64 within 1e-12 (fields, violations, the same stopping step and reason), FP32 within float
rounding, and a YAML Simulation run with both diagnostics."""
import io

import numpy as np
import pytest

import oracle as O
from openpixi_amd import FastTYMSolver, OpxError

from .helpers import device_grid, make_state, oracle_grid, rel_err

pytestmark = pytest.mark.gpu


def violating_pair(nc, dims, precision="fp64", with_rho=False, active=None, seed=0):
    """U random, E random (violating Gauss's law), Unext from stepLinks; optional charges."""
    U, E = make_state(nc, dims, 0.4, 0.3, seed)
    og = oracle_grid(nc, dims, U, E)
    dg = device_grid(nc, dims, U, E, precision=precision)
    if with_rho:
        n = int(np.prod(dims))
        rho = O.algs_from_coeffs(nc, np.random.default_rng(seed + 50).normal(0, 0.2, size=(n, nc * nc - 1)))
        og.rho[:] = rho
        dg.setRho(rho)
    if active is not None:
        og.set_active_region(*active)
        dg.setActiveRegion(*active)
    og.step_links(0.25)
    FastTYMSolver().stepLinks(dg, 0.25)
    return og, dg


CASES = [(2, False, None), (3, False, None), (2, True, ((1, 0, 1), (3, 3, 4))), (3, True, ((0, 1, 0), (4, 2, 5)))]
CASE_IDS = ["su2", "su3", "su2-rho-region", "su3-rho-region"]


@pytest.mark.parametrize("nc,with_rho,active", CASES, ids=CASE_IDS)
def test_restoration_matches_oracle_fp64(gpu, nc, with_rho, active):
    """A fixed number of descending corrections (reason 2): every iteration's state is compared
    through the final E, Unext and the violations."""
    dims = (5, 4, 6)
    print(f"gauss restore fp64 nc={nc} rho={with_rho} active={active}", flush=True)
    og, dg = violating_pair(nc, dims, with_rho=with_rho, active=active, seed=3)
    ko, ro, bo, ao, hist = og.gauss_restore(0.1, 12, 1e-30)
    kd, rd, bd, ad = dg.restoreGaussConstraint(0.1, 12, 1e-30)
    assert (ko, ro) == (12, 2) and (kd, rd) == (ko, ro)
    assert abs(bd - bo) < 1e-12 * bo and abs(ad - ao) < 1e-12 * ao
    # the reference's correction leaves the charges out, so it descends fastest without them
    assert ao < (0.5 if not with_rho else 0.8) * bo
    assert rel_err(dg.getE(), og.E) < 1e-12
    assert rel_err(dg.getUnext(), og.Unext) < 1e-12
    assert rel_err(dg.getU(), og.U) == 0.0


@pytest.mark.parametrize("nc", [2, 3])
def test_restoration_stops_like_the_oracle(gpu, nc):
    """The absolute-value stop (a goal between two of the oracle's iterates, so rounding cannot
    move it) and the undone overshoot (gamma far too large: E and Unext come back unchanged)."""
    dims = (4, 5, 3)
    og, dg = violating_pair(nc, dims, seed=4)
    _, _, _, _, hist = violating_pair(nc, dims, seed=4)[0].gauss_restore(0.1, 8, 1e-30)
    goal = float(np.sqrt(hist[5] * hist[6]))
    ko, ro, bo, ao, _ = og.gauss_restore(0.1, 50, goal)
    kd, rd, bd, ad = dg.restoreGaussConstraint(0.1, 50, goal)
    assert (ko, ro) == (6, 0) and (kd, rd) == (ko, ro)
    assert abs(ad - ao) < 1e-12 * ao and ad < goal
    assert rel_err(dg.getE(), og.E) < 1e-12
    # overshoot: undone, nothing changes
    E0, Un0 = dg.getE(), dg.getUnext()
    kd, rd, bd, ad = dg.restoreGaussConstraint(40.0, 5, 1e-30)
    assert (kd, rd) == (0, 1) and ad == bd
    assert np.array_equal(dg.getE(), E0) and np.array_equal(dg.getUnext(), Un0)
    assert dg.restoreGaussConstraint(0.1, 0, 1e-30)[:2] == (0, 3)


@pytest.mark.parametrize("nc", [2, 3])
def test_restoration_fp32(gpu, nc):
    dims = (6, 4, 5)
    og, dg = violating_pair(nc, dims, precision="fp32", seed=5)
    ko, ro, bo, ao, _ = og.gauss_restore(0.1, 10, 1e-30)
    kd, rd, bd, ad = dg.restoreGaussConstraint(0.1, 10, 1e-30)
    assert (kd, rd) == (ko, ro) == (10, 2)
    assert abs(bd - bo) < 1e-5 * bo and abs(ad - ao) < 1e-4 * ao
    assert rel_err(dg.getE(), og.E) < 2e-5


def test_restoration_refuses_uneven_spacings(gpu):
    U, E = make_state(2, (3, 3, 3), 0.2, 0.1, seed=6)
    dg = device_grid(2, (3, 3, 3), U, E, spacing=(1.0, 1.0, 0.5))
    with pytest.raises(OpxError, match="uneven"):
        dg.restoreGaussConstraint(0.1, 3, 1e-9)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("nc", [2, 3])
def test_unitarity_test_matches_oracle(gpu, nc, precision):
    dims = (3, 4, 5)
    U, E = make_state(nc, dims, 0.6, 0.0, seed=7)
    og = oracle_grid(nc, dims, U, E)
    ident = O.links_from_coeffs(nc, np.zeros(nc * nc - 1))
    bad = U.copy()
    bad[4, 1] = 1.001 * ident            # |Re tr| = 1.001 N (1.001^2 in the rebuilt FP32 third row)
    bad[9, 0, 0] = np.nan
    badn = U.copy()
    # SU(2): a negative trace beyond -N; SU(3) FP32 keeps two rows and rebuilds the third, so it
    # only represents scaled group elements (-1.001 times the identity would read back with trace -1)
    badn[9, 2] = (-1.001 if nc == 2 else 1.001) * ident
    badn[17, 0] = 1.001 * ident
    badn[20, 1] = ident                   # |Re tr| = N exactly: below N + 1e-8, not flagged
    og.U[:] = bad
    og.Unext[:] = badn
    dg = device_grid(nc, dims, bad, E, precision=precision)
    dg.setUnext(badn)
    cnt_o, flags = og.unitarity_flags(1e-8)
    cnt_d, cells = dg.unitarityTest(1e-8)
    want = [i for i in range(og.n) for j in range(3) if flags[i, j]]
    assert cnt_o == cnt_d == len(want) and cnt_o >= 4
    assert list(cells) == want
    # the links of an untouched lattice pass
    clean = device_grid(nc, dims, U, E, precision=precision)
    clean.setUnext(U)
    assert clean.unitarityTest(1e-8)[0] == 0


def test_yaml_simulation_with_restoration_and_unitarity_tester(gpu):
    """output.gaussConstraintRestoration (applyOnlyOnce at offset 0.5 = step 2) and
    output.unitarityTester in a YAML run: the fields after the run match the oracle run that
    applies the restoration after the same step, and the reference's message is printed."""
    from openpixi_amd.simulation import Settings, Simulation

    nc, dims = 2, (5, 4, 4)
    s = Settings.from_yaml(f"""
simulationType: temporal yang-mills
timeStep: 0.25
duration: 1.0
numberOfColors: {nc}
gridCells: {list(dims)}
output:
  unitarityTester:
    - interval: 0.5
  gaussConstraintRestoration:
    - gamma: 0.1
      maxIterations: 6
      absoluteValue: 1e-30
      offset: 0.5
      applyOnlyOnce: true
""")
    U, E = make_state(nc, dims, 0.4, 0.3, seed=8)
    sim = Simulation(s)
    sim.grid.setU(U)
    sim.grid.setE(E)
    out = io.StringIO()
    for d in s.diagnostics:
        d.out = out
    sim.initialize()
    sim.run()
    assert sim.totalSimulationSteps == 5
    og = oracle_grid(nc, dims, U, E)
    og.initialize()
    for step in range(1, 6):
        og.sim_step(1)
        if step == 2:
            assert og.gauss_restore(0.1, 6, 1e-30)[:2] == (6, 2)
    assert rel_err(sim.grid.getE(), og.E) < 1e-12
    assert rel_err(sim.grid.getU(), og.U) < 1e-12
    assert out.getvalue() == "GaussConstraintRestoration: Reached maximum number of iterations at step #5\n"
