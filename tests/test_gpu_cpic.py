"""CPIC on the device (super-particle NGP deposit, parallel transport, charge
density, Simulation step ordering) against the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
from openpixi_amd import FastTYMSolver, FieldMeasurements

from .helpers import device_grid, make_state, rel_err
from .test_oracle_cpic import two_nuclei

pytestmark = pytest.mark.gpu


def setup_pair(nc, dims, ppc, seed, precision="fp64", sigma_a=0.3, subdivisions=2, width_cells=1):
    U, E = make_state(nc, dims, sigma_a, 0.0, seed=seed)
    at = 1.0 / ppc
    og = O.OracleGrid(nc, dims, (1.0, 1.0, 1.0), at, 1.0)
    og.U[:] = U
    og.E[:] = E
    cols, charges = two_nuclei(nc, dims, ppc, seed, subdivisions, width_cells)
    args = (cols["orientation"], cols["num"], cols["index_offset"], cols["per_plane"], cols["shift"], cols["ppc"])
    og.set_superparticles(*args, charges)
    dg = device_grid(nc, dims, U, E, at, precision=precision)
    dg.setSuperParticles(*args, np.concatenate(charges))
    return og, dg


@pytest.mark.parametrize("kernel", ["link", "plaq", "stream", "tile", "cell"])
@pytest.mark.parametrize("nc,ppc", [(2, 4), (3, 4), (2, 8), (2, 1), (3, 2)])
def test_cpic_simulation_parity(gpu, nc, ppc, kernel):
    if kernel == "tile" and nc == 3:
        pytest.skip("the tile form is SU(2) only")
    dims = (14, 4, 3)
    print(f"cpic parity nc={nc} ppc={ppc} kernel={kernel}", flush=True)
    og, dg = setup_pair(nc, dims, ppc, seed=3, subdivisions=2, width_cells=2)
    dg.setKernel(kernel)
    og.initialize()
    dg.simulationInitialize()
    steps = 4 * ppc + 3
    og.sim_step(steps)
    assert dg.simulationStep(steps) == steps
    # integer parity: the cells that carry current and charge are identical
    Jo, Jd = og.J, dg.getJ()
    ro, rd = og.rho, dg.getRho()
    assert np.array_equal(np.abs(Jo).sum(axis=2) > 0, np.abs(Jd).sum(axis=2) > 0)
    assert np.array_equal(np.abs(ro).sum(axis=1) > 0, np.abs(rd).sum(axis=1) > 0)
    assert np.max(np.abs(Jo)) > 0 and np.max(np.abs(ro)) > 0
    tol = 1e-12
    assert rel_err(Jd, Jo) < tol
    assert rel_err(rd, ro) < tol
    assert rel_err(dg.superParticleCharges(), og.superparticle_charges()) < tol
    assert rel_err(dg.getE(), og.E) < tol
    assert rel_err(dg.getU(), og.U) < tol
    m_o, m_d = og.measure(), FieldMeasurements().measure(dg)
    for k in (0, 1, 2, 3, 4, 5, 6, 7):
        assert abs(m_d[k] - m_o[k]) <= 1e-11 * max(abs(m_o[k]), 1e-30), (k, m_d[k], m_o[k])


@pytest.mark.parametrize("kernel", ["link", "cell", "plaq"])
def test_current_everywhere_then_reset_then_particles(gpu, kernel):
    """The device keeps J zero outside the x0 planes it was written on and reads it only there
    (cell and per-link forms). A current uploaded on every cell enters the update everywhere;
    resetCurrent then clears all of it; particle deposits afterwards match the oracle."""
    nc, dims, ppc = 2, (14, 4, 3), 4
    og, dg = setup_pair(nc, dims, ppc, seed=7, subdivisions=2, width_cells=2)
    dg.setKernel(kernel)
    J = np.random.default_rng(1).normal(0, 0.1, size=og.J.shape)
    og.J[:] = J
    dg.setJ(J)
    og.store_fields()
    og.step(og.at)
    dg.storeFields()
    FastTYMSolver().step(dg, og.at)
    assert rel_err(dg.getE(), og.E) < 1e-12
    og.reset_current()
    dg.resetCurrent()
    assert not np.any(dg.getJ())
    og.initialize()
    dg.simulationInitialize()
    steps = 2 * ppc + 1
    og.sim_step(steps)
    assert dg.simulationStep(steps) == steps
    assert np.max(np.abs(og.J)) > 0
    assert rel_err(dg.getJ(), og.J) < 1e-12
    assert rel_err(dg.getE(), og.E) < 1e-12
    assert rel_err(dg.getU(), og.U) < 1e-12


@pytest.mark.parametrize("kernel", ["link", "cell"])
@pytest.mark.parametrize("nc", [2, 3])
def test_cpic_fp32_conserves_gauss(gpu, nc, kernel):
    """FP32 CPIC with reunitarization: G(x) stays at its initial value to FP32 rounding."""
    dims = (16, 6, 5)
    og, dg = setup_pair(nc, dims, 4, seed=5, precision="fp32")
    dg.setKernel(kernel)
    dg.simulationInitialize()
    m0 = FieldMeasurements().measure(dg)
    for _ in range(10):
        dg.simulationStep(4)
        dg.reunitarize()
    m1 = FieldMeasurements().measure(dg)
    assert m0[6] > 1e-3
    assert abs(m1[6] - m0[6]) < 1e-5 * m0[6]


def setup_particle_pair(nc, dims, ppc, seed, precision="fp64"):
    from .test_oracle_cpic import particle_nuclei

    U, E = make_state(nc, dims, 0.3, 0.0, seed=seed)
    at = 1.0 / ppc
    og = O.OracleGrid(nc, dims, (1.0, 1.0, 1.0), at, 1.0)
    og.U[:] = U
    og.E[:] = E
    parts = particle_nuclei(nc, dims, ppc, seed)
    og.set_particles(*parts)
    dg = device_grid(nc, dims, U, E, at, precision=precision)
    dg.setParticles(*parts)
    return og, dg


@pytest.mark.parametrize("kernel", ["link", "plaq", "stream", "tile", "cell"])
@pytest.mark.parametrize("nc,ppc", [(2, 4), (3, 4), (2, 1), (3, 2)])
def test_per_particle_simulation_parity(gpu, nc, ppc, kernel):
    if kernel == "tile" and nc == 3:
        pytest.skip("the tile form is SU(2) only")
    """Per-particle CGC NGP (CGCParticleInterpolationNGP + CGCParticleSolver) through
    Simulation steps: positions bit-exact, fields and charges within 1e-12."""
    dims = (12, 5, 4)
    print(f"per-particle parity nc={nc} ppc={ppc} kernel={kernel}", flush=True)
    og, dg = setup_particle_pair(nc, dims, ppc, seed=5)
    dg.setKernel(kernel)
    og.initialize()
    dg.simulationInitialize()
    steps = 3 * ppc + 2
    og.sim_step(steps)
    dg.simulationStep(steps)
    po0, po1, qo0, qo1 = og.particle_state()
    pd0, pd1, qd0, qd1 = dg.particleState()
    assert np.array_equal(pd0, po0) and np.array_equal(pd1, po1)
    assert np.array_equal(np.rint(pd1), np.rint(po1))
    tol = 1e-12
    assert rel_err(qd0, qo0) < tol and rel_err(qd1, qo1) < tol
    Jo, Jd = og.J, dg.getJ()
    ro, rd = og.rho, dg.getRho()
    assert np.array_equal(np.abs(Jo).sum(axis=2) > 0, np.abs(Jd).sum(axis=2) > 0)
    assert np.array_equal(np.abs(ro).sum(axis=1) > 0, np.abs(rd).sum(axis=1) > 0)
    assert np.max(np.abs(Jo)) > 0 and np.max(np.abs(ro)) > 0
    assert rel_err(Jd, Jo) < tol and rel_err(rd, ro) < tol
    assert rel_err(dg.getE(), og.E) < tol
    assert rel_err(dg.getU(), og.U) < tol


@pytest.mark.parametrize("nc", [2, 3])
def test_per_particle_fp32(gpu, nc):
    """FP32 fields and charges, double positions: positions stay bit-exact with the FP64
    oracle and G(x) keeps its initial value to FP32 rounding."""
    dims, ppc = (16, 6, 5), 4
    og, dg = setup_particle_pair(nc, dims, ppc, seed=6, precision="fp32")
    og.initialize()
    dg.simulationInitialize()
    m0 = FieldMeasurements().measure(dg)
    for _ in range(5):
        og.sim_step(4)
        dg.simulationStep(4)
        dg.reunitarize()
    m1 = FieldMeasurements().measure(dg)
    po0, po1, _, qo1 = og.particle_state()
    pd0, pd1, _, qd1 = dg.particleState()
    assert np.array_equal(pd0, po0) and np.array_equal(pd1, po1)
    assert rel_err(qd1, qo1) < 1e-5
    assert m0[6] > 1e-3
    assert abs(m1[6] - m0[6]) < 1e-5 * m0[6]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("nc", [2, 3])
def test_per_particle_deposits_are_deterministic(gpu, nc, precision):
    """The per-particle charge density and current are ordered gathers (a stable sort of the
    particles by target cell, each cell summed in particle order): two runs from the same state
    give bit-identical rho and J, with several particles per cell (ppc 4)."""
    dims, ppc = (16, 6, 5), 4
    runs = []
    for _ in range(2):
        og, dg = setup_particle_pair(nc, dims, ppc, seed=7, precision=precision)
        dg.simulationInitialize()
        dg.simulationStep(3 * ppc + 1)
        runs.append((dg.getRho(), dg.getJ(), dg.getE()))
    (r0, j0, e0), (r1, j1, e1) = runs
    assert np.max(np.abs(r0)) > 0 and np.max(np.abs(j0)) > 0
    assert np.array_equal(r0, r1) and np.array_equal(j0, j1) and np.array_equal(e0, e1)
    # cells do hold several particles' charges
    from .test_oracle_cpic import particle_nuclei

    pos0 = particle_nuclei(nc, dims, ppc, 7)[1]
    cells = [tuple(int(np.rint(x)) % n for x, n in zip(p, dims)) for p in np.asarray(pos0).reshape(-1, 3)]
    assert max(cells.count(c) for c in set(cells)) >= ppc


@pytest.mark.parametrize("nc", [2, 3])
def test_sp_chargedensity_ragged_layout(gpu, nc):
    """Charge density of super particles whose cell ranges are not plane aligned: lengths that
    are not multiples of a plane, offsets anywhere (ranges cut at both lattice ends, ranges
    that nest, overlap in part or leave gaps), both beams, several ppc and shifts. The device
    gathers per cell segment in particle-list order, as the oracle's sequential loop adds."""
    dims = (12, 5, 3)
    pp = dims[1] * dims[2]
    n = dims[0] * pp
    rng = np.random.default_rng(11)
    cols = dict(orientation=[], num=[], index_offset=[], per_plane=[], shift=[], ppc=[])
    charges = []
    for p in range(40):
        ppc = int(rng.choice([1, 2, 4]))
        cols["orientation"].append(int(rng.choice([1, -1])))
        num = int(rng.integers(1, 3 * pp))
        cols["num"].append(num)
        cols["index_offset"].append(int(rng.integers(-pp, n)))
        cols["per_plane"].append(pp)
        cols["shift"].append(int(rng.integers(0, ppc)))
        cols["ppc"].append(ppc)
        charges.append(O.algs_from_coeffs(nc, rng.normal(0, 1, (num, nc * nc - 1))))
    args = (cols["orientation"], cols["num"], cols["index_offset"], cols["per_plane"], cols["shift"], cols["ppc"])
    U, E = make_state(nc, dims, 0.3, 0.0, seed=2)
    og = O.OracleGrid(nc, dims, (1.0, 1.0, 1.0), 0.5, 1.0)
    og.U[:] = U
    og.E[:] = E
    og.set_superparticles(*args, charges)
    dg = device_grid(nc, dims, U, E, 0.5)
    dg.setSuperParticles(*args, np.concatenate(charges))
    print(f"ragged charge density nc={nc}", flush=True)
    for t in (0, 1, 3, 7, 12):
        og.reset_charge()
        dg.resetCharge()
        og.sp_interpolate_chargedensity(t)
        dg.interpolateChargedensity(t)
        ro, rd = og.rho, dg.getRho()
        assert np.max(np.abs(ro)) > 0
        assert np.array_equal(np.abs(ro).sum(axis=1) > 0, np.abs(rd).sum(axis=1) > 0), t
        if nc == 2:
            assert np.array_equal(rd, ro), t
        else:
            assert rel_err(rd, ro) < 1e-14, t


@pytest.mark.parametrize("kernel", ["cell", "link"])
@pytest.mark.parametrize("nc", [2, 3])
def test_simulation_reunit_option_equals_manual_reprojection(gpu, nc, kernel):
    """opx_simulation_step with OPX_OPT_REUNIT_EVERY = 3 (the cell form reprojects the new links
    inside its update, the per-link form by a pass after it) against single steps with
    opx_reunitarize after every third: after the step that follows a reprojection both runs hold
    the same U, Unext and E (the old links, which opx_reunitarize reprojects as well, are
    overwritten by that step): bit for bit with the pass after the per-link form, to FP32 rounding
    with the reprojection inside the cell form (a different compilation of the same arithmetic)."""
    dims = (8, 6, 5)
    U, E = make_state(nc, dims, 0.3, 0.2, seed=6)
    a = device_grid(nc, dims, U, E, precision="fp32")
    b = device_grid(nc, dims, U, E, precision="fp32")
    for g in (a, b):
        g.setKernel(kernel)
        g.simulationInitialize()
    a.setReunitarizeEvery(3)
    assert a.simulationStep(7) == 7
    for t in range(1, 8):
        b.simulationStep(1)
        if t % 3 == 0:
            b.reunitarize()
    for get in ("getU", "getUnext", "getE"):
        if kernel == "link":
            assert np.array_equal(getattr(a, get)(), getattr(b, get)()), get
        else:
            assert rel_err(getattr(a, get)(), getattr(b, get)()) < 1e-6, get
    # and the reprojection did happen: without it the links drift from the group differently
    c = device_grid(nc, dims, U, E, precision="fp32")
    c.setKernel(kernel)
    c.simulationInitialize()
    c.simulationStep(7)
    assert not np.array_equal(c.getU(), b.getU())
