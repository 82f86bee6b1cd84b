"""The CGC initial condition on the device against the oracle (beam along x0):

* MV charges from given noise, on power-of-two and other axis lengths (radix-2 and direct
  transforms) and with an active region;
* the light-cone solve on non-trivial links with a charge density already on the grid:
  U, Unext, E and the Gauss violation of the Wilson-line copy, both orientations, FP64 and FP32;
* the super particles: descriptors, refined charges, uneven subdivisions, a second nucleus
  appended to the list;
* a seeded YAML run of "temporal optimized cgc ngp" (JavaRandom noise, two nuclei, particles,
  initialize and steps) against the same run of the oracle;
* the same YAML run through Simulation.run() with the tadpole, dipole and bulk files written.

FP64 bars are relative to the largest value: 1e-12 for the charges, fields and Gauss
violation, 1e-11 for the refined charges (200 sweeps of a 4- and a 6-point stencil each).
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from openpixi_amd import Grid, _lib
from openpixi_amd.cgc import (CGCInitialCondition, ChargeDensity, JavaRandom, LightConeNGPSuperParticleCreator,
                              LightConePoissonSolverImprovedFull, MVModel)
from openpixi_amd.simulation import Settings, Simulation

from .helpers import rel_err

pytestmark = pytest.mark.gpu

SPACING = (1.0, 0.9, 0.9)
AT = 0.25
G = 1.3


def grids(nc, dims, precision="fp64", spacing=SPACING, at=AT):
    og = O.OracleGrid(nc, dims, spacing, at, G)
    dg = Grid(nc, dims, list(spacing), at, G, precision)
    return og, dg


def mv_params(dims):
    L = dims[0] * SPACING[0]
    return dict(location=0.45 * L, width=0.1 * L, mu=0.8, uvt=2.5, long_width=0.6, ir=0.4)


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("dims,active", [((16, 8, 8), None), ((12, 6, 10), None), ((9, 5, 7), ((1, 0, 1), (7, 4, 5)))])
def test_mv_charge_matches_oracle(gpu, nc, dims, active):
    print(f"MV charge nc={nc} dims={dims} active={active}", flush=True)
    og, dg = grids(nc, dims)
    if active:
        og.set_active_region(*active)
        dg.setActiveRegion(*active)
    noise = np.random.default_rng(3).normal(size=(nc * nc - 1) * og.n)
    p = mv_params(dims)
    want = og.mv_charge(p["location"], p["width"], p["mu"], p["uvt"], p["long_width"], p["ir"], noise)
    m = MVModel(0, 1, p["location"], p["width"], p["mu"], True, 0, p["uvt"], p["long_width"], p["ir"])
    m.initialize(dg, noise=noise)
    got = m.getChargeDensity()
    assert np.max(np.abs(want)) > 0
    assert rel_err(got, want) < 1e-12


def solve_setup(nc, dims, precision, seed=7):
    """Both lattices with the same random links at both time levels and a grid charge density;
    returns (og, dg, rho) with rho an MV charge density of the oracle."""
    og, dg = grids(nc, dims, precision)
    rng = np.random.default_rng(seed)
    n = og.n
    U = O.links_from_coeffs(nc, rng.normal(0, 0.3, (n, 3, nc * nc - 1)))
    Un = O.links_from_coeffs(nc, rng.normal(0, 0.3, (n, 3, nc * nc - 1)))
    grid_rho = O.algs_from_coeffs(nc, rng.normal(0, 0.05, (n, nc * nc - 1)))
    og.U[:] = U
    og.Unext[:] = Un
    og.rho[:] = grid_rho
    dg.setU(U)
    dg.setUnext(Un)
    dg.setRho(grid_rho)
    p = mv_params(dims)
    rho = og.mv_charge(p["location"], p["width"], p["mu"], p["uvt"], p["long_width"], p["ir"],
                       rng.normal(size=(nc * nc - 1) * n))
    return og, dg, rho


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
@pytest.mark.parametrize("orientation", [1, -1])
@pytest.mark.parametrize("nc", [2, 3])
def test_lightcone_solve_matches_oracle(gpu, nc, orientation, precision, tol):
    dims = (10, 6, 8)
    print(f"light-cone solve nc={nc} orientation={orientation} {precision}", flush=True)
    og, dg, rho = solve_setup(nc, dims, precision)
    gauss = og.lightcone_solve(rho, orientation)
    density = ChargeDensity(rho, 0, orientation)
    density.initialize(dg)
    solver = LightConePoissonSolverImprovedFull()
    solver.initialize(dg)
    solver.solve(density)
    assert rel_err(dg.getU(), og.U) < tol
    assert rel_err(dg.getUnext(), og.Unext) < tol
    assert rel_err(dg.getE(), og.E) < tol
    # the FP32 Gauss violation is evaluated in double from float-stored longitudinal links and rho
    assert rel_err(solver.getGaussViolation(), gauss) < (tol if precision == "fp64" else 1e-4)


def sp_lists_equal(sim_sp, arrs, charges, tol):
    assert list(sim_sp.orientation) == list(arrs["orientation"])
    assert list(sim_sp.number) == list(arrs["num"])
    assert list(sim_sp.index_offset) == list(arrs["index_offset"])
    assert list(sim_sp.per_plane) == list(arrs["per_plane"])
    assert list(sim_sp.sub_lattice_shift) == list(arrs["shift"])
    assert list(sim_sp.particles_per_cell) == list(arrs["ppc"])
    want = np.concatenate(charges)
    assert sim_sp.charges.shape == want.shape
    assert rel_err(sim_sp.charges, want) < tol


class _Sim:
    """The parts of a Simulation the super-particle creator reads."""

    def __init__(self, grid, threads):
        self.grid = grid

        class _S:
            number_of_threads = threads
            simulation_type = "TemporalOptimizedCGCNGP"

        self.settings = _S()


def device_superparticles(dg):
    sim = Simulation.__new__(Simulation)
    sim.grid = dg
    return Simulation.superparticle_descriptors(sim)


@pytest.mark.parametrize("nc,orientation,threads", [(2, 1, 3), (2, -1, 1), (3, 1, 4), (3, -1, 2)])
def test_superparticles_match_oracle(gpu, nc, orientation, threads):
    dims = (16, 4, 6)
    print(f"super particles nc={nc} orientation={orientation} threads={threads}", flush=True)
    og, dg, rho = solve_setup(nc, dims, "fp64", seed=11)
    gauss = og.lightcone_solve(rho, orientation)
    arrs, charges, (zs, bw) = og.lightcone_superparticles(gauss, orientation, threads)
    density = ChargeDensity(rho, 0, orientation)
    density.initialize(dg)
    solver = LightConePoissonSolverImprovedFull()
    solver.initialize(dg)
    solver.solve(density)
    creator = LightConeNGPSuperParticleCreator()
    creator.initialize(_Sim(dg, threads), 0, orientation)
    assert (creator.zStart, creator.blockWidth) == (zs, bw)
    assert creator.added == len(arrs["num"])
    sp_lists_equal(device_superparticles(dg), arrs, charges, 1e-11)


def test_superparticles_of_a_narrow_block(gpu):
    """A charge on a few planes only: the block is found from the plane maxima, and ws = 0
    subdivisions (more threads than planes) take one plane each."""
    nc, dims, orientation = 2, (20, 4, 4), 1
    og, dg = grids(nc, dims)
    rng = np.random.default_rng(5)
    rho = np.zeros((og.n, 3))
    T = dims[1] * dims[2]
    rho[8 * T:11 * T] = rng.normal(0, 1.0, (3 * T, 3))
    gauss = og.lightcone_solve(rho, orientation)
    arrs, charges, (zs, bw) = og.lightcone_superparticles(gauss, orientation, 16)
    assert 0 < bw < dims[0] - 2
    ChargeDensity(rho, 0, orientation).initialize(dg)
    solver = LightConePoissonSolverImprovedFull()
    solver.initialize(dg)
    solver.solve(ChargeDensity(rho, 0, orientation))
    creator = LightConeNGPSuperParticleCreator()
    creator.initialize(_Sim(dg, 16), 0, orientation)
    assert (creator.zStart, creator.blockWidth) == (zs, bw)
    sp_lists_equal(device_superparticles(dg), arrs, charges, 1e-11)


def test_superparticles_of_a_long_block(gpu):
    """A block of 598 planes at 4 charges per cell: lists of 2392 charges, above the length up to
    which the device refines through the sweep matrices' powers, so the lists are swept one by
    one as the reference does."""
    nc, dims, orientation = 2, (600, 2, 2), -1
    print("super particles, long block", flush=True)
    og, dg = grids(nc, dims)
    rho = np.random.default_rng(9).normal(0, 1.0, (og.n, 3))
    gauss = og.lightcone_solve(rho, orientation)
    arrs, charges, (zs, bw) = og.lightcone_superparticles(gauss, orientation, 5)
    assert bw * 4 > 2048
    ChargeDensity(rho, 0, orientation).initialize(dg)
    solver = LightConePoissonSolverImprovedFull()
    solver.initialize(dg)
    solver.solve(ChargeDensity(rho, 0, orientation))
    creator = LightConeNGPSuperParticleCreator()
    creator.initialize(_Sim(dg, 5), 0, orientation)
    assert (creator.zStart, creator.blockWidth) == (zs, bw)
    sp_lists_equal(device_superparticles(dg), arrs, charges, 1e-11)


YAML = """
simulationType: temporal optimized cgc ngp
gridStep: 1
timeStep: 0.25
duration: 1.0
numberOfDimensions: 3
numberOfColors: {nc}
numberOfThreads: 3
couplingConstant: 1.0
gridCells: [24, 6, 6]
initialConditions:
  CGC:
    poissonSolver: improved full
    MVModel:
      - direction: 0
        orientation: 1
        longitudinalLocation: 7.0
        longitudinalWidth: 1.2
        mu: 0.5
        randomSeed: 11
        ultravioletCutoffTransverse: 3.0
        longitudinalCoherenceLength: 0.4
        infraredCoefficient: 0.3
      - direction: 0
        orientation: -1
        longitudinalLocation: 17.0
        longitudinalWidth: 1.2
        mu: 0.5
        randomSeed: 12
        ultravioletCutoffTransverse: 3.0
        longitudinalCoherenceLength: 0.4
        infraredCoefficient: 0.3
"""


@pytest.mark.parametrize("nc", [2, 3])
def test_seeded_yaml_collision_matches_oracle(gpu, nc):
    """Two nuclei from a reference-style YAML file: the device run (java.util.Random noise drawn
    by JavaRandom, charges, fields, particle list, initialize and four steps) against the oracle
    fed the same noise."""
    print(f"YAML CGC collision nc={nc}", flush=True)
    settings = Settings.from_yaml(YAML.format(nc=nc))
    sim = Simulation(settings)
    dims = tuple(settings.grid_cells)
    og = O.OracleGrid(nc, dims, (1.0, 1.0, 1.0), 0.25, 1.0)
    n = og.n
    all_arrs, all_charges = None, []
    for seed, orient, loc in ((11, 1, 7.0), (12, -1, 17.0)):
        noise = JavaRandom(seed).gaussians((nc * nc - 1) * n)
        rho = og.mv_charge(loc, 1.2, 0.5, 3.0, 0.4, 0.3, noise)
        gauss = og.lightcone_solve(rho, orient)
        arrs, charges, _ = og.lightcone_superparticles(gauss, orient, 3)
        all_charges += charges
        all_arrs = arrs if all_arrs is None else {k: np.concatenate([all_arrs[k], arrs[k]]) for k in arrs}
    sp_lists_equal(sim.superparticle_descriptors(), all_arrs, all_charges, 1e-11)
    assert rel_err(sim.grid.getU(), og.U) < 1e-12
    assert rel_err(sim.grid.getUnext(), og.Unext) < 1e-12
    assert rel_err(sim.grid.getE(), og.E) < 1e-12
    og.set_superparticles(all_arrs["orientation"], all_arrs["num"], all_arrs["index_offset"],
                          all_arrs["per_plane"], all_arrs["shift"], all_arrs["ppc"], all_charges)
    og.initialize()
    sim.initialize()
    og.sim_step(4)
    sim.step(4)
    assert rel_err(sim.grid.getE(), og.E) < 1e-11
    assert rel_err(sim.grid.getU(), og.U) < 1e-11
    assert rel_err(sim.grid.getRho(), og.rho) < 1e-11
    assert rel_err(sim.superparticle_charges(), og.superparticle_charges()) < 1e-11


def test_cgc_entry_points_reject_bad_calls(gpu):
    dg = Grid(2, (8, 4, 4), [1.0, 1.0, 1.0], 0.25, 1.0)
    L = _lib.load()
    assert L.opx_cgc_lightcone_solve(dg.handle, 1, None, 0) == _lib.OPX_E_STATE      # no charge density yet
    assert L.opx_cgc_create_superparticles(dg.handle, 1, 1, None, None, None) == _lib.OPX_E_STATE
    ChargeDensity(np.zeros((dg.total_cells, 3))).initialize(dg)
    assert L.opx_cgc_lightcone_solve(dg.handle, 0, None, 0) == _lib.OPX_E_ARG         # orientation 0
    assert L.opx_cgc_lightcone_solve(dg.handle, 1, None, 0) == 0
    # no plane exceeds the cutoff of a zero violation: the reference's block is then every plane but
    # the last, with zero charges
    zs, bw = ctypes.c_int(), ctypes.c_int()
    assert L.opx_cgc_create_superparticles(dg.handle, 1, 2, None, ctypes.byref(zs), ctypes.byref(bw)) == 0
    assert (zs.value, bw.value) == (0, 7)
    assert L.opx_cgc_create_superparticles(dg.handle, 1, 0, None, None, None) == _lib.OPX_E_ARG
    z = np.zeros(5)
    assert L.opx_cgc_mv_charge(dg.handle, 1.0, 1.0, 1.0, 1.0, 0.0, 0.0, _lib.dptr(z), z.size, None, 0) == _lib.OPX_E_ARG


@pytest.mark.parametrize("nc,orientation", [(2, 1), (3, -1)])
def test_wilson_observables_match_oracle(gpu, nc, orientation, tmp_path, monkeypatch):
    """Tadpole and dipole of the device's lines behind the nucleus against the oracle's, and the
    files WilsonLineObservables appends (info line, header, one line per bin)."""
    from openpixi_amd.cgc import WilsonLineObservables
    from openpixi_amd.diagnostics import java_sci

    dims = (12, 6, 8)
    print(f"Wilson observables nc={nc} orientation={orientation}", flush=True)
    og, dg, rho = solve_setup(nc, dims, "fp64", seed=23)
    og.lightcone_solve(rho, orientation)
    density = ChargeDensity(rho, 0, orientation)
    density.initialize(dg)
    solver = LightConePoissonSolverImprovedFull()
    solver.initialize(dg)
    solver.solve(density)
    w = WilsonLineObservables(dg, solver, density)
    want_t = O.wilson_tadpole(nc, og.VT)
    assert abs(w.tadpole() - want_t) < 1e-12 * nc
    d, corr, counts = w.dipole()
    nbins = int(np.sqrt(dims[1] * dims[2])) // 2
    want_c, want_n = O.wilson_dipole(nc, dims[1], dims[2], SPACING[1], SPACING[2], og.VT, nbins)
    assert len(corr) == nbins and list(counts) == list(want_n)
    assert np.max(np.abs(corr - want_c)) < 1e-12 * nc
    monkeypatch.chdir(tmp_path)
    w.computeTadpole("t.dat")
    w.computeDipole("d.dat")
    t = (tmp_path / "output" / "t.dat").read_text().splitlines()
    assert len(t) == 1 and t[0].startswith("charge density, V: ")
    assert float(t[0].split("V: ")[1]) == pytest.approx(want_t, abs=1e-12 * nc)
    lines = (tmp_path / "output" / "d.dat").read_text().splitlines()
    assert lines[:2] == ["charge density", "d\ttr(V V^t)"] and len(lines) == 2 + nbins
    for b, line in enumerate(lines[2:]):
        x, v = line.split("\t")
        assert x == java_sci(d[b])
        assert float(v) == pytest.approx(want_c[b], abs=1e-12 * nc)


@pytest.mark.parametrize("seed,count", [(42, 1), (0, 2), (-7, 1001), (12345, 100000), (2**40 + 3, 30001)])
def test_device_java_gaussians_equal_the_host_sequence(gpu, seed, count):
    """java.util.Random(seed).nextGaussian() on the device: bit for bit the host JavaRandom
    sequence (whose first values the JDK's are), odd and even counts."""
    got = np.zeros(count)
    assert _lib.load().opx_java_gaussians(seed, count, _lib.dptr(got)) == 0
    assert np.array_equal(got, JavaRandom(seed).gaussians(count))
    if seed == 42:
        assert got[0] == 1.1419053154730547


@pytest.mark.parametrize("nc", [2, 3])
def test_seeded_mv_charge_equals_host_noise(gpu, nc):
    dims = (12, 6, 10)
    og, dg = grids(nc, dims)
    p = mv_params(dims)
    m = MVModel(0, 1, p["location"], p["width"], p["mu"], True, 77, p["uvt"], p["long_width"], p["ir"])
    m.initialize(dg)
    seeded = m.getChargeDensity()
    m.initialize(dg, noise=JavaRandom(77).gaussians((nc * nc - 1) * og.n))
    assert np.array_equal(seeded, m.getChargeDensity())
    want = og.mv_charge(p["location"], p["width"], p["mu"], p["uvt"], p["long_width"], p["ir"],
                        JavaRandom(77).gaussians((nc * nc - 1) * og.n))
    assert rel_err(seeded, want) < 1e-12


@pytest.mark.parametrize("nc,orientation", [(2, 1), (2, -1), (3, 1), (3, -1)])
def test_particles_match_oracle(gpu, nc, orientation):
    """LightConeNGPParticleCreator on the device: positions at t = 0 and t = at bit for bit,
    charges (Q0 = Q1) against the oracle's refined list."""
    from openpixi_amd.cgc import LightConeNGPParticleCreator

    dims = (16, 4, 6)
    print(f"particles nc={nc} orientation={orientation}", flush=True)
    og, dg, rho = solve_setup(nc, dims, "fp64", seed=21)
    gauss = og.lightcone_solve(rho, orientation)
    _, p0, p1, _, q, (zs, ze) = og.lightcone_particles(gauss, orientation)
    density = ChargeDensity(rho, 0, orientation)
    density.initialize(dg)
    solver = LightConePoissonSolverImprovedFull()
    solver.initialize(dg)
    solver.solve(density)
    creator = LightConeNGPParticleCreator()
    creator.initialize(_Sim(dg, 1), 0, orientation)
    assert (creator.zStart, creator.zEnd, creator.added) == (zs, ze, len(q))
    d0, d1, dq0, dq1 = dg.particleState()
    assert np.array_equal(d0, p0) and np.array_equal(d1, p1)
    assert rel_err(dq0, q) < 1e-11 and np.array_equal(dq0, dq1)


@pytest.mark.parametrize("nc", [2, 3])
def test_seeded_yaml_collision_per_particle_matches_oracle(gpu, nc):
    """The YAML collision of "temporal cgc ngp": per-particle CGC particles for two nuclei,
    initialize and four steps, against the oracle fed the same noise."""
    print(f"YAML CGC per-particle collision nc={nc}", flush=True)
    text = YAML.format(nc=nc).replace("temporal optimized cgc ngp", "temporal cgc ngp")
    settings = Settings.from_yaml(text)
    sim = Simulation(settings)
    dims = tuple(settings.grid_cells)
    og = O.OracleGrid(nc, dims, (1.0, 1.0, 1.0), 0.25, 1.0)
    parts = []
    for seed, orient, loc in ((11, 1, 7.0), (12, -1, 17.0)):
        noise = JavaRandom(seed).gaussians((nc * nc - 1) * og.n)
        rho = og.mv_charge(loc, 1.2, 0.5, 3.0, 0.4, 0.3, noise)
        gauss = og.lightcone_solve(rho, orient)
        parts.append(og.lightcone_particles(gauss, orient))
    d = np.concatenate([p[0] for p in parts])
    p0 = np.concatenate([p[1] for p in parts])
    v = np.concatenate([p[3] for p in parts])
    q = np.concatenate([p[4] for p in parts])
    g0, _, gq0, _ = sim.grid.particleState()
    assert np.array_equal(g0, p0) and rel_err(gq0, q) < 1e-11
    assert rel_err(sim.grid.getE(), og.E) < 1e-12
    og.set_particles(d, p0, v, q)
    og.initialize()
    sim.initialize()
    og.sim_step(4)
    sim.step(4)
    assert rel_err(sim.grid.getE(), og.E) < 1e-11
    assert rel_err(sim.grid.getU(), og.U) < 1e-11
    assert rel_err(sim.grid.getRho(), og.rho) < 1e-11
    pos0, _, _, Q1 = og.particle_state()
    g0, _, _, gQ1 = sim.grid.particleState()
    assert np.array_equal(g0, pos0) and rel_err(gQ1, Q1) < 1e-11


def test_yaml_collision_run_writes_reference_files(gpu, tmp_path, monkeypatch):
    """The whole reference-style run: a YAML file with two seeded nuclei asking for the tadpole
    and dipole of each nucleus's Wilson lines and for bulkQuantitiesInTime, through
    Simulation(...), initialize() and run(). The three files under output/ hold what the oracle
    run of the same noise gives: one tadpole line and one dipole block per nucleus (in the order
    of the MVModel list), and the bulk rows of steps 0, 2 and 4."""
    from openpixi_amd.diagnostics import java_sci

    nc = 3
    print(f"YAML CGC run with output files nc={nc}", flush=True)
    text = YAML.format(nc=nc).replace("    poissonSolver: improved full",
                                      "    poissonSolver: improved full\n    computeTadpole: true\n"
                                      "    computeDipole: true\n    dipoleFilename: wilson/dipole.dat")
    text += "output:\n  bulkQuantitiesInTime:\n    - path: bulk.dat\n      interval: 0.5\n"
    settings = Settings.from_yaml(text)
    monkeypatch.chdir(tmp_path)
    sim = Simulation(settings)
    sim.initialize()
    sim.run()
    assert sim.totalSimulationSteps == 5   # duration / timeStep + 1 steps

    dims = tuple(settings.grid_cells)
    og = O.OracleGrid(nc, dims, (1.0, 1.0, 1.0), 0.25, 1.0)
    nbins = int(np.sqrt(dims[1] * dims[2])) // 2
    all_arrs, all_charges, tadpoles, dipoles = None, [], [], []
    for seed, orient, loc in ((11, 1, 7.0), (12, -1, 17.0)):
        rho = og.mv_charge(loc, 1.2, 0.5, 3.0, 0.4, 0.3, JavaRandom(seed).gaussians((nc * nc - 1) * og.n))
        gauss = og.lightcone_solve(rho, orient)
        tadpoles.append(O.wilson_tadpole(nc, og.VT))
        dipoles.append(O.wilson_dipole(nc, dims[1], dims[2], 1.0, 1.0, og.VT, nbins))
        arrs, charges, _ = og.lightcone_superparticles(gauss, orient, 3)
        all_charges += charges
        all_arrs = arrs if all_arrs is None else {k: np.concatenate([all_arrs[k], arrs[k]]) for k in arrs}
    og.set_superparticles(all_arrs["orientation"], all_arrs["num"], all_arrs["index_offset"],
                          all_arrs["per_plane"], all_arrs["shift"], all_arrs["ppc"], all_charges)
    og.initialize()
    want_bulk = []
    for step in range(5):
        if step > 0:
            og.sim_step(1)
        if step % 2 == 0:
            m = og.measure()
            es, bs = m[0:3], m[3:6]
            want_bulk.append([step * 0.25, es.sum(), bs.sum(), -es[0] + es[1] + es[2] - bs[0] + bs[1] + bs[2],
                              es[0] - es[1] + es[2] + bs[0] - bs[1] + bs[2],
                              es[0] + es[1] - es[2] + bs[0] + bs[1] - bs[2], m[6], m[7]])

    info = "MV, mu: 0.500000, w: 1.200000, UVT: 3.000000, c: 0.400000, m: 0.300000"
    out = tmp_path / "output"
    t = (out / "tadpole.txt").read_text().splitlines()
    assert len(t) == 2
    for line, want in zip(t, tadpoles):
        head, v = line.split(", V: ")
        assert head == info
        assert float(v) == pytest.approx(want, abs=1e-12 * nc)
    lines = (out / "wilson" / "dipole.dat").read_text().splitlines()
    assert len(lines) == 2 * (2 + nbins)
    for k, (want_c, want_n) in enumerate(dipoles):
        block = lines[k * (2 + nbins): (k + 1) * (2 + nbins)]
        assert block[:2] == [info, "d\ttr(V V^t)"]
        for b, line in enumerate(block[2:]):
            x, v = line.split("\t")
            assert x == java_sci(b * min(dims[1], dims[2]) * 0.5 / nbins)   # bin * ds, unit spacing
            assert float(v) == pytest.approx(want_c[b], abs=1e-12 * nc)

    lines = (out / "bulk.dat").read_text().splitlines()
    assert lines[0] == "#time \t E^2 \t B^2 \t P_x \t P_y \t P_z \t G \t rho^2"
    rows = [np.array([float(x) for x in l.split("\t")]) for l in lines[1:]]
    assert len(rows) == len(want_bulk) == 3
    for got, want in zip(rows, want_bulk):
        assert got[0] == want[0]
        assert np.allclose(got[1:6], want[1:6], rtol=1e-11, atol=0)
        assert abs(got[6] - want[6]) < 1e-11 * want[7]   # Gauss violation: round-off of rho^2
        assert got[7] == pytest.approx(want[7], rel=1e-11) and want[7] > 0
