"""Parity of the gfx950 leapfrog (through the C ABI) against the FP64 oracle.

FP64 device results must agree with the oracle within 1e-12 relative (max-norm
over the lattice) after N steps; the FP32 form within the stated FP32 bounds.
"""
import numpy as np
import pytest

import oracle as O
from openpixi_amd import FieldMeasurements, FastTYMSolver

from .helpers import device_grid, make_state, oracle_grid, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


# kernel forms of the fused update: (kernel, tile, chunks)
KERNELS = [("link", None, None), ("plaq", None, None), ("cell", None, None), ("stream", 0, None), ("stream", 1, 3),
           ("stream", 4, 1), ("cell", 1, None), ("cell", 2, None)]
KERNEL_IDS = ["link", "plaq", "cell", "stream16x16", "stream8x32c3", "stream8x16c1", "cell16rows", "cell8rows"]


def run_both(nc, dims, steps, at=0.25, spacing=(1.0, 1.0, 1.0), sigma_a=0.1, sigma_e=0.05, seed=0, J=None,
             precision="fp64", active=None, kernel=("auto", None, None)):
    U, E = make_state(nc, dims, sigma_a, sigma_e, seed)
    og = oracle_grid(nc, dims, U, E, at, spacing, J=J)
    dg = device_grid(nc, dims, U, E, at, spacing, precision=precision, J=J)
    dg.setKernel(*kernel)
    if active is not None:
        og.set_active_region(*active)
        dg.setActiveRegion(*active)
    solver = FastTYMSolver()
    # Simulation.initialize: U(-dt/2) -> Unext = U(+dt/2)
    og.step_links(at)
    solver.stepLinks(dg, at)
    for _ in range(steps):
        og.store_fields()
        og.step(at)
        dg.storeFields()
        solver.step(dg, at)
    return og, dg


@pytest.mark.parametrize("kernel", KERNELS, ids=KERNEL_IDS)
@pytest.mark.parametrize("nc,dims,steps", [(2, (32, 32, 32), 100), (2, (5, 7, 6), 60), (3, (6, 6, 6), 40),
                                           (3, (4, 5, 7), 30), (3, (12, 20, 17), 10)])
def test_leapfrog_fp64_parity(gpu, nc, dims, steps, kernel):
    print(f"leapfrog parity nc={nc} dims={dims} steps={steps} kernel={kernel}", flush=True)
    og, dg = run_both(nc, dims, steps, kernel=kernel)
    assert rel_err(dg.getU(), og.U) < FP64_TOL
    assert rel_err(dg.getUnext(), og.Unext) < FP64_TOL
    assert rel_err(dg.getE(), og.E) < FP64_TOL


@pytest.mark.parametrize("kernel", KERNELS[:3], ids=KERNEL_IDS[:3])
def test_su2_32cube_plain_method(gpu, kernel):
    """SU(2) 32^3 FP64: links exp(i A_a sigma_a/2) with A_a ~ N(0, 0.1), E = 0, dt = 0.25,
    100 steps; final links, E, total energy and Gauss residual against the oracle."""
    print(f"su2 32^3 plain method kernel={kernel}", flush=True)
    og, dg = run_both(2, (32, 32, 32), 100, sigma_a=0.1, sigma_e=0.0, seed=3, kernel=kernel)
    assert rel_err(dg.getU(), og.U) < FP64_TOL
    assert rel_err(dg.getUnext(), og.Unext) < FP64_TOL
    assert rel_err(dg.getE(), og.E) < FP64_TOL
    md = FieldMeasurements().measure(dg)
    mo = og.measure()
    energy_d, energy_o = md[:6].sum(), mo[:6].sum()
    assert abs(energy_d - energy_o) < 1e-12 * abs(energy_o)
    assert md[6] < 1e-26 and mo[6] < 1e-26   # E starts at 0: Gauss holds to rounding through 100 steps
    assert abs(md[6] - mo[6]) < 1e-26


@pytest.mark.parametrize("kernel", KERNELS[:3], ids=KERNEL_IDS[:3])
@pytest.mark.parametrize("nc", [2, 3])
def test_leapfrog_strong_fields_anisotropic(gpu, nc, kernel):
    """Large amplitudes (SU(3) exp beyond the Taylor branch) and uneven spacings."""
    print(f"strong fields nc={nc} kernel={kernel}", flush=True)
    og, dg = run_both(nc, (4, 6, 5), 10, at=0.1, spacing=(1.0, 0.8, 1.3), sigma_a=1.0, sigma_e=3.0, seed=7,
                      kernel=kernel)
    assert rel_err(dg.getU(), og.U) < 1e-11
    assert rel_err(dg.getE(), og.E) < 1e-11


@pytest.mark.parametrize("kernel", KERNELS[:3], ids=KERNEL_IDS[:3])
@pytest.mark.parametrize("nc", [2, 3])
def test_current_and_active_region(gpu, nc, kernel):
    dims = (6, 5, 4)
    n = int(np.prod(dims))
    rng = np.random.default_rng(3)
    J = rng.normal(0, 0.1, size=(n, 3, O.asize(nc)))
    if nc == 3:
        from openpixi_amd.inputs import su3_coeffs_to_storage

        J = su3_coeffs_to_storage(rng.normal(0, 0.1, size=(n, 3, 8)))
    og, dg = run_both(nc, dims, 12, J=J, active=((1, 0, 1), (4, 4, 2)), seed=5, kernel=kernel)
    assert rel_err(dg.getU(), og.U) < FP64_TOL
    assert rel_err(dg.getUnext(), og.Unext) < FP64_TOL
    assert rel_err(dg.getE(), og.E) < FP64_TOL


@pytest.mark.parametrize("nc", [2, 3])
def test_step_links_parity(gpu, nc):
    dims = (4, 4, 6)
    U, E = make_state(nc, dims, 0.3, 0.2, seed=11)
    og = oracle_grid(nc, dims, U, E)
    dg = device_grid(nc, dims, U, E)
    og.step_links(0.25)
    FastTYMSolver().stepLinks(dg, 0.25)
    assert rel_err(dg.getUnext(), og.Unext) < FP64_TOL


def test_step_links_fp32_exponential_branches(gpu):
    """FP32 SU(3) exponential on both sides of its closed-form range: the field is drawn so that
    tr(Q^2) of Q = -dt E spans 0.03 to 1.7 (the closed form below 0.25, the series above,
    scaling and squaring above 1). The new links agree with the FP64 oracle's getLink (Taylor-15 /
    spectral) to FP32 rounding."""
    dims = (4, 6, 8)
    U, E = make_state(3, dims, 0.3, 1.4, seed=41)
    og = oracle_grid(3, dims, U, E)
    dg = device_grid(3, dims, U, E, precision="fp32")
    coeffs = np.array([[O.alg_coeffs(3, E[i, d]) for d in range(3)] for i in range(E.shape[0])])
    trq2 = 0.25 ** 2 / 2 * np.sum(coeffs ** 2, axis=2)
    assert trq2.min() < 0.1 and trq2.max() > 1.0 and np.mean(trq2 <= 0.25) > 0.1
    og.step_links(0.25)
    FastTYMSolver().stepLinks(dg, 0.25)
    assert rel_err(dg.getUnext(), og.Unext) < 2e-6


@pytest.mark.parametrize("kernel", ["link", "cell", "plaq", "tile"])
def test_fp32_su2_exponential_branches(gpu, kernel):
    """FP32 SU(2) exponential on both sides of its series range: m^2 = (dt |E| / 2)^2 spans
    0.005 to 2.3 (the Taylor series up to 0.25, sine and cosine above). The links of stepLinks
    and of two fused steps agree with the FP64 oracle's getLink (cos m, sin m / m) to FP32
    rounding."""
    dims = (4, 6, 8)
    U, E = make_state(2, dims, 0.3, 3.0, seed=41)
    coeffs = np.array([[O.alg_coeffs(2, E[i, d]) for d in range(3)] for i in range(E.shape[0])])
    m2 = 0.25 ** 2 / 4 * np.sum(coeffs ** 2, axis=2)
    assert m2.min() < 0.05 and m2.max() > 1.0 and 0.2 < np.mean(m2 <= 0.25) < 0.8
    og, dg = run_both(2, dims, 0, sigma_a=0.3, sigma_e=3.0, seed=41, precision="fp32", kernel=(kernel, None, None))
    assert rel_err(dg.getUnext(), og.Unext) < 2e-6
    solver = FastTYMSolver()
    for _ in range(2):
        og.store_fields()
        og.step(0.25)
        dg.storeFields()
        solver.step(dg, 0.25)
    assert rel_err(dg.getUnext(), og.Unext) < 1e-5
    assert rel_err(dg.getE(), og.E) < 1e-5


@pytest.mark.parametrize("nc", [2, 3])
def test_measurements_parity(gpu, nc):
    dims = (6, 5, 7)
    og, dg = run_both(nc, dims, 5, sigma_a=0.3, sigma_e=0.2, seed=2)
    m_o = og.measure()
    m_d = FieldMeasurements().measure(dg)
    for k in range(7):
        assert abs(m_d[k] - m_o[k]) <= 1e-12 * max(abs(m_o[k]), 1e-30), (k, m_d[k], m_o[k])
    # evaluation region
    og.set_evaluation_region((1, 1, 1), (4, 3, 5))
    dg.setEvaluationRegion((1, 1, 1), (4, 3, 5))
    m_o = og.measure()
    m_d = FieldMeasurements().measure(dg)
    for k in range(7):
        assert abs(m_d[k] - m_o[k]) <= 1e-12 * max(abs(m_o[k]), 1e-30), (k, m_d[k], m_o[k])


@pytest.mark.parametrize("nc", [2, 3])
def test_projected_energy_parity(gpu, nc):
    dims = (6, 4, 5)
    og, dg = run_both(nc, dims, 3, sigma_a=0.3, sigma_e=0.2, seed=4)
    for d in range(3):
        ref = og.projected_energy_density(d)
        got = FieldMeasurements().projectedEnergyDensity(dg, d)
        for a, b in zip(got, ref):
            assert rel_err(a, b) < 1e-12


@pytest.mark.parametrize("kernel", KERNELS[:3] + [("stream", 2, None)], ids=KERNEL_IDS[:3] + ["stream16x32"])
@pytest.mark.parametrize("nc", [2, 3])
def test_fp32_form_tracks_fp64(gpu, nc, kernel):
    """FP32 links/E with FP64 reunitarization every 10 steps stay within FP32
    rounding of the FP64 oracle, and Gauss's law stays conserved."""
    dims = (8, 8, 8)
    steps = 50
    U, E = make_state(nc, dims, 0.1, 0.05, seed=1)
    og = oracle_grid(nc, dims, U, E)
    dg = device_grid(nc, dims, U, E, precision="fp32")
    dg.setKernel(*kernel)
    g0 = og.measure()[6]
    from openpixi_amd import _lib

    L = _lib.load()
    og.step_links(0.25)
    FastTYMSolver().stepLinks(dg, 0.25)
    for s in range(steps):
        og.store_fields()
        og.step(0.25)
    dg._check(L.opx_leapfrog(dg.handle, steps, 10))
    assert rel_err(dg.getU(), og.U) < 2e-5
    assert rel_err(dg.getE(), og.E) < 2e-4
    m_d = FieldMeasurements().measure(dg)
    m_o = og.measure()
    assert abs(m_d[6] - m_o[6]) <= 1e-4 * abs(g0) + 1e-12
    e_o = m_o[0:6].sum()
    assert abs(m_d[0:6].sum() - e_o) <= 1e-5 * abs(e_o)


@pytest.mark.parametrize("kernel", [("cell", 0, None), ("cell", 1, None), ("link", None, None)],
                         ids=["cell32rows", "cell16rows", "link"])
@pytest.mark.parametrize("nc", [2, 3])
def test_fp32_anisotropic_spacing_with_current(gpu, nc, kernel):
    """Uneven spacings (the float cell form keeps each staple sum as s_j1 + (a_j1/a_j2)^2 s_j2 and
    puts at / a_j1^2 into the E update) and a current, against the FP64 oracle."""
    dims = (6, 40, 9)
    n = int(np.prod(dims))
    spacing = (1.0, 0.8, 1.3)
    at, steps = 0.1, 20
    print(f"fp32 anisotropic nc={nc} kernel={kernel}", flush=True)
    U, E = make_state(nc, dims, 0.2, 0.5, seed=19)
    rng = np.random.default_rng(19)
    if nc == 3:
        from openpixi_amd.inputs import su3_coeffs_to_storage

        J = su3_coeffs_to_storage(rng.normal(0, 0.1, size=(n, 3, 8)))
    else:
        J = rng.normal(0, 0.1, size=(n, 3, O.asize(nc)))
    og = oracle_grid(nc, dims, U, E, at, spacing, J=J)
    dg = device_grid(nc, dims, U, E, at, spacing, precision="fp32", J=J)
    dg.setKernel(*kernel)
    from openpixi_amd import _lib

    L = _lib.load()
    og.step_links(at)
    FastTYMSolver().stepLinks(dg, at)
    for _ in range(steps):
        og.store_fields()
        og.step(at)
    dg._check(L.opx_leapfrog(dg.handle, steps, 10))
    assert rel_err(dg.getU(), og.U) < 2e-5
    assert rel_err(dg.getE(), og.E) < 2e-5
    assert rel_err(dg.getUnext(), og.Unext) < 2e-5


@pytest.mark.parametrize("dims,active,with_j", [((4, 3, 64), None, False), ((3, 2, 128), None, True),
                                                ((5, 3, 64), ((1, 1, 0), (3, 2, 63)), True),
                                                ((3, 2, 128), ((0, 0, 1), (2, 1, 100)), False)],
                         ids=["row64", "row128-J", "box-full-x2-J", "box-part-x2"])
def test_fp32_cell_form_on_64_cell_rows(gpu, dims, active, with_j):
    """FP32 SU(3) cell form on x2 rows a multiple of 64 long, where each wave holds 64 cells of one
    row (a row of 64 wraps onto its own wave): against the FP64 oracle within FP32 rounding and
    against the FP32 per-link form, with an active region spanning x2 (whole waves skip) or not,
    with and without a current."""
    nc, steps, at = 3, 12, 0.25
    print(f"fp32 cell rows of 64 dims={dims} active={active} J={with_j}", flush=True)
    n = int(np.prod(dims))
    U, E = make_state(nc, dims, 0.1, 0.05, seed=21)
    J = None
    if with_j:
        from openpixi_amd.inputs import su3_coeffs_to_storage

        J = su3_coeffs_to_storage(np.random.default_rng(4).normal(0, 0.05, size=(n, 3, 8)))
    og = oracle_grid(nc, dims, U, E, at, J=J)
    grids = []
    for kernel in ("cell", "link"):
        dg = device_grid(nc, dims, U, E, at, precision="fp32", J=J)
        dg.setKernel(kernel)
        if active is not None:
            dg.setActiveRegion(*active)
        FastTYMSolver().stepLinks(dg, at)
        for _ in range(steps):
            dg.storeFields()
            FastTYMSolver().step(dg, at)
        grids.append(dg)
    if active is not None:
        og.set_active_region(*active)
    og.step_links(at)
    for _ in range(steps):
        og.store_fields()
        og.step(at)
    cell, link = grids
    assert rel_err(cell.getU(), og.U) < 2e-5 and rel_err(cell.getUnext(), og.Unext) < 2e-5
    assert rel_err(cell.getE(), og.E) < 2e-4
    assert rel_err(cell.getUnext(), link.getUnext()) < 1e-5
    assert rel_err(cell.getE(), link.getE()) < 1e-4


# Row tiles of the cell form (32, 16 and 8 x1 rows per tile, swept along x0): lattices whose x1
# extent spans several tiles with a partial last one (80 = 2 x 32 + 16, 40 = 32 + 8 = 2 x 16 + 8,
# 45 = 32 + 13 = 2 x 16 + 13 = 5 x 8 + 5), and x2 rows of 70 (waves straddle rows), 128 (two waves
# per row) and 33 cells.
CELL_TILE_LATTICES = [((5, 80, 70), 6), ((4, 40, 128), 6), ((3, 45, 33), 8)]
CELL_TILE_IDS = ["5x80x70", "4x40x128", "3x45x33"]


def su3_current(n, seed, sigma=0.05):
    from openpixi_amd.inputs import su3_coeffs_to_storage

    return su3_coeffs_to_storage(np.random.default_rng(seed).normal(0, sigma, size=(n, 3, 8)))


def cell_tile_region(dims):
    return (1, 3, 2), (dims[0] - 2, dims[1] - 4, dims[2] - 3)


@pytest.mark.parametrize("with_region", [False, True], ids=["all", "J-region"])
@pytest.mark.parametrize("tile", [0, 1, 2], ids=["rows32", "rows16", "rows8"])
@pytest.mark.parametrize("dims,steps", CELL_TILE_LATTICES, ids=CELL_TILE_IDS)
def test_cell_form_row_tiles_fp64(gpu, dims, steps, tile, with_region):
    """SU(3) FP64 cell form with every row tile on lattices of several tiles and a partial last
    tile, bare and with a current plus an active region: within 1e-12 of the oracle."""
    print(f"cell row tiles fp64 dims={dims} tile={tile} region={with_region}", flush=True)
    n = int(np.prod(dims))
    J = su3_current(n, 17) if with_region else None
    active = cell_tile_region(dims) if with_region else None
    og, dg = run_both(3, dims, steps, J=J, active=active, seed=19, kernel=("cell", tile, None))
    assert dg.kernelConfig()["kernel"] == "cell" and dg.kernelConfig()["tile"] == tile
    assert rel_err(dg.getU(), og.U) < FP64_TOL
    assert rel_err(dg.getUnext(), og.Unext) < FP64_TOL
    assert rel_err(dg.getE(), og.E) < FP64_TOL


@pytest.mark.parametrize("with_region", [False, True], ids=["all", "J-region"])
@pytest.mark.parametrize("tile", [0, 1, 2], ids=["rows32", "rows16", "rows8"])
@pytest.mark.parametrize("dims,steps", CELL_TILE_LATTICES, ids=CELL_TILE_IDS)
def test_cell_form_row_tiles_fp32(gpu, dims, steps, tile, with_region):
    """SU(3) FP32 cell form (12-real links, two-row products) with every row tile, through
    opx_leapfrog with FP64 reunitarization every 5 steps: against the FP64 oracle within FP32
    rounding (links 2e-5, E 2e-4 of the largest entry) and against the FP32 per-link form
    (1e-5, 1e-4)."""
    from openpixi_amd import _lib

    print(f"cell row tiles fp32 dims={dims} tile={tile} region={with_region}", flush=True)
    nc, at = 3, 0.25
    n = int(np.prod(dims))
    U, E = make_state(nc, dims, 0.1, 0.05, seed=23)
    J = su3_current(n, 29) if with_region else None
    active = cell_tile_region(dims) if with_region else None
    og = oracle_grid(nc, dims, U, E, at, J=J)
    if active is not None:
        og.set_active_region(*active)
    og.step_links(at)
    for _ in range(steps):
        og.store_fields()
        og.step(at)
    L = _lib.load()
    grids = []
    for kernel in (("cell", tile, None), ("link", None, None)):
        dg = device_grid(nc, dims, U, E, at, precision="fp32", J=J)
        dg.setKernel(*kernel)
        if active is not None:
            dg.setActiveRegion(*active)
        FastTYMSolver().stepLinks(dg, at)
        dg._check(L.opx_leapfrog(dg.handle, steps, 5))
        grids.append(dg)
    cell, link = grids
    assert rel_err(cell.getU(), og.U) < 2e-5 and rel_err(cell.getUnext(), og.Unext) < 2e-5
    assert rel_err(cell.getE(), og.E) < 2e-4
    assert rel_err(cell.getUnext(), link.getUnext()) < 1e-5
    assert rel_err(cell.getE(), link.getE()) < 1e-4


@pytest.mark.parametrize("nc", [2, 3])
def test_cell_form_fused_reunitarization(gpu, nc):
    """opx_leapfrog with the cell form reprojects the new links inside the update kernel on
    every 10th step. The result equals the kernel without it followed by the separate
    reprojection pass (to FP32 rounding: the two are different compilations of the same
    arithmetic), the step count carries across calls, and the reprojected links are closer to the
    group as they are stored."""
    from openpixi_amd import _lib

    dims = (6, 40, 36)
    L = _lib.load()
    U, E = make_state(nc, dims, 0.3, 0.2, seed=37)

    def grid():
        dg = device_grid(nc, dims, U, E, precision="fp32")
        dg.setKernel("cell", 0)
        FastTYMSolver().stepLinks(dg, 0.25)
        return dg

    fused = grid()
    fused._check(L.opx_leapfrog(fused.handle, 4, 10))
    fused._check(L.opx_leapfrog(fused.handle, 6, 10))      # the 10th step reprojects
    plain = grid()
    plain._check(L.opx_leapfrog(plain.handle, 10, 0))
    plain.reunitarize()                                     # U and Unext reprojected in a pass
    assert rel_err(fused.getUnext(), plain.getUnext()) < 1e-6
    assert rel_err(fused.getE(), plain.getE()) < 1e-6
    assert fused.checkHealth(1e-5) == (0, 0)
    # the pass and the fused kernel apply the same double-precision projection: starting from the
    # same float links (plain's Unext, copied into a fresh grid as U), one more step of each form
    # agrees to rounding as well
    once = device_grid(nc, dims, plain.getUnext(), plain.getE(), precision="fp32")
    once.setKernel("cell", 0)
    once.setUnext(plain.getUnext())
    twice = device_grid(nc, dims, plain.getUnext(), plain.getE(), precision="fp32")
    twice.setKernel("cell", 0)
    twice.setUnext(plain.getUnext())
    once._check(L.opx_leapfrog(once.handle, 10, 10))
    twice._check(L.opx_leapfrog(twice.handle, 10, 0))
    twice.reunitarize()
    assert rel_err(once.getUnext(), twice.getUnext()) < 1e-6


def test_su3_128cube_fp32_cell_against_fp64_link(gpu):
    """The measured workload at its own size: SU(3) 128^3 built on the device from seeded
    algebra coefficients (A ~ N(0, 0.1), E ~ N(0, 0.05) projected onto Gauss's law), the
    initial half step, then 20 leapfrog steps. The FP32 cell form (32-row tiles, four full
    tiles along x1, FP64 reunitarization every 10 steps) and the FP64 per-link form give the
    same E^2 and B^2 per direction within 1e-7 relative, and both keep the Gauss residual below
    1e-10."""
    from openpixi_amd import Grid, _lib, inputs

    dims, nc, at, steps = (128, 128, 128), 3, 0.25, 20
    n = int(np.prod(dims))
    A = inputs.random_links_coefficients(n, nc, 0.1, 0)
    Ec = inputs.random_efield_coefficients(n, nc, 0.05, 0)
    L = _lib.load()
    meas = {}
    for precision, kernel in (("fp32", ("cell", 0, None)), ("fp64", ("link", None, None))):
        print(f"128^3 {precision} {kernel}", flush=True)
        g = Grid(nc, dims, 1.0, at, 1.0, precision)
        g.setLinksFromCoefficients(A)
        g.setEFromCoefficients(Ec)
        it, g0, g1 = g.projectGauss(tol=1e-10, max_iter=4000)
        assert g0 > 1e-2 and g1 < 1e-15, (it, g0, g1)
        g.setKernel(*kernel)
        FastTYMSolver().stepLinks(g, at)
        g._check(L.opx_leapfrog(g.handle, steps, 10 if precision == "fp32" else 0))
        m = FieldMeasurements().measure(g)
        assert m[6] < 1e-10, (precision, m[6])
        meas[precision] = m
        g.close()
    m32, m64 = meas["fp32"], meas["fp64"]
    for k in range(6):
        assert abs(m32[k] - m64[k]) <= 1e-7 * abs(m64[k]), (k, m32[k], m64[k])


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)], ids=["fp64", "fp32"])
def test_su3_128cube_cell_form_fields_against_per_link_form(gpu, precision, tol):
    """The benchmark's kernel at its own launch shape: SU(3) 128^3 (the bench's seeded inputs,
    Gauss-projected E, the initial half step), 20 leapfrog steps of the cell form with 16-row
    tiles (eight tiles along x1, the autotuner's pick on MI355X) against the per-link form, every
    entry of E and Unext compared (relative to the largest): 1e-12 in FP64, 1e-5 in FP32 (FP64
    reunitarization every 10 steps in both)."""
    from openpixi_amd import Grid, _lib, inputs

    dims, nc, at, steps = (128, 128, 128), 3, 0.25, 20
    n = int(np.prod(dims))
    A = inputs.random_links_coefficients(n, nc, 0.1, 0)
    Ec = inputs.random_efield_coefficients(n, nc, 0.05, 0)
    L = _lib.load()
    fields = []
    for kernel in (("cell", 1, None), ("link", None, None)):
        print(f"128^3 fields {precision} {kernel}", flush=True)
        g = Grid(nc, dims, 1.0, at, 1.0, precision)
        g.setLinksFromCoefficients(A)
        g.setEFromCoefficients(Ec)
        g.projectGauss(tol=1e-10, max_iter=4000)
        g.setKernel(*kernel)
        FastTYMSolver().stepLinks(g, at)
        g._check(L.opx_leapfrog(g.handle, steps, 10 if precision == "fp32" else 0))
        assert g.kernelConfig()["kernel"] == kernel[0]
        fields.append((g.getE(), g.getUnext()))
        g.close()
    (e_cell, u_cell), (e_link, u_link) = fields
    assert rel_err(e_cell, e_link) < tol
    assert rel_err(u_cell, u_link) < tol


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("nc", [2, 3])
def test_fields_from_coefficients(gpu, nc, precision):
    """opx_set_links_from_coefficients / opx_set_efield_from_coefficients (the device builds
    U = exp(i A_a t_a) and stores E coefficients) against the oracle's getLink and algebra storage:
    1e-13 in FP64 (float rounding in FP32), including coefficients with tr A^2 > 1 (the SU(3)
    exponential's scaled branch), for U and Unext."""
    dims = (5, 4, 6)
    n = int(np.prod(dims))
    ad = nc * nc - 1
    rng = np.random.default_rng(31 + nc)
    A = rng.normal(0, 0.1, size=(n, 3, ad))
    A[: n // 2] *= 30.0            # half the links far from the identity (tr A^2 up to ~100)
    Ec = rng.normal(0, 0.5, size=(n, 3, ad))
    from openpixi_amd import Grid
    from openpixi_amd._lib import FIELD_UNEXT

    big = np.einsum("ndk,ndk->nd", A, A) / 2
    assert big.max() > 1.0 and big.min() < 1.0
    g = Grid(nc, dims, 1.0, 0.25, 1.0, precision)
    g.setLinksFromCoefficients(A)
    g.setLinksFromCoefficients(-A, field=FIELD_UNEXT)
    g.setEFromCoefficients(Ec)
    tol_u, tol_e = (1e-13, 1e-15) if precision == "fp64" else (2e-6, 1e-7)
    assert rel_err(g.getU(), O.links_from_coeffs(nc, A)) < tol_u
    assert rel_err(g.getUnext(), O.links_from_coeffs(nc, -A)) < tol_u
    assert rel_err(g.getE(), O.algs_from_coeffs(nc, Ec)) < tol_e


@pytest.mark.parametrize("nc", [2, 3])
def test_reunitarize_and_health(gpu, nc):
    dims = (4, 4, 4)
    U, E = make_state(nc, dims, 0.5, 0.0, seed=9)
    dg = device_grid(nc, dims, U, E)
    nf, nu = dg.checkHealth(1e-12)
    assert nf == 0 and nu == 0
    # unitary input: reunitarization moves it by rounding only
    dg.reunitarize()
    assert rel_err(dg.getU(), U) < 1e-14
    # a scaled (non-unitary) link is flagged and then repaired
    bad = U.copy()
    bad[5, 1] *= 1.01
    dg.setU(bad)
    nf, nu = dg.checkHealth(1e-6)
    assert nf == 0 and nu >= 1
    dg.reunitarize()
    nf, nu = dg.checkHealth(1e-12)
    assert nu == 0
    # a NaN is flagged
    bad[3, 0, 0] = np.nan
    dg.setU(bad)
    nf, nu = dg.checkHealth(1e-6)
    assert nf >= 1


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("nc", [2, 3])
def test_autotune_keeps_state_and_parity(gpu, nc, precision):
    """The autotuner runs on scratch buffers: the state is unchanged and the chosen kernel is exact."""
    dims = (16, 12, 20)
    print(f"autotune nc={nc} {precision}", flush=True)
    U, E = make_state(nc, dims, 0.1, 0.05, seed=6)
    og = oracle_grid(nc, dims, U, E)
    dg = device_grid(nc, dims, U, E, precision=precision)
    ms = dg.autotune(reps=2)
    assert ms > 0
    tol = FP64_TOL if precision == "fp64" else 2e-5
    assert rel_err(dg.getU(), og.U) < tol
    assert rel_err(dg.getE(), og.E) < (FP64_TOL if precision == "fp64" else 1e-6)
    og.step_links(0.25)
    FastTYMSolver().stepLinks(dg, 0.25)
    for _ in range(5):
        og.store_fields()
        og.step(0.25)
        dg.storeFields()
        FastTYMSolver().step(dg, 0.25)
    assert rel_err(dg.getU(), og.U) < tol
    print("chosen", dg.kernelConfig(), flush=True)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("nc", [2, 3])
def test_gauss_projection(gpu, nc, precision):
    """E' = E - D lambda reaches Gauss's law (checked by the oracle's own Gauss
    constraint), is the nearest such field (E - E' is orthogonal to E'), and the
    leapfrog then keeps |G|^2 at rounding level."""
    dims = (8, 6, 10)
    print(f"gauss projection nc={nc} {precision}", flush=True)
    U, E = make_state(nc, dims, 0.3, 0.2, seed=13)
    dg = device_grid(nc, dims, U, E, precision=precision)
    it, g0, g1 = dg.projectGauss(tol=1e-12, max_iter=3000)
    assert g0 > 1e-3
    assert g1 < (1e-22 if precision == "fp64" else 1e-12) * g0, (it, g0, g1)
    Ep = dg.getE()
    og = oracle_grid(nc, dims, dg.getU(), Ep)
    gauss_oracle = og.measure()[6]
    assert gauss_oracle < (1e-22 if precision == "fp64" else 1e-12) * g0
    c = lambda X: np.array([[O.alg_coeffs(nc, X[i, d]) for d in range(3)] for i in range(X.shape[0])])  # noqa: E731
    ce, cp = c(E), c(Ep)
    assert abs(np.sum((ce - cp) * cp)) < (1e-10 if precision == "fp64" else 1e-4) * np.sum(cp * cp)
    assert np.sum(cp * cp) < np.sum(ce * ce)
    FastTYMSolver().stepLinks(dg, 0.25)
    from openpixi_amd import _lib

    dg._check(_lib.load().opx_leapfrog(dg.handle, 20, 10 if precision == "fp32" else 0))
    g_end = FieldMeasurements().measure(dg)[6]
    assert g_end < (1e-20 if precision == "fp64" else 1e-10) * g0


@pytest.mark.parametrize("policy", [0, 1], ids=["nontemporal", "cached"])
@pytest.mark.parametrize("chunks", [None, 1, 3])
@pytest.mark.parametrize("dims,steps", [((6, 5, 7), 20), ((20, 33, 17), 8), ((3, 40, 70), 6), ((1, 18, 20), 5),
                                        ((32, 32, 32), 10)])
def test_tile_form_su2_parity(gpu, dims, steps, chunks, policy):
    """SU(2) LDS-tile form: lattices smaller than a tile, not multiples of it, a single
    x0 plane, and several x0 chunkings, both store policies, against the oracle (FP64, 1e-12)."""
    print(f"tile parity dims={dims} steps={steps} chunks={chunks} policy={policy}", flush=True)
    og, dg = run_both(2, dims, steps, sigma_a=0.3, sigma_e=0.2, seed=12, kernel=("tile", policy, chunks))
    assert rel_err(dg.getU(), og.U) < FP64_TOL
    assert rel_err(dg.getUnext(), og.Unext) < FP64_TOL
    assert rel_err(dg.getE(), og.E) < FP64_TOL


def test_tile_form_current_active_region_fp32(gpu):
    dims = (9, 18, 21)
    n = int(np.prod(dims))
    J = np.random.default_rng(4).normal(0, 0.1, size=(n, 3, 3))
    og, dg = run_both(2, dims, 12, J=J, active=((1, 2, 0), (7, 15, 19)), seed=6, kernel=("tile", None, 2))
    assert rel_err(dg.getU(), og.U) < FP64_TOL
    assert rel_err(dg.getE(), og.E) < FP64_TOL
    U, E = make_state(2, (16, 16, 16), 0.1, 0.05, seed=2)
    og = oracle_grid(2, (16, 16, 16), U, E)
    dg = device_grid(2, (16, 16, 16), U, E, precision="fp32")
    dg.setKernel("tile")
    from openpixi_amd import _lib

    og.step_links(0.25)
    FastTYMSolver().stepLinks(dg, 0.25)
    for _ in range(40):
        og.store_fields()
        og.step(0.25)
    dg._check(_lib.load().opx_leapfrog(dg.handle, 40, 10))
    assert rel_err(dg.getU(), og.U) < 2e-5
    assert rel_err(dg.getE(), og.E) < 2e-4


def test_tile_form_rejected_for_su3(gpu):
    from openpixi_amd import OpxError

    U, E = make_state(3, (4, 4, 4), 0.1, 0.0, seed=1)
    dg = device_grid(3, (4, 4, 4), U, E)
    with pytest.raises(OpxError):
        dg.setKernel("tile")


@pytest.mark.parametrize("shape", [0, 1, 2])
@pytest.mark.parametrize("nc,dims,steps", [(2, (6, 5, 7), 20), (2, (3, 40, 70), 6), (2, (1, 18, 20), 5),
                                           (3, (4, 5, 7), 12), (3, (12, 20, 17), 6), (3, (2, 9, 35), 4)])
def test_plane_form_parity(gpu, nc, dims, steps, shape):
    """Plane-tile form (LDS tile of one x0 plane with halo): lattices smaller than a tile,
    not multiples of it, one or two x0 planes, every tile shape, against the oracle (FP64)."""
    print(f"plane parity nc={nc} dims={dims} steps={steps} shape={shape}", flush=True)
    og, dg = run_both(nc, dims, steps, sigma_a=0.3, sigma_e=0.2, seed=13, kernel=("plane", shape, None))
    assert rel_err(dg.getU(), og.U) < FP64_TOL
    assert rel_err(dg.getUnext(), og.Unext) < FP64_TOL
    assert rel_err(dg.getE(), og.E) < FP64_TOL


@pytest.mark.parametrize("nc", [2, 3])
def test_plane_form_current_active_region(gpu, nc):
    dims = (7, 18, 37)
    n = int(np.prod(dims))
    rng = np.random.default_rng(8)
    J = rng.normal(0, 0.1, size=(n, 3, O.asize(nc)))
    if nc == 3:
        from openpixi_amd.inputs import su3_coeffs_to_storage

        J = su3_coeffs_to_storage(rng.normal(0, 0.1, size=(n, 3, 8)))
    og, dg = run_both(nc, dims, 6, J=J, active=((1, 2, 0), (5, 15, 33)), seed=9, kernel=("plane", 0, None))
    assert rel_err(dg.getU(), og.U) < FP64_TOL
    assert rel_err(dg.getUnext(), og.Unext) < FP64_TOL
    assert rel_err(dg.getE(), og.E) < FP64_TOL


@pytest.mark.parametrize("shape", [0, 1, 2])
@pytest.mark.parametrize("nc", [2, 3])
def test_plane_form_fp32_matches_link_form(gpu, nc, shape):
    """FP32: the plane form does the per-link kernel's arithmetic in the same order, so
    after 30 steps (reunitarization every 10) it agrees with it to FP32 rounding."""
    dims = (9, 20, 40)
    U, E = make_state(nc, dims, 0.1, 0.05, seed=2)
    from openpixi_amd import _lib

    L = _lib.load()
    out = []
    for kernel in (("link", None, None), ("plane", shape, None)):
        dg = device_grid(nc, dims, U, E, precision="fp32")
        dg.setKernel(*kernel)
        FastTYMSolver().stepLinks(dg, 0.25)
        dg._check(L.opx_leapfrog(dg.handle, 30, 10))
        out.append((dg.getU(), dg.getE()))
    assert rel_err(out[1][0], out[0][0]) < 1e-5
    assert rel_err(out[1][1], out[0][1]) < 1e-4


@pytest.mark.parametrize("kernel", [("link", None, None), ("cell", 0, None), ("plaq", None, None)],
                         ids=["link", "cell", "plaq"])
@pytest.mark.parametrize("nc", [2, 3])
def test_global_pointer_variants(gpu, nc, kernel):
    """The plain-pointer variants of the fused kernels (taken for fields of 4 GiB or more), forced
    on a small lattice with a current and an active region: FP64 within 1e-12 of the oracle, and
    FP32 equal to the buffer-instruction variants (the same arithmetic, other loads): bit for bit in
    the per-link and plaquette forms; in the cell form the compiler schedules the two differently
    and fuses a few multiply-adds differently, so within float rounding (1e-6)."""
    dims = (6, 5, 7)
    n = int(np.prod(dims))
    rng = np.random.default_rng(47)
    J = rng.normal(0, 0.1, size=(n, 3, O.asize(nc)))
    if nc == 3:
        J = su3_current(n, 48, 0.1)
    active = ((1, 0, 1), (4, 4, 5))
    U, E = make_state(nc, dims, 0.1, 0.05, 49)
    og = oracle_grid(nc, dims, U, E, J=J)
    og.set_active_region(*active)
    dg = device_grid(nc, dims, U, E, J=J)
    dg.setKernel(*kernel)
    dg.setActiveRegion(*active)
    dg.setGlobalPointers(True)
    og.step_links(0.25)
    FastTYMSolver().stepLinks(dg, 0.25)
    for _ in range(8):
        og.store_fields()
        og.step(0.25)
        dg.storeFields()
        FastTYMSolver().step(dg, 0.25)
    assert rel_err(dg.getU(), og.U) < FP64_TOL and rel_err(dg.getE(), og.E) < FP64_TOL
    from openpixi_amd import _lib

    out = []
    for forced in (True, False):
        g = device_grid(nc, dims, U, E, precision="fp32", J=J)
        g.setKernel(*kernel)
        g.setActiveRegion(*active)
        g.setGlobalPointers(forced)
        FastTYMSolver().stepLinks(g, 0.25)
        g._check(_lib.load().opx_leapfrog(g.handle, 8, 0))
        out.append((g.getUnext(), g.getE()))
    if kernel[0] == "cell":
        assert rel_err(out[0][0], out[1][0]) < 1e-6 and rel_err(out[0][1], out[1][1]) < 1e-6
    else:
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
