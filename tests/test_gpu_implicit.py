"""ImplicitTYMSolver on the device against the oracle: the group power (every branch of the
logarithm), single implicit steps with iterations, damping, a current and an active region,
FP32, and a Simulation of an implicit simulation type.

FP64 bars: 1e-12, and 1e-11 for SU(3) steps with damping, whose logarithm takes eigenvectors
of (U - l_j)(U - l_k) and so amplifies the rounding of its input links."""
import numpy as np
import pytest

import oracle as O
from openpixi_amd import FastTYMSolver, ImplicitTYMSolver, _lib
from openpixi_amd.simulation import Settings, Simulation

from .dense import group_to_matrix, matrix_to_group
from .helpers import device_grid, make_state, oracle_grid, rel_err

pytestmark = pytest.mark.gpu


def elements(nc, rng):
    out = []
    for sigma in (0.03, 0.3, 0.8, 1.5, 2.5):
        out += [O.links_from_coeffs(nc, rng.normal(0, sigma, (nc * nc - 1,))) for _ in range(12)]
    if nc == 3:
        for t in (0.05, 0.2, 0.4):   # doubly degenerate spectra in random bases
            V = group_to_matrix(3, O.links_from_coeffs(3, rng.normal(0, 1.0, (8,))))
            out.append(matrix_to_group(3, V @ np.diag(np.exp(1j * np.array([t, t, -2 * t]))) @ V.conj().T))
    return np.array(out)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
@pytest.mark.parametrize("nc", [2, 3])
def test_group_pow_matches_oracle(gpu, nc, precision, tol):
    """Every branch of the logarithm, powers 0 to 1.7. Left out: SU(3) powers above 1 of the
    degenerate-spectrum elements. Their logarithm is the 50-term series with a residual
    trace of its truncation size (about 1e-8), and for tr (xH)^2 > 1 the reference
    exponentiates it through its spectral branch, which assumes a traceless matrix; the
    device exponentiates the traceless part. Damping factors in [0, 1] give powers <= 1."""
    print(f"group pow nc={nc} {precision}", flush=True)
    us = elements(nc, np.random.default_rng(4))
    L = _lib.load()
    prec = _lib.OPX_FP64 if precision == "fp64" else _lib.OPX_FP32
    degenerate = np.zeros(len(us), bool)
    if nc == 3:
        degenerate[-3:] = True
    for x in (0.0, 0.3, 1.0, 1.7):
        got = np.zeros_like(us)
        assert L.opx_group_pow(nc, prec, len(us), _lib.dptr(np.ascontiguousarray(us)), x, _lib.dptr(got)) == 0
        # the FP32 form rounds the elements to float and raises them in double
        ins = us if precision == "fp64" else us.astype(np.float32).astype(np.float64)
        want = np.array([O.group_pow(nc, u, x) for u in ins])
        keep = ~degenerate if x > 1 else np.ones(len(us), bool)
        assert np.max(np.abs(got - want)[keep]) < tol, x


def implicit_pair(nc, dims, seed, precision="fp64", J=True, active=None, sigma_a=0.3):
    U, E = make_state(nc, dims, sigma_a, 0.2, seed)
    n = int(np.prod(dims))
    rng = np.random.default_rng(seed + 3)
    Un = O.links_from_coeffs(nc, rng.normal(0, sigma_a, (n, 3, nc * nc - 1)))
    Jv = O.algs_from_coeffs(nc, rng.normal(0, 0.1, (n, 3, nc * nc - 1))) if J else None
    og = oracle_grid(nc, dims, U, E, 0.25, (1.0, 0.9, 1.2), 1.3, J=Jv)
    og.Unext[:] = Un
    dg = device_grid(nc, dims, U, E, 0.25, (1.0, 0.9, 1.2), 1.3, precision=precision, J=Jv)
    dg.setUnext(Un)
    if active is not None:
        og.set_active_region(*active)
        dg.setActiveRegion(*active)
    return og, dg


@pytest.mark.parametrize("iterations,damping", [(0, 0.0), (1, 0.0), (3, 0.0), (2, 0.3), (4, 0.7), (2, 1.0)])
@pytest.mark.parametrize("nc,dims", [(2, (5, 6, 7)), (3, (4, 5, 3)), (2, (1, 9, 4))])
def test_implicit_step_matches_oracle(gpu, nc, dims, iterations, damping):
    print(f"implicit nc={nc} dims={dims} K={iterations} damping={damping}", flush=True)
    og, dg = implicit_pair(nc, dims, seed=5)
    og.implicit_step(iterations=iterations, damping=damping)
    ImplicitTYMSolver(iterations, damping).step(dg, 0.25)
    tol = 1e-11 if nc == 3 and damping > 0 else 1e-12
    assert rel_err(dg.getE(), og.E) < tol
    assert rel_err(dg.getUnext(), og.Unext) < tol
    assert rel_err(dg.getU(), og.U) < 1e-15


@pytest.mark.parametrize("nc", [2, 3])
def test_implicit_step_active_region(gpu, nc):
    og, dg = implicit_pair(nc, (6, 5, 4), seed=8, active=((1, 1, 0), (4, 3, 2)))
    og.implicit_step(iterations=3, damping=0.2)
    ImplicitTYMSolver(3, 0.2).step(dg, 0.25)
    tol = 1e-11 if nc == 3 else 1e-12
    assert rel_err(dg.getE(), og.E) < tol
    assert rel_err(dg.getUnext(), og.Unext) < tol


@pytest.mark.parametrize("damping", [0.1, 0.0, 1.0])
@pytest.mark.parametrize("nc", [2, 3])
def test_implicit_fp32(gpu, nc, damping):
    """FP32 links: the powers evaluated in double (damping 0.1), and damping 0 or 1, where the
    link is only projected onto the group in double."""
    og, dg = implicit_pair(nc, (6, 6, 6), seed=2, precision="fp32")
    og.implicit_step(iterations=3, damping=damping)
    ImplicitTYMSolver(3, damping).step(dg, 0.25)
    assert rel_err(dg.getE(), og.E) < 2e-5
    assert rel_err(dg.getUnext(), og.Unext) < 2e-6


def test_fast_solver_after_implicit(gpu):
    """The solver objects select the context's solver on every step."""
    og, dg = implicit_pair(2, (4, 4, 4), seed=1)
    ImplicitTYMSolver(2, 0.0).step(dg, 0.25)
    og.implicit_step(iterations=2)
    dg.storeFields()
    og.store_fields()
    FastTYMSolver().step(dg, 0.25)
    og.step(0.25)
    assert rel_err(dg.getE(), og.E) < 1e-12


@pytest.mark.parametrize("nc", [2, 3])
def test_implicit_simulation_type(gpu, nc):
    """A 'temporal optimized implicit cgc ngp' Simulation (no particles here) over 6 steps."""
    dims = (6, 4, 5)
    s = Settings.from_yaml(f"""
simulationType: temporal optimized implicit cgc ngp
timeStep: 0.25
duration: 1.25
numberOfColors: {nc}
gridCells: {list(dims)}
gridStep: 1
implicitSolverIterations: 2
implicitSolverDamping: 0.25
""")
    U, E = make_state(nc, dims, 0.3, 0.2, seed=6)
    sim = Simulation(s)
    assert isinstance(sim.fsolver, ImplicitTYMSolver)
    sim.grid.setU(U)
    sim.grid.setE(E)
    sim.initialize()
    sim.run()
    og = oracle_grid(nc, dims, U, E)
    og.step_links(0.25)
    for _ in range(sim.totalSimulationSteps):
        og.store_fields()
        og.implicit_step(iterations=2, damping=0.25)
    assert sim.totalSimulationSteps == 6
    assert rel_err(sim.grid.getE(), og.E) < 1e-11
    assert rel_err(sim.grid.getUnext(), og.Unext) < 1e-11
