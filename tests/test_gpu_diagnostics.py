"""Output writers on the device: the longitudinal Poynting vector against the oracle,
and the BulkQuantitiesInTime / ProjectedEnergyDensity files written by a Simulation run,
parsed and compared with the oracle's values at the same steps."""
import numpy as np
import pytest

from openpixi_amd import FieldMeasurements, FastTYMSolver, OpxError
from openpixi_amd.diagnostics import java_double_to_string, java_sci
from openpixi_amd.simulation import Settings, Simulation

from .helpers import device_grid, make_state, oracle_grid

pytestmark = pytest.mark.gpu


def stepped_pair(nc, dims, spacing, g, steps, precision="fp64", seed=4):
    U, E = make_state(nc, dims, 0.4, 0.2, seed=seed)
    og = oracle_grid(nc, dims, U, E, 0.25, spacing, g)
    dg = device_grid(nc, dims, U, E, 0.25, spacing, g, precision=precision)
    og.step_links(0.25)
    FastTYMSolver().stepLinks(dg, 0.25)
    for _ in range(steps):
        og.store_fields()
        og.step(0.25)
        dg.storeFields()
        FastTYMSolver().step(dg, 0.25)
    return og, dg


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_poynting_matches_oracle(gpu, precision, tol):
    print(f"poynting {precision}", flush=True)
    og, dg = stepped_pair(2, (6, 5, 9), (1.0, 0.8, 1.3), 1.7, 4, precision)
    for d in range(3):
        ref = og.projected_poynting(d)
        got = FieldMeasurements().projectedPoynting(dg, d)
        for a, b in zip(got, ref):
            scale = np.max(np.abs(b))
            assert scale > 0
            assert np.max(np.abs(a - b)) < tol * scale, d


def test_poynting_su3_is_rejected(gpu):
    U, E = make_state(3, (3, 3, 3), 0.1, 0.1, seed=1)
    dg = device_grid(3, (3, 3, 3), U, E)
    with pytest.raises(OpxError):
        FieldMeasurements().projectedPoynting(dg, 0)


def parse_tokens(line):
    toks = line.rstrip("\n").split("\t")
    vals = [float(t) for t in toks]
    # every token is the canonical Java rendering of the value it holds
    assert [java_sci(v) for v in vals] == toks
    return np.array(vals)


def test_simulation_writes_reference_files(gpu, tmp_path):
    print("simulation output files", flush=True)
    nc, dims, spacing, g = 2, (6, 4, 5), (1.0, 0.9, 1.1), 1.3
    bulk, ped = tmp_path / "bulk.dat", tmp_path / "ped.dat"
    s = Settings.from_yaml(f"""
simulationType: temporal yang-mills
timeStep: 0.25
duration: 1.5
numberOfColors: {nc}
couplingConstant: {g}
gridCells: {list(dims)}
gridSteps: {list(spacing)}
output:
  bulkQuantitiesInTime:
    - path: {bulk}
      interval: 0.5
  projectedEnergyDensity:
    - path: {ped}
      interval: 0.75
      direction: 1
""")
    U, E = make_state(nc, dims, 0.4, 0.2, seed=6)
    sim = Simulation(s)
    sim.grid.setU(U)
    sim.grid.setE(E)
    sim.initialize()
    sim.run()
    assert sim.totalSimulationSteps == 7   # ceil(1.5 / 0.25) + 1 steps

    og = oracle_grid(nc, dims, U, E, 0.25, spacing, g)
    og.initialize()
    want_bulk, want_ped = [], []
    for step in range(8):
        if step > 0:
            og.sim_step(1)
        if step % 2 == 0:
            m = og.measure()
            es, bs = m[0:3], m[3:6]
            want_bulk.append([step * 0.25, es.sum(), bs.sum(), -es[0] + es[1] + es[2] - bs[0] + bs[1] + bs[2],
                              es[0] - es[1] + es[2] + bs[0] - bs[1] + bs[2],
                              es[0] + es[1] - es[2] + bs[0] + bs[1] - bs[2], m[6], m[7]])
        if step % 3 == 0:
            want_ped.append((step * 0.25, og.projected_energy_density(1) + list(og.projected_poynting(1))))

    lines = bulk.read_text().splitlines()
    assert lines[0] == "#time \t E^2 \t B^2 \t P_x \t P_y \t P_z \t G \t rho^2"
    rows = [parse_tokens(l) for l in lines[1:]]
    assert len(rows) == len(want_bulk) == 4
    for got, want in zip(rows, want_bulk):
        assert got[0] == want[0]
        assert np.allclose(got[1:7], want[1:7], rtol=1e-12, atol=0)
        assert got[7] == 0.0   # no charges

    lines = ped.read_text().splitlines()
    assert len(lines) == 7 * len(want_ped) == 21
    for k, (t, arrays) in enumerate(want_ped):
        block = lines[7 * k: 7 * k + 7]
        assert block[0] == java_double_to_string(t)
        for line, want in zip(block[1:], arrays):
            got = parse_tokens(line)
            assert got.shape == (dims[1],)
            scale = max(np.max(np.abs(want)), 1e-300)
            assert np.max(np.abs(got - want)) < 1e-12 * scale
