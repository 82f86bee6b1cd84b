"""Pins the oracle's colored-particle (CPIC) functions:

* the integer trajectory of CGCSuperParticle (getCurrentNGPOffset, needsUpdate;
  physics/particles/CGCSuperParticle.java:71-105) against the continuous
  trajectory z(t) = (k + 1/2)/ppc +- t/ppc rounded to the nearest grid point,
  for both beam directions and every sub-lattice shift;
* covariant charge conservation: with the deposit and parallel transport of
  CGCSuperParticleInterpolationNGP (physics/grid/CGCSuperParticleInterpolationNGP.java:19-93)
  the Gauss constraint G(x) (which includes rho) stays constant through
  Simulation steps, the defining property of the CPIC scheme;
* |Q|^2 of each charge is conserved by the transport;
* per-particle NGP (GridFunctions.nearestGridPoint) rounds half to even as Math.rint.
"""
import numpy as np
import pytest

import oracle as O

from .helpers import make_state


@pytest.mark.parametrize("ppc", [1, 2, 4, 6, 8])
@pytest.mark.parametrize("orientation", [1, -1])
def test_trajectory_matches_continuous_motion(ppc, orientation):
    for k in range(ppc):
        for t in range(0, 4 * ppc + 9):
            z = (k + 0.5) / ppc + orientation * t / ppc
            if ppc % 2 == 0:  # positions never sit on a cell midpoint
                assert O.sp_ngp_offset(orientation, 0, 1, k, ppc, t) == int(np.rint(z)), (k, t)
            crosses = O.sp_ngp_offset(orientation, 0, 1, k, ppc, t + 1) != O.sp_ngp_offset(orientation, 0, 1, k, ppc, t)
            assert crosses == O.sp_needs_update(orientation, k, ppc, t), (k, t)


def test_offsets_scale_with_plane_size():
    for t in range(20):
        for o in (1, -1):
            assert O.sp_current_offset(o, 100, 16, 1, 4, t) == 100 + 16 * O.sp_current_offset(o, 0, 1, 1, 4, t)
            assert O.sp_ngp_offset(o, 100, 16, 3, 4, t) == 100 + 16 * O.sp_ngp_offset(o, 0, 1, 3, 4, t)


def two_nuclei(nc, dims, ppc, seed, subdivisions=2, width_cells=1):
    """+x nucleus near x0 = 2 and -x nucleus near x0 = N0 - 5, `subdivisions` blocks of
    `width_cells` longitudinal cells, one super particle per sub-lattice shift."""
    rng = np.random.default_rng(seed)
    pp = dims[1] * dims[2]
    cols = dict(orientation=[], num=[], index_offset=[], per_plane=[], shift=[], ppc=[])
    charges = []
    for o, start in ((1, 2), (-1, dims[0] - 3 - 2 * width_cells)):
        for j in range(subdivisions):
            for k in range(ppc):
                cols["orientation"].append(o)
                cols["num"].append(pp * width_cells)
                cols["index_offset"].append((start + j * width_cells) * pp)
                cols["per_plane"].append(pp)
                cols["shift"].append(k)
                cols["ppc"].append(ppc)
                charges.append(O.algs_from_coeffs(nc, rng.normal(0, 1, (pp * width_cells, nc * nc - 1)) / ppc))
    return cols, charges


def oracle_with_particles(nc, dims, ppc, seed=0, sigma_a=0.3):
    U, E = make_state(nc, dims, sigma_a, 0.0, seed=seed)
    og = O.OracleGrid(nc, dims, (1.0, 1.0, 1.0), 1.0 / ppc, 1.0)
    og.U[:] = U
    og.E[:] = E
    cols, charges = two_nuclei(nc, dims, ppc, seed)
    og.set_superparticles(cols["orientation"], cols["num"], cols["index_offset"], cols["per_plane"],
                          cols["shift"], cols["ppc"], charges)
    return og


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("ppc", [4, 8])
def test_gauss_law_conserved_with_particles(nc, ppc):
    dims = (12, 3, 2)
    og = oracle_with_particles(nc, dims, ppc, seed=1)
    og.initialize()
    G0 = np.array([O.alg_coeffs(nc, og.gauss_constraint(i)) for i in range(og.n)])
    assert np.max(np.abs(G0)) > 0.1
    assert np.max(np.abs(og.J)) > 0.1  # charges do move through the lattice
    for _ in range(3 * ppc):
        og.sim_step(1)
        G = np.array([O.alg_coeffs(nc, og.gauss_constraint(i)) for i in range(og.n)])
        assert np.max(np.abs(G - G0)) < 1e-13


@pytest.mark.parametrize("nc", [2, 3])
def test_transport_conserves_charge_norm(nc):
    dims = (10, 3, 2)
    og = oracle_with_particles(nc, dims, 4, seed=2)
    q0 = og.superparticle_charges().copy()
    n0 = np.array([O.alg_square(nc, q) for q in q0])
    og.initialize()
    og.sim_step(9)
    q1 = og.superparticle_charges()
    n1 = np.array([O.alg_square(nc, q) for q in q1])
    assert np.max(np.abs(q1 - q0)) > 1e-3  # charges were rotated
    assert np.allclose(n0, n1, rtol=1e-13, atol=1e-15)


def test_nearest_grid_point_rounds_half_to_even():
    assert O.nearest_grid_point(2.5, 1.0) == 2
    assert O.nearest_grid_point(3.5, 1.0) == 4
    assert O.nearest_grid_point(-0.5, 1.0) == 0
    assert O.nearest_grid_point(1.25, 0.5) == 2
    assert O.nearest_grid_point(0.7499999, 0.5) == 1


def particle_nuclei(nc, dims, ppc, seed, block=2):
    """Per-particle CGC nuclei laid out as LightConeNGPParticleCreator does
    (physics/initial/CGC/LightConeNGPParticleCreator.java:155-200): ppc particles per cell
    at x0 = z + (j - ppc/2)/ppc, every coordinate shifted by 1e-11 a, v = +-1 along x0;
    a +x nucleus near x0 = 2 and a -x nucleus near x0 = N0 - 4, `block` cells long."""
    rng = np.random.default_rng(seed)
    direction, pos, vel, q = [], [], [], []
    for o, z0 in ((1, 2), (-1, dims[0] - 2 - block)):
        for z in range(z0, z0 + block):
            for x1 in range(dims[1]):
                for x2 in range(dims[2]):
                    for j in range(ppc):
                        x = (1.0 * j - ppc // 2) / ppc
                        direction.append(0)
                        pos.append([z + 10e-12 + x, x1 + 10e-12, x2 + 10e-12])
                        vel.append([float(o), 0.0, 0.0])
                        q.append(O.alg_from_coeffs(nc, rng.normal(0, 1, nc * nc - 1) / ppc))
    return np.array(direction, np.int32), np.array(pos), np.array(vel), np.array(q)


def oracle_with_per_particle(nc, dims, ppc, seed=0, sigma_a=0.3):
    U, E = make_state(nc, dims, sigma_a, 0.0, seed=seed)
    og = O.OracleGrid(nc, dims, (1.0, 1.0, 1.0), 1.0 / ppc, 1.0)
    og.U[:] = U
    og.E[:] = E
    og.set_particles(*particle_nuclei(nc, dims, ppc, seed))
    return og


@pytest.mark.parametrize("nc", [2, 3])
@pytest.mark.parametrize("ppc", [1, 2, 4])
def test_per_particle_gauss_law_conserved(nc, ppc):
    """CGCParticleInterpolationNGP + CGCParticleSolver keep G(x) (with rho) constant."""
    dims = (12, 3, 2)
    og = oracle_with_per_particle(nc, dims, ppc, seed=3)
    og.initialize()
    G0 = np.array([O.alg_coeffs(nc, og.gauss_constraint(i)) for i in range(og.n)])
    assert np.max(np.abs(G0)) > 0.1
    assert np.max(np.abs(og.J)) > 0.1
    for _ in range(3 * ppc):
        og.sim_step(1)
        G = np.array([O.alg_coeffs(nc, og.gauss_constraint(i)) for i in range(og.n)])
        assert np.max(np.abs(G - G0)) < 1e-13


@pytest.mark.parametrize("nc", [2, 3])
def test_per_particle_positions_and_charges(nc):
    """Each particle advances by v dt per step (pos0 after n steps is the n-fold sum),
    |Q|^2 is conserved, and the charge is rotated only when its NGP changes. At the
    end of a step the current charge is Q1 (reassignValues swaps it into Q0 next step)."""
    dims, ppc = (10, 2, 2), 4
    d, p0, v, q = particle_nuclei(nc, dims, ppc, seed=4)
    og = O.OracleGrid(nc, dims, (1.0, 1.0, 1.0), 1.0 / ppc, 1.0)
    U, _ = make_state(nc, dims, 0.3, 0.0, seed=4)
    og.U[:] = U
    og.set_particles(d, p0, v, q)
    og.initialize()
    steps = 2   # the charges are current after steps + 1 pushes (initialize pushes once)
    og.sim_step(steps)
    pos0, pos1, _, Q1 = og.particle_state()
    expect = p0.copy()
    for _ in range(steps):
        expect = expect + v * (1.0 / ppc)
    assert np.array_equal(pos0, expect)
    assert np.array_equal(pos1, expect + v * (1.0 / ppc))
    n0 = np.array([O.alg_square(nc, x) for x in q])
    n1 = np.array([O.alg_square(nc, x) for x in Q1])
    assert np.allclose(n0, n1, rtol=1e-13, atol=1e-15)
    crossed = np.rint(pos1[:, 0]) != np.rint(p0[:, 0])
    assert crossed.any() and (~crossed).any()
    assert np.array_equal(Q1[~crossed], q[~crossed])
    assert np.min(np.max(np.abs(Q1[crossed] - q[crossed]), axis=1)) > 1e-6
