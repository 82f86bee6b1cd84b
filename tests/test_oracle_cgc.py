"""Pins the oracle's CGC initial condition (oracle/opx_oracle_cgc.c) with checks independent
of its code:

* the Fourier regulation against numpy's FFT, and with every regulator off it is the
  removal of each x0 plane's transverse mean;
* the 2D Poisson solve satisfies the lattice Poisson equation -Laplace(phi) = rho - mean(rho);
* a Wilson-line segment: the closed form for commuting (single-colour) fields, and
  second-order convergence to the path-ordered exponential (an ODE solution) otherwise;
* the super-particle charges: the refinement only moves charge within an NGP cell, so the
  particles deposit exactly the Gauss violation of the Poisson solution, and the complete
  initialization (fields, particles, charge density) satisfies Gauss's law;
* the MV charge profile and box: charge only inside the box, zero transverse mean.
* the Wilson-line observables: tadpole and dipole against dense numpy brute force, and the
  lines they use are those of the plane behind the nucleus.

The reference draws its random charges with java.util.Random and transforms them with
JTransforms. Here the noise is an input: the package draws it with JavaRandom, pinned to
the JDK's outputs in test_cgc_host.py, and the transforms are pinned to numpy's FFT above.
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

import oracle as O

from .dense import coeffs_to_herm, group_to_matrix

RNG = np.random.default_rng(17)


def np_regulate(rho, dims, uvt, lw, ir, aT, aL):
    f = np.fft.fftn(rho.reshape(dims))
    f[0, 0, 0] = 0
    z, p1, p2 = np.meshgrid(np.arange(dims[0]), np.arange(dims[1]), np.arange(dims[2]), indexing="ij")
    kte = (4.0 - 2 * np.cos(2 * np.pi * p1 / dims[1]) - 2 * np.cos(2 * np.pi * p2 / dims[2])) / aT ** 2
    c1 = np.where(p1 < dims[1] // 2, p1, dims[1] - p1) * 2 * np.pi / (aT * dims[1])
    c2 = np.where(p2 < dims[2] // 2, p2, dims[2] - p2) * 2 * np.pi / (aT * dims[2])
    kt = c1 ** 2 + c2 ** 2
    delta = z / dims[0]
    kl = np.abs(np.where(delta < 0.5, 2 * delta * np.pi / aL, 2 * (delta - 1) * np.pi / aL))
    keep = (kt <= uvt * uvt) & (kt > 0)
    with np.errstate(invalid="ignore", divide="ignore"):
        f = np.where(keep, f * (kte / (kte + ir * ir)) * np.exp(-0.25 * kl * kl * lw * lw), 0)
    return np.real(np.fft.ifftn(f)).reshape(-1)


@pytest.mark.parametrize("dims", [(8, 4, 8), (6, 5, 3), (16, 8, 8)])
def test_regulation_matches_numpy_fft(dims):
    rho = RNG.normal(size=int(np.prod(dims)))
    for uvt, lw, ir, aT, aL in [(100.0, 0.0, 0.0, 1.0, 1.0), (2.0, 1.5, 0.3, 0.9, 1.2), (0.8, 0.4, 1.0, 1.0, 0.5)]:
        got = O.regulate_charge_gaussian(dims, uvt, lw, ir, aT, aL, rho)
        assert np.allclose(got, np_regulate(rho, dims, uvt, lw, ir, aT, aL), atol=1e-12)


def test_regulation_without_regulators_removes_transverse_means():
    dims = (6, 5, 7)
    rho = RNG.normal(size=int(np.prod(dims)))
    got = O.regulate_charge_gaussian(dims, 1e6, 0.0, 0.0, 1.0, 1.0, rho).reshape(dims)
    want = rho.reshape(dims) - rho.reshape(dims).mean(axis=(1, 2), keepdims=True)
    assert np.allclose(got, want, atol=1e-12)


@pytest.mark.parametrize("n1,n2,aT", [(8, 8, 1.0), (6, 10, 0.7), (5, 3, 1.3)])
def test_poisson_2d_solves_the_lattice_equation(n1, n2, aT):
    rho = RNG.normal(size=n1 * n2)
    phi = O.poisson_2d(n1, n2, aT, rho).reshape(n1, n2)
    lap = (np.roll(phi, 1, 0) + np.roll(phi, -1, 0) + np.roll(phi, 1, 1) + np.roll(phi, -1, 1) - 4 * phi) / aT ** 2
    assert np.allclose(-lap, rho.reshape(n1, n2) - rho.mean(), atol=1e-12)
    assert abs(phi.mean()) < 1e-13


def segment(nc, z1, z2, M, P1, P2, aL=1.0, g=1.0):
    og = O.OracleGrid(nc, (2, 2, 2), (aL, 1.0, 1.0), 0.25, g)
    W = np.zeros(O.gsize(nc))
    O.lib().orc_wilson_segment(og._g, z1, z2, M, O._ptr(np.ascontiguousarray(P1, float)),
                               O._ptr(np.ascontiguousarray(P2, float)), O._ptr(W))
    return group_to_matrix(nc, W)


@pytest.mark.parametrize("nc", [2, 3])
def test_wilson_segment_commuting_fields_closed_form(nc):
    """Single-colour fields commute: the product is exp(i H) of the summed exponent. The
    field is the linear ramp through P1 at half a cell below and P2 at half a cell above the
    integer nearest the segment's middle (a staggered grid); the midpoint rule integrates a
    linear ramp exactly, so H = -g aL |z2 - z1| times the ramp's mean over the segment."""
    from scipy.linalg import expm

    a = nc * nc - 1
    c = np.zeros(a)
    c[2] = 1.0
    p1, p2 = 0.7, -0.4
    for z1, z2, M, aL, g in [(3.0, 2.5, 16, 1.0, 1.0), (2.5, 2.0, 16, 0.8, 1.7), (4.0, 4.125, 4, 1.2, 2.0)]:
        P1, P2 = O.alg_from_coeffs(nc, p1 * c), O.alg_from_coeffs(nc, p2 * c)
        W = segment(nc, z1, z2, M, P1, P2, aL, g)
        centre = np.floor((z1 + z2) / 2 + 0.5)
        ramp = lambda z: p1 + (p2 - p1) * (z - centre + 0.5)
        mean = (ramp(z1) + ramp(z2)) / 2
        H = coeffs_to_herm(nc, -g * abs(z2 - z1) * aL * mean * c)
        assert np.allclose(W, expm(1j * H), atol=1e-13)


@pytest.mark.parametrize("nc", [2, 3])
def test_wilson_segment_converges_to_the_path_ordered_exponential(nc):
    """dW/ds = W A(s) with A the linear ramp from P1 to P2 (times -i g aL per unit z): the
    midpoint product approaches the ODE solution at second order in 1/M."""
    a = nc * nc - 1
    c1, c2 = RNG.normal(0, 0.8, a), RNG.normal(0, 0.8, a)
    P1, P2 = O.alg_from_coeffs(nc, c1), O.alg_from_coeffs(nc, c2)
    z1, z2, aL, g = 1.0, 0.5, 1.0, 1.3
    H1, H2 = coeffs_to_herm(nc, c1), coeffs_to_herm(nc, c2)
    ui = np.floor((z1 + z2) * 0.5 + 0.5)
    u1, u2 = z1 - ui, z2 - ui
    L = abs(z2 - z1) * aL

    def rhs(s, y):
        W = y.reshape(nc, nc)
        u = u1 + (u2 - u1) * s
        Pu = (H1 + H2) / 2 + (H2 - H1) * u
        return (W @ (1j * (-g * L) * Pu)).reshape(-1)

    sol = solve_ivp(rhs, (0.0, 1.0), np.eye(nc, dtype=complex).reshape(-1), rtol=1e-12, atol=1e-13)
    exact = sol.y[:, -1].reshape(nc, nc)
    errs = [np.max(np.abs(segment(nc, z1, z2, M, P1, P2, aL, g) - exact)) for M in (4, 8, 16, 32)]
    for e0, e1 in zip(errs, errs[1:]):
        assert e1 < 0.3 * e0   # ~1/4 per halving
    assert errs[-1] < 1e-3


def cgc_state(nc, dims, orientation, seed, at=0.25, spacing=(1.0, 1.0, 1.0)):
    og = O.OracleGrid(nc, dims, spacing, at, 1.0)
    noise = np.random.default_rng(seed).normal(size=(nc * nc - 1, og.n))
    rho = og.mv_charge(dims[0] * spacing[0] / 2, 2.0, 1.0, 10.0, 0.5, 0.2, noise)
    gauss = og.lightcone_solve(rho, orientation)
    return og, rho, gauss


@pytest.mark.parametrize("nc", [2, 3])
def test_mv_charge_box_and_zero_transverse_mean(nc):
    dims = (16, 4, 4)
    og = O.OracleGrid(nc, dims, (1.0, 1.0, 1.0), 0.25, 1.0)
    noise = np.random.default_rng(1).normal(size=(nc * nc - 1, og.n))
    loc = 5.0
    rho = og.mv_charge(loc, 1.5, 1.0, 10.0, 0.5, 0.2, noise).reshape(dims[0], -1, O.asize(nc))
    # box of the simulation width centred on the nucleus, clipped to the lattice: planes 0 < z < 13
    assert np.all(rho[0] == 0) and np.all(rho[13:] == 0)
    assert np.max(np.abs(rho[1:13])) > 0
    # the regulation removes every kT = 0 mode: zero transverse mean in each plane
    coeffs = np.array([[O.alg_coeffs(nc, v) for v in plane] for plane in rho])
    assert np.max(np.abs(coeffs.mean(axis=1))) < 1e-12


@pytest.mark.parametrize("nc,orientation", [(2, 1), (2, -1), (3, 1)])
def test_particles_deposit_the_gauss_violation_and_gauss_law_holds(nc, orientation):
    """Inside the particle block every NGP cell receives ppc charges of G / ppc and the
    refinement moves charge only within a cell, so the deposit is the Gauss violation and the
    initialized fields satisfy Gauss's law there. The block's first plane and the cell past its
    last plane receive half their charge (particles exist on one side only), and the
    refinement's periodic list couples exactly those two cells: the reference's accepted
    boundary error."""
    dims = (24, 4, 4)
    og, rho, gauss = cgc_state(nc, dims, orientation, seed=3)
    arrs, charges, (zs, bw) = og.lightcone_superparticles(gauss, orientation, subdivisions=3)
    assert bw > 2 and len(charges) == 3 * 4
    og.set_superparticles(arrs["orientation"], arrs["num"], arrs["index_offset"], arrs["per_plane"], arrs["shift"],
                          arrs["ppc"], charges)
    og.rho[:] = 0
    O.lib().orc_sp_interpolate_chargedensity(og._g, og._sp, 0)
    scale = np.max(np.abs(gauss))
    T = dims[1] * dims[2]
    interior = slice((zs + 1) * T, (zs + bw) * T)
    assert np.max(np.abs(og.rho[interior] - gauss[interior])) < 1e-12 * scale
    # outside the block and its two edge cells nothing is deposited
    outside = np.ones(og.n, bool)
    outside[zs * T:(zs + bw + 1) * T] = False
    assert np.all(og.rho[outside] == 0)
    # Gauss's law on the interior planes (the links, E and the deposited charge of one nucleus)
    g2 = np.array([O.lib().orc_gauss_constraint_squared(og._g, c) for c in range(og.n)])
    q2 = np.array([O.alg_square(nc, og.rho[c]) for c in range(og.n)])
    assert np.max(g2[interior]) < 1e-20 * np.max(q2)


def test_refinement_conserves_the_charge_of_each_transverse_column():
    nc, dims = 2, (20, 3, 3)
    og, rho, gauss = cgc_state(nc, dims, 1, seed=5)
    arrs, charges, (zs, bw) = og.lightcone_superparticles(gauss, 1, subdivisions=2)
    q = np.concatenate(charges)
    T = dims[1] * dims[2]
    # before refinement particle k of cell index holds G(ngp) / ppc; each column's total is
    # the sum of G over the block planes' NGP cells, unchanged by the refinement
    total = np.zeros((T, 3))
    for s_, c in enumerate(charges):
        off = arrs["index_offset"][s_]
        for i in range(arrs["num"][s_]):
            total[(off + i) % T] += c[i]
    want = np.zeros((T, 3))
    ppc = arrs["ppc"][0]
    for idx in range(zs * T, (zs + bw) * T):
        for k in range(ppc):
            ngp = idx if k < ppc // 2 else og.shift(idx, 0, 1)
            want[idx % T] += gauss[ngp] / ppc
    assert np.allclose(total, want, atol=1e-12 * np.max(np.abs(gauss)))
    assert q.shape[0] == bw * T * ppc


@pytest.mark.parametrize("nc", [2, 3])
def test_wilson_observables_match_dense_brute_force(nc):
    """Tadpole: the mean real trace; dipole: tr(V_x V_y^+) over all ordered pairs binned by the
    periodic transverse distance (a coordinate difference above half the extent wraps), both
    from dense matrices with numpy."""
    n1, n2, a1, a2 = 7, 6, 0.9, 1.1
    T = n1 * n2
    VT = O.links_from_coeffs(nc, RNG.normal(0, 0.8, (T, nc * nc - 1)))
    M = np.array([group_to_matrix(nc, v) for v in VT])
    assert O.wilson_tadpole(nc, VT) == pytest.approx(np.mean(np.trace(M, axis1=1, axis2=2).real), abs=1e-14)
    nbins = int(np.sqrt(T)) // 2
    ds = min(n1 * a1, n2 * a2) * 0.5 / nbins
    x = np.stack([np.arange(T) // n2, np.arange(T) % n2], axis=1)
    d = np.abs(x[:, None, :] - x[None, :, :]).astype(float)
    half = np.array([n1 // 2, n2 // 2])
    d = np.where(d > half, d - np.array([n1, n2]), d)
    dist = np.sqrt((d[..., 0] * a1) ** 2 + (d[..., 1] * a2) ** 2)
    bins = (dist / ds).astype(int)
    tr = np.einsum("iab,jab->ij", M, M.conj()).real
    corr, counts = O.wilson_dipole(nc, n1, n2, a1, a2, VT, nbins)
    for b in range(nbins):
        sel = bins == b
        assert counts[b] == sel.sum()
        assert corr[b] == pytest.approx(tr[sel].mean(), abs=1e-13)
    # the nearest bin holds the x = y pairs, tr(V V^+) = N
    assert counts[0] >= T


@pytest.mark.parametrize("orientation", [1, -1])
def test_wilson_lines_kept_from_the_plane_behind_the_nucleus(orientation):
    """A charge on planes 8-10 of 20: the lines start as identities on the plane the path begins
    at and pick the charge up on their way, so the plane behind the nucleus (x0 = 0 for +1,
    N0 - 1 for -1) holds non-trivial lines, and a zero charge leaves identities."""
    dims = (20, 4, 4)
    og = O.OracleGrid(2, dims, (1.0, 1.0, 1.0), 0.25, 1.0)
    T = dims[1] * dims[2]
    rho = np.zeros((og.n, 3))
    rho[8 * T:11 * T] = RNG.normal(0, 1.0, (3 * T, 3))
    og.lightcone_solve(rho, orientation)
    assert np.max(np.abs(og.VT - np.array([1.0, 0, 0, 0]))) > 1e-2
    og2 = O.OracleGrid(2, dims, (1.0, 1.0, 1.0), 0.25, 1.0)
    og2.lightcone_solve(np.zeros((og2.n, 3)), orientation)
    assert np.max(np.abs(og2.VT - np.array([1.0, 0, 0, 0]))) == 0.0


@pytest.mark.parametrize("nc,orientation", [(2, 1), (2, -1), (3, 1)])
def test_per_particle_creator_deposits_the_gauss_violation(nc, orientation):
    """LightConeNGPParticleCreator: every particle's nearest grid point is its own cell, ppc of
    them per cell of the planes z_start..z_end, and the refinement (which never moves the last
    particle of a cell) keeps each cell's total, so the particles deposit exactly the Gauss
    violation on those planes and the initialized fields satisfy Gauss's law there."""
    dims = (24, 4, 4)
    og, rho, gauss = cgc_state(nc, dims, orientation, seed=7)
    d, p0, p1, v, q, (zs, ze) = og.lightcone_particles(gauss, orientation)
    T = dims[1] * dims[2]
    ppc = int(round(1.0 / 0.25))
    assert len(q) == (ze - zs + 1) * T * ppc and ze > zs
    # layout: (z, transverse cell, j) order, NGP = own cell, v = orientation along x0
    ngp = np.rint(p0).astype(int)
    cells = (ngp[:, 0] * dims[1] + ngp[:, 1]) * dims[2] + ngp[:, 2]
    assert np.array_equal(cells, np.repeat(np.arange(zs * T, (ze + 1) * T), ppc))
    assert np.array_equal(v, np.tile([orientation, 0.0, 0.0], (len(q), 1)))
    assert np.allclose(p1[:, 0] - p0[:, 0], 0.25 * orientation, atol=1e-12)
    sub = p0[:ppc, 0] - zs
    assert np.allclose(np.sort(sub), (np.arange(ppc) - ppc // 2) / ppc, atol=1e-10)
    # the deposit is the Gauss violation on the block planes, nothing elsewhere
    dep = np.zeros_like(gauss)
    np.add.at(dep, cells, q)
    scale = np.max(np.abs(gauss))
    assert np.max(np.abs(dep[zs * T:(ze + 1) * T] - gauss[zs * T:(ze + 1) * T])) < 1e-12 * scale
    assert np.all(dep[:zs * T] == 0) and np.all(dep[(ze + 1) * T:] == 0)
    # Gauss's law with that charge density on the block planes; outside them the violation
    # stays (the block never includes the last plane, whose violation the scan skips)
    og.rho[:] = dep
    g2 = np.array([O.lib().orc_gauss_constraint_squared(og._g, c) for c in range(og.n)])
    q2 = np.array([O.alg_square(nc, x) for x in gauss])
    assert np.max(g2[zs * T:(ze + 1) * T]) < 1e-20 * np.max(q2)
    assert np.allclose(g2[(ze + 1) * T:], q2[(ze + 1) * T:], rtol=1e-9, atol=1e-30)
    # the refinement moved charge between the particles of a cell
    assert np.max(np.abs(q.reshape(-1, ppc, O.asize(nc)) - q.reshape(-1, ppc, O.asize(nc)).mean(axis=1, keepdims=True))) > 0


def window_off_polynomial(q, i, lo, width, degree):
    """Distance of the last value of the periodic window q[i+lo .. i+lo+width) from the polynomial of
    `degree` through the others (numpy interpolation), relative to the list's scale."""
    nn = len(q)
    xs = np.arange(width, dtype=float)
    ys = np.array([q[(i + lo + m) % nn] for m in range(width)])
    coef = np.polyfit(xs[:-1], ys[:-1], degree)
    return abs(np.polyval(coef, xs[-1]) - ys[-1]) / np.max(np.abs(q))


@pytest.mark.parametrize("ppc,skip", [(4, 2), (4, 0), (2, 1), (8, 4)])
def test_refinement_sweeps_converge_to_polynomial_windows(ppc, skip):
    """LightConeNGPSuperParticleCreator.refine2 / refine4 (:245-300): each sweep moves charge between
    neighbours i, i+1 where (i + skip) % ppc < ppc - 1 so that, at the fixed point, every such
    window of 4 charges (i-1 .. i+2) lies on a quadratic (2nd order) and every window of 6
    (i-2 .. i+3) on a quartic (4th order), the charges summing to the same total. Checked with
    numpy polynomial interpolation of the windows, after enough sweeps to converge; a wrong
    coefficient in either stencil moves the fixed point off the polynomials."""
    nn = 10 * ppc
    q0 = np.random.default_rng(ppc + skip).normal(size=(nn, 1))
    eligible = [i for i in range(nn) if (i + skip) % ppc < ppc - 1]
    q2 = O.refine_list(q0, ppc, skip, 2, 4000)
    assert abs(q2.sum() - q0.sum()) < 1e-12 * np.abs(q0).sum()
    assert max(window_off_polynomial(q2[:, 0], i, -1, 4, 2) for i in eligible) < 1e-9
    q4 = O.refine_list(q2, ppc, skip, 4, 40000)
    assert abs(q4.sum() - q0.sum()) < 1e-12 * np.abs(q0).sum()
    assert max(window_off_polynomial(q4[:, 0], i, -2, 6, 4) for i in eligible) < 1e-9
    # the starting list is far from both
    assert max(window_off_polynomial(q0[:, 0], i, -2, 6, 4) for i in eligible) > 1e-2
    # one sweep moves a generic list but leaves a fixed point where it is
    assert np.max(np.abs(O.refine_list(q0, ppc, skip, 2, 1) - q0)) > 1e-3
    assert np.max(np.abs(O.refine_list(q4, ppc, skip, 4, 1) - q4)) < 1e-9


def test_creator_refinement_reaches_the_quartic_fixed_point():
    """The charges the super-particle creator returns, arranged as each transverse cell's list (plane,
    then shift), have every eligible 6-window on a quartic to the accuracy of the reference's 100
    sweeps, and the ineligible ones do not."""
    nc, dims = 2, (20, 3, 3)
    og, rho, gauss = cgc_state(nc, dims, 1, seed=5)
    arrs, charges, (zs, bw) = og.lightcone_superparticles(gauss, 1, subdivisions=2)
    T, ppc = dims[1] * dims[2], int(arrs["ppc"][0])
    nn = bw * ppc
    lists = np.zeros((T, nn, 3))
    for s_, c in enumerate(charges):
        off, k = arrs["index_offset"][s_], arrs["shift"][s_]
        for i in range(arrs["num"][s_]):
            idx = off + i
            lists[idx % T, (idx // T - zs) * ppc + k] = c[i]
    eligible = [i for i in range(nn) if (i + ppc // 2) % ppc < ppc - 1]
    other = [i for i in range(nn) if i not in eligible]
    worst = max(window_off_polynomial(lists[t, :, a], i, -2, 6, 4) for t in range(T) for a in range(3) for i in eligible)
    assert worst < 1e-5
    assert max(window_off_polynomial(lists[t, :, a], i, -2, 6, 4) for t in range(T) for a in range(3)
               for i in other) > 1e-3
