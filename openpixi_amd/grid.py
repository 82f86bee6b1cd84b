"""Host-side mirror of OpenPixi's Grid / FieldSolver / FieldMeasurements seams.

The lattice lives on the GPU inside an ``opx_ctx``; these classes keep the
reference's method names and argument meaning so a caller written against
``org.openpixi.pixi.physics.grid.Grid`` and ``FastTYMSolver`` reads the same.
Whole-lattice getters/setters exchange numpy arrays in the reference's per-cell
element order (see include/opx.h).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import FIELD_E, FIELD_J, FIELD_RHO, FIELD_U, FIELD_UNEXT, OpxError


def group_size(nc: int) -> int:
    return 4 if nc == 2 else 18


def algebra_size(nc: int) -> int:
    """Reals per algebra element in the reference storage (3 or 9)."""
    return 3 if nc == 2 else 9


def algebra_dim(nc: int) -> int:
    return nc * nc - 1


class Grid:
    """Device-resident lattice (physics/grid/Grid.java:443-473).

    numCells = [N0, N1, N2] with x0 slowest; spacing a_i; time step a_t; coupling g.
    precision: "fp64" or "fp32".
    """

    def __init__(self, num_colors, num_cells, spacing=1.0, time_step=0.25, coupling=1.0, precision="fp64"):
        L = _lib.load()
        if np.isscalar(spacing):
            spacing = [float(spacing)] * 3
        self.num_colors = int(num_colors)
        self.num_cells = [int(x) for x in num_cells]
        self.spacing = [float(x) for x in spacing]
        self.time_step = float(time_step)
        self.coupling = float(coupling)
        self.precision = precision
        prec = {"fp64": _lib.OPX_FP64, "fp32": _lib.OPX_FP32}[precision]
        h = ctypes.c_void_p()
        rc = L.opx_create(self.num_colors, len(self.num_cells), _lib.ivec(self.num_cells), _lib.dvec(self.spacing),
                          self.time_step, self.coupling, prec, ctypes.byref(h))
        if rc != 0:
            raise OpxError(rc, "opx_create failed (colours, dimensions or precision unsupported, or device error)")
        self._h = h
        self.total_cells = int(np.prod(self.num_cells))
        self.simulation_steps = 0

    # ---- lifetime ----
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().opx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def _check(self, rc):
        if rc != 0:
            msg = _lib.load().opx_last_error(self._h)
            raise OpxError(rc, msg.decode() if msg else "")

    # ---- reference accessors ----
    def getNumberOfColors(self):
        return self.num_colors

    def getNumCells(self, i=None):
        return self.num_cells if i is None else self.num_cells[i]

    def getTotalNumberOfCells(self):
        return self.total_cells

    def getLatticeSpacing(self, i=0):
        return self.spacing[i]

    def getTemporalSpacing(self):
        return self.time_step

    def getGaugeCoupling(self):
        return self.coupling

    def getLatticeUnitFactor(self, i):
        return self.coupling * self.spacing[i]

    # ---- whole-lattice data exchange ----
    def _shape(self, field):
        nc = self.num_colors
        if field in (FIELD_U, FIELD_UNEXT):
            return (self.total_cells, 3, group_size(nc))
        if field == FIELD_RHO:
            return (self.total_cells, algebra_size(nc))
        return (self.total_cells, 3, algebra_size(nc))

    def upload(self, field, values):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if a.size != int(np.prod(self._shape(field))):
            raise ValueError(f"field {field}: expected {self._shape(field)} values, got {a.size}")
        self._check(_lib.load().opx_upload_field(self._h, field, _lib.dptr(a), a.size))

    def download(self, field):
        a = np.zeros(self._shape(field), dtype=np.float64)
        self._check(_lib.load().opx_download_field(self._h, field, _lib.dptr(a.reshape(-1)), a.size))
        return a

    def setU(self, values):
        self.upload(FIELD_U, values)

    def setUnext(self, values):
        self.upload(FIELD_UNEXT, values)

    def setE(self, values):
        self.upload(FIELD_E, values)

    def setJ(self, values):
        self.upload(FIELD_J, values)

    def setRho(self, values):
        self.upload(FIELD_RHO, values)

    def getU(self):
        return self.download(FIELD_U)

    def getUnext(self):
        return self.download(FIELD_UNEXT)

    def getE(self):
        return self.download(FIELD_E)

    def getJ(self):
        return self.download(FIELD_J)

    def getRho(self):
        return self.download(FIELD_RHO)

    def setLinksFromCoefficients(self, coeffs, field=FIELD_U):
        """U = exp(i sum_a c_a t_a), coefficients [cell][dir][nc^2-1], computed on the device."""
        a = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(-1)
        self._check(_lib.load().opx_set_links_from_coefficients(self._h, field, _lib.dptr(a), a.size))

    def setEFromCoefficients(self, coeffs):
        a = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(-1)
        self._check(_lib.load().opx_set_efield_from_coefficients(self._h, _lib.dptr(a), a.size))

    # ---- regions (physics/grid/Grid.java:992-1020) ----
    def setActiveRegion(self, p1, p2):
        self._check(_lib.load().opx_set_active_region(self._h, _lib.ivec(p1), _lib.ivec(p2)))

    def setEvaluationRegion(self, p1, p2):
        self._check(_lib.load().opx_set_evaluation_region(self._h, _lib.ivec(p1), _lib.ivec(p2)))

    # ---- Grid actions ----
    def storeFields(self):
        self._check(_lib.load().opx_store_fields(self._h))

    def resetCurrent(self):
        self._check(_lib.load().opx_reset_current(self._h))

    def resetCharge(self):
        self._check(_lib.load().opx_reset_charge(self._h))

    def updateGrid(self, tstep, solver=None):
        (solver or FastTYMSolver()).step(self, tstep)

    def updateLinks(self, tstep, solver=None):
        (solver or FastTYMSolver()).stepLinks(self, tstep)

    # ---- kernel form (no counterpart in the reference) ----
    def setKernel(self, kernel="auto", tile=None, chunks=None):
        """kernel: "auto", "link", "stream", "plaq", "tile" (SU(2)), "plane" or "cell"; tile: stream tile index,
        tile store policy or plane tile shape;
        chunks: x0 chunk count of the stream and tile forms (0 = auto)."""
        L = _lib.load()
        k = {"auto": _lib.KERNEL_AUTO, "link": _lib.KERNEL_LINK, "stream": _lib.KERNEL_STREAM,
             "plaq": _lib.KERNEL_PLAQ, "tile": _lib.KERNEL_TILE,
             "plane": _lib.KERNEL_PLANE, "cell": _lib.KERNEL_CELL}[kernel]
        self._check(L.opx_set_option(self._h, _lib.OPT_KERNEL, k))
        if tile is not None:
            self._check(L.opx_set_option(self._h, _lib.OPT_TILE, int(tile)))
        if chunks is not None:
            self._check(L.opx_set_option(self._h, _lib.OPT_CHUNKS, int(chunks)))

    def setGlobalPointers(self, on=True):
        """Read the fields through plain global pointers in the per-link, cell and plaquette forms,
        as they do for fields of 4 GiB or more (testing that path on small lattices)."""
        self._check(_lib.load().opx_set_option(self._h, _lib.OPT_GLOBAL_POINTERS, 1 if on else 0))

    def setReunitarizeEvery(self, n):
        """Every n-th Simulation step (simulationStep) reprojects the new links onto the group in
        double precision, inside the update where the kernel form can (0: never)."""
        self._check(_lib.load().opx_set_option(self._h, _lib.OPT_REUNIT_EVERY, int(n)))

    def kernelConfig(self):
        L = _lib.load()
        out = []
        for opt in (_lib.OPT_KERNEL, _lib.OPT_TILE, _lib.OPT_CHUNKS):
            v = ctypes.c_int()
            self._check(L.opx_get_option(self._h, opt, ctypes.byref(v)))
            out.append(v.value)
        return {"kernel": ["auto", "link", "stream", "plaq", "tile", "plane", "cell"][out[0]], "tile": out[1],
                "chunks": out[2]}

    def autotune(self, reps=5):
        """Times every kernel form / tile / chunk candidate and keeps the fastest; returns ms per launch."""
        ms = ctypes.c_double()
        self._check(_lib.load().opx_autotune(self._h, int(reps), ctypes.byref(ms)))
        return ms.value

    def projectGauss(self, tol=1e-10, max_iter=2000):
        """Nearest E satisfying Gauss's law for the current U and rho (conjugate gradients, double
        precision). Returns (iterations, mean |G|^2 before, mean |G|^2 after)."""
        it = ctypes.c_int()
        g0, g1 = ctypes.c_double(), ctypes.c_double()
        self._check(_lib.load().opx_gauss_project(self._h, float(tol), int(max_iter), ctypes.byref(it),
                                                  ctypes.byref(g0), ctypes.byref(g1)))
        return it.value, g0.value, g1.value

    def restoreGaussConstraint(self, gamma, maxIterations, absoluteValue):
        """GaussConstraintRestoration.iterateRestorationAlgorithm
        (diagnostics/methods/GaussConstraintRestoration.java:79-112) on the device. Returns
        (corrections kept, reason, violation before, violation after); reason 0: the absolute
        value was reached, 1: a correction made the violation worse and was undone, 2: the
        maximum number of iterations, 3: no iteration ran."""
        it, why = ctypes.c_int(), ctypes.c_int()
        v0, v1 = ctypes.c_double(), ctypes.c_double()
        self._check(_lib.load().opx_gauss_restore(self._h, float(gamma), int(maxIterations), float(absoluteValue),
                                                  ctypes.byref(it), ctypes.byref(why), ctypes.byref(v0),
                                                  ctypes.byref(v1)))
        return it.value, why.value, v0.value, v1.value

    def unitarityTest(self, tol=1e-8, max_cells=1 << 20):
        """UnitarityTester (diagnostics/methods/UnitarityTester.java:31-46): (number of links whose
        |Re tr U| or |Re tr Unext| is not below N + tol, their cell indices in the reference's loop
        order, at most max_cells of them)."""
        cnt = ctypes.c_int64()
        cells = np.zeros(max(int(max_cells), 1), np.int64)
        self._check(_lib.load().opx_unitarity_test(self._h, float(tol), ctypes.byref(cnt),
                                                   cells.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                   int(max_cells)))
        return cnt.value, cells[: min(cnt.value, int(max_cells))]

    def reunitarize(self):
        self._check(_lib.load().opx_reunitarize(self._h))

    def sync(self):
        self._check(_lib.load().opx_sync(self._h))

    def checkHealth(self, tol=1e-6):
        nf, nu = ctypes.c_int64(), ctypes.c_int64()
        rc = _lib.load().opx_check_health(self._h, tol, ctypes.byref(nf), ctypes.byref(nu))
        if rc not in (0, _lib.OPX_E_NUMERIC):
            self._check(rc)
        return nf.value, nu.value


    # ---- colored particles: CGC super particles (physics/particles/CGCSuperParticle.java:16-46) ----
    def setSuperParticles(self, orientation, number, index_offset, per_plane, sub_lattice_shift,
                          particles_per_cell, charges):
        """Upload a particle list; charges: (sum(number), A) reference algebra storage, particle after particle."""
        L = _lib.load()
        arrs = [np.ascontiguousarray(a, np.int32) for a in (orientation, number, index_offset, per_plane,
                                                             sub_lattice_shift, particles_per_cell)]
        q = np.ascontiguousarray(charges, np.float64).reshape(-1)
        self._check(L.opx_superparticles_upload(self._h, len(arrs[0]), *[_lib.iptr(a) for a in arrs],
                                                 _lib.dptr(q), q.size))
        self._sp_total = int(np.sum(arrs[1]))

    def superParticleCharges(self):
        A = algebra_size(self.num_colors)
        out = np.zeros(self._sp_total * A)
        self._check(_lib.load().opx_superparticles_download_charges(self._h, _lib.dptr(out), out.size))
        return out.reshape(-1, A)

    def interpolateToGrid(self, t):
        """CGCSuperParticleInterpolationNGP.interpolateToGrid for every particle at integer clock t."""
        self._check(_lib.load().opx_interpolate_to_grid(self._h, int(t)))

    def interpolateChargedensity(self, t):
        self._check(_lib.load().opx_interpolate_chargedensity(self._h, int(t)))

    def setParticles(self, direction, pos0, vel, charges):
        """Upload per-particle CGC NGP particles (CGCParticle): direction (count,), pos0 and
        vel (count, 3), charges (count, A) reference algebra storage."""
        d = np.ascontiguousarray(direction, np.int32)
        p0 = np.ascontiguousarray(pos0, np.float64).reshape(-1)
        v = np.ascontiguousarray(vel, np.float64).reshape(-1)
        q = np.ascontiguousarray(charges, np.float64).reshape(-1)
        self._check(_lib.load().opx_particles_upload(self._h, len(d), _lib.iptr(d), _lib.dptr(p0), _lib.dptr(v),
                                                     _lib.dptr(q), q.size))
        self._p_count = len(d)

    def particleCount(self):
        c = ctypes.c_int64()
        self._check(_lib.load().opx_particles_count(self._h, ctypes.byref(c)))
        return c.value

    def particleState(self):
        """(pos0, pos1, Q0, Q1) of the per-particle list."""
        c = self.particleCount()
        A = algebra_size(self.num_colors)
        p0, p1, q0, q1 = np.zeros(c * 3), np.zeros(c * 3), np.zeros(c * A), np.zeros(c * A)
        self._check(_lib.load().opx_particles_download(self._h, _lib.dptr(p0), _lib.dptr(p1), _lib.dptr(q0),
                                                       _lib.dptr(q1), c))
        return p0.reshape(-1, 3), p1.reshape(-1, 3), q0.reshape(-1, A), q1.reshape(-1, A)

    def simulationInitialize(self):
        """Simulation.initialize (physics/Simulation.java:318-352) on the device."""
        self._check(_lib.load().opx_simulation_initialize(self._h))

    def simulationStep(self, n=1):
        """n x Simulation.step (physics/Simulation.java:373-435) without diagnostics."""
        self._check(_lib.load().opx_simulation_step(self._h, int(n)))
        return _lib.load().opx_simulation_steps_done(self._h)


class FieldSolver:
    """physics/fields/FieldSolver.java:114-156: step and stepLinks on a Grid."""

    def clone(self):
        return type(self)()

    def step(self, grid: Grid, timeStep: float):
        pass

    def stepLinks(self, grid: Grid, timeStep: float):
        pass


class FastTYMSolver(FieldSolver):
    """physics/fields/FastTYMSolver.java:7-85 on the gfx950 kernels."""

    def step(self, grid: Grid, timeStep: float):
        L = _lib.load()
        grid._check(L.opx_set_field_solver(grid.handle, _lib.SOLVER_FAST, 0, 0.0))
        grid._check(L.opx_step(grid.handle, float(timeStep)))

    def stepLinks(self, grid: Grid, timeStep: float):
        grid._check(_lib.load().opx_step_links(grid.handle, float(timeStep)))


class ImplicitTYMSolver(FieldSolver):
    """physics/fields/ImplicitTYMSolver.java:7-263 on the gfx950 kernels: the explicit update,
    then `implicitIterations` passes with plaquettes averaged over the past and future links
    (beam direction x0), each damped by Unext = U^damping Unext^(1 - damping)."""

    def __init__(self, implicitIterations=0, implicitDampingFactor=0.0):
        self.implicitIterations = int(implicitIterations)
        self.implicitDampingFactor = float(implicitDampingFactor)

    def clone(self):
        return ImplicitTYMSolver(self.implicitIterations, self.implicitDampingFactor)

    def step(self, grid: Grid, timeStep: float):
        L = _lib.load()
        grid._check(L.opx_set_field_solver(grid.handle, _lib.SOLVER_IMPLICIT, self.implicitIterations,
                                           self.implicitDampingFactor))
        grid._check(L.opx_step(grid.handle, float(timeStep)))

    def stepLinks(self, grid: Grid, timeStep: float):
        grid._check(_lib.load().opx_step_links(grid.handle, float(timeStep)))


class FieldMeasurements:
    """physics/measurements/FieldMeasurements.java:26-66 computed on the device in double precision."""

    def _all(self, grid: Grid):
        out = np.zeros(8)
        grid._check(_lib.load().opx_measure(grid.handle, _lib.dptr(out)))
        return out

    def measure(self, grid: Grid):
        """[E2_0,E2_1,E2_2,B2_0,B2_1,B2_2,gauss_mean_sq,total_charge_squared]."""
        return self._all(grid)

    def calculateEsquared(self, grid: Grid, direction=None):
        m = self._all(grid)
        return float(m[direction]) if direction is not None else float(m[0:3].sum() * grid.total_cells)

    def calculateBsquared(self, grid: Grid, direction=None):
        m = self._all(grid)
        return float(m[3 + direction]) if direction is not None else float(m[3:6].sum() * grid.total_cells)

    def calculateGaussConstraint(self, grid: Grid):
        return float(self._all(grid)[6])

    def calculateTotalChargeSquared(self, grid: Grid):
        return float(self._all(grid)[7])

    def projectedPoynting(self, grid: Grid, direction: int):
        """Longitudinal Poynting vector per slice, (averaged, time averaged)
        (diagnostics/methods/ProjectedEnergyDensity.java:178-262). SU(2) only."""
        n = grid.num_cells[direction]
        avg, tavg = np.zeros(n), np.zeros(n)
        grid._check(_lib.load().opx_projected_poynting(grid.handle, direction, _lib.dptr(avg), _lib.dptr(tavg)))
        return avg, tavg

    def projectedEnergyDensity(self, grid: Grid, direction: int):
        n = grid.num_cells[direction]
        arrs = [np.zeros(n) for _ in range(4)]
        grid._check(_lib.load().opx_projected_energy_density(grid.handle, direction, *[_lib.dptr(a) for a in arrs]))
        return arrs
