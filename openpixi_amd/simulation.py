"""Host-side mirror of OpenPixi's Settings (YAML) and Simulation driver for the
FastTYMSolver path.

``Settings.from_yaml`` accepts the reference's YAML keys
(ui/util/yaml/YamlSettings.java:14-191) that concern this path, including the
``output`` diagnostics of this path (bulkQuantitiesInTime, unitarityTester,
gaussConstraintRestoration, projectedEnergyDensity), and
ignores panel sections. ``Simulation`` keeps the reference's step ordering
(physics/Simulation.java:318-452): ``initialize()`` once, then ``step()`` while
``continues()``, which runs ceil(duration/timeStep) + 1 steps like the reference;
diagnostics run before the first step and after every step.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from .grid import FieldMeasurements, FastTYMSolver, Grid, ImplicitTYMSolver

SIMULATION_TYPES = {
    "temporal yang-mills": "TemporalYangMills",
    "temporal cgc": "TemporalCGC",
    "temporal cgc ngp": "TemporalCGCNGP",
    "temporal implicit cgc ngp": "TemporalImplicitCGCNGP",
    "temporal optimized cgc ngp": "TemporalOptimizedCGCNGP",
    "temporal optimized implicit cgc ngp": "TemporalOptimizedImplicitCGCNGP",
    "boost-invariant cgc": "BoostInvariantCGC",
}

# simulation types whose field solver is FastTYMSolver or ImplicitTYMSolver (physics/Settings.java:540-581)
FAST_TYM_TYPES = {"TemporalYangMills", "TemporalCGC", "TemporalCGCNGP", "TemporalOptimizedCGCNGP"}
IMPLICIT_TYM_TYPES = {"TemporalImplicitCGCNGP", "TemporalOptimizedImplicitCGCNGP"}


@dataclass
class Region:
    enabled: bool = False
    point1: Optional[List[int]] = None
    point2: Optional[List[int]] = None

    def resolved(self, grid_cells):
        """YamlRegion.getPoint1/2: negative coordinates count from the end."""
        p1 = [p + n if p < 0 else p for p, n in zip(self.point1, grid_cells)]
        p2 = [p + n if p < 0 else p for p, n in zip(self.point2, grid_cells)]
        return p1, p2


@dataclass
class Settings:
    """Subset of physics/Settings.java used by the FastTYMSolver path, with the
    YamlSettings defaults (3 dimensions, 1 colour, coupling 1, YamlSettings.java:49-56)."""
    simulation_type: str = "TemporalYangMills"
    time_step: float = 0.1
    duration: float = 0.0
    number_of_dimensions: int = 3
    number_of_colors: int = 1
    coupling_constant: float = 1.0
    grid_steps: List[float] = field(default_factory=lambda: [1.0, 1.0, 1.0])
    grid_cells: List[int] = field(default_factory=lambda: [1, 1, 1])
    evaluation_region: Region = field(default_factory=Region)
    active_region: Region = field(default_factory=Region)
    precision: str = "fp64"          # device precision (not a reference key)
    reunitarize_every: int = 0       # FP64 reprojection period (not a reference key)
    diagnostics: list = field(default_factory=list)   # Settings.addDiagnostics (YamlOutput.applyTo)
    implicit_solver_iterations: int = 0               # physics/Settings.java:79
    implicit_solver_damping: float = 0.0              # physics/Settings.java:80
    number_of_threads: int = 1                        # physics/Settings.java:100 (super-particle subdivisions)
    initial_conditions: list = field(default_factory=list)   # Settings.addInitialConditions (YamlInitialConditions)

    @property
    def iterations(self) -> int:
        """physics/Settings.java:374-379: ceil(tMax / timeStep)."""
        return int(math.ceil(self.duration / self.time_step))

    @classmethod
    def from_yaml(cls, text: str) -> "Settings":
        import yaml

        y = yaml.safe_load(text) or {}
        s = cls()
        if "simulationType" in y:
            key = y["simulationType"]
            if key not in SIMULATION_TYPES:
                raise RuntimeError("Unknown simulation type specified in YAML file.")
            s.simulation_type = SIMULATION_TYPES[key]
        if "timeStep" in y:
            s.time_step = float(y["timeStep"])
        if "duration" in y:
            s.duration = float(y["duration"])
        if "numberOfDimensions" in y:
            s.number_of_dimensions = int(y["numberOfDimensions"])
        if "numberOfColors" in y:
            s.number_of_colors = int(y["numberOfColors"])
        if "couplingConstant" in y:
            s.coupling_constant = float(y["couplingConstant"])
        cells = y.get("gridCells")
        if "gridSteps" in y:
            if len(y["gridSteps"]) != s.number_of_dimensions:
                raise RuntimeError("Size of gridSteps does not match numberOfDimensions.")
            s.grid_steps = [float(v) for v in y["gridSteps"]]
        elif "gridStep" in y:
            s.grid_steps = [float(y["gridStep"])] * s.number_of_dimensions
        if cells is not None:
            if len(cells) != s.number_of_dimensions:
                raise RuntimeError("Size of gridCells does not match numberOfDimensions.")
            s.grid_cells = [int(v) for v in cells]
        for key, attr in (("evaluationRegion", "evaluation_region"), ("activeRegion", "active_region")):
            r = y.get(key)
            if r and r.get("point1") is not None and r.get("point2") is not None and \
                    len(r["point1"]) == s.number_of_dimensions and len(r["point2"]) == s.number_of_dimensions:
                setattr(s, attr, Region(bool(r.get("enabled", False)), list(r["point1"]), list(r["point2"])))
        if y.get("implicitSolverIterations") is not None:
            s.implicit_solver_iterations = int(y["implicitSolverIterations"])
        if y.get("implicitSolverDamping") is not None:
            s.implicit_solver_damping = float(y["implicitSolverDamping"])
        if y.get("numberOfThreads") is not None:
            s.number_of_threads = int(y["numberOfThreads"])
        ic = y.get("initialConditions") or {}
        if ic.get("YM"):
            raise NotImplementedError("Yang-Mills initial conditions (YamlYM) are not built")
        if ic.get("CGC"):
            from .cgc import from_yaml_cgc

            s.initial_conditions = from_yaml_cgc(ic["CGC"])
        if y.get("output"):
            from .diagnostics import from_yaml_output

            s.diagnostics = from_yaml_output(y["output"])
        dev = y.get("device") or {}
        if "precision" in dev:
            s.precision = str(dev["precision"])
        if "reunitarizeEvery" in dev:
            s.reunitarize_every = int(dev["reunitarizeEvery"])
        return s


@dataclass
class ParticleSet:
    """A list of CGCParticle (physics/particles/CGCParticle.java) for CGCParticleInterpolationNGP."""
    direction: np.ndarray        # (count,) beam axis of each particle
    pos0: np.ndarray             # (count, 3) position at t = 0
    vel: np.ndarray              # (count, 3)
    charges: np.ndarray          # (count, A) reference algebra storage (Q0)


@dataclass
class SuperParticleSet:
    """A list of CGCSuperParticle (physics/particles/CGCSuperParticle.java:16-46)."""
    orientation: np.ndarray
    number: np.ndarray
    index_offset: np.ndarray
    per_plane: np.ndarray
    sub_lattice_shift: np.ndarray
    particles_per_cell: np.ndarray
    charges: np.ndarray          # (sum(number), A) reference algebra storage


class Simulation:
    """physics/Simulation.java driver over the device grid (FastTYMSolver types only)."""

    def __init__(self, settings: Settings, superparticles: Optional[SuperParticleSet] = None,
                 particles: Optional[ParticleSet] = None):
        if settings.simulation_type not in FAST_TYM_TYPES | IMPLICIT_TYM_TYPES:
            raise NotImplementedError(f"{settings.simulation_type} uses neither FastTYMSolver nor ImplicitTYMSolver")
        if settings.number_of_dimensions != 3:
            raise NotImplementedError("only 3 spatial dimensions are supported")
        self.settings = settings
        self.tstep = settings.time_step
        self.iterations = settings.iterations
        self.totalSimulationSteps = 0
        self._diagnostics_started = False
        self.grid = Grid(settings.number_of_colors, settings.grid_cells, settings.grid_steps, settings.time_step,
                         settings.coupling_constant, settings.precision)
        if settings.simulation_type in IMPLICIT_TYM_TYPES:
            self.fsolver = ImplicitTYMSolver(settings.implicit_solver_iterations, settings.implicit_solver_damping)
            self.grid._check(_lib.load().opx_set_field_solver(self.grid.handle, _lib.SOLVER_IMPLICIT,
                                                              settings.implicit_solver_iterations,
                                                              settings.implicit_solver_damping))
        else:
            self.fsolver = FastTYMSolver()
        # FP64 reprojection of the new links every reunitarize_every steps, inside opx_simulation_step
        self.grid.setReunitarizeEvery(settings.reunitarize_every)
        if settings.evaluation_region.enabled:
            self.grid.setEvaluationRegion(*settings.evaluation_region.resolved(settings.grid_cells))
        if settings.active_region.enabled:
            self.grid.setActiveRegion(*settings.active_region.resolved(settings.grid_cells))
        self.superparticles = None
        if superparticles is not None:
            self.set_superparticles(superparticles)
        self.particles = None
        if particles is not None:
            self.set_particles(particles)
        # physics/Simulation.java:274-276: initial conditions after the regions and particle lists
        for ic in settings.initial_conditions:
            ic.applyInitialCondition(self)

    def set_particles(self, p: ParticleSet):
        self.grid.setParticles(p.direction, p.pos0, p.vel, p.charges)
        self.particles = p

    def set_superparticles(self, sp: SuperParticleSet):
        L = _lib.load()
        arrs = [np.ascontiguousarray(a, np.int32) for a in (sp.orientation, sp.number, sp.index_offset,
                                                             sp.per_plane, sp.sub_lattice_shift,
                                                             sp.particles_per_cell)]
        q = np.ascontiguousarray(sp.charges, np.float64).reshape(-1)
        self.grid._check(L.opx_superparticles_upload(self.grid.handle, len(arrs[0]), *[_lib.iptr(a) for a in arrs],
                                                     _lib.dptr(q), q.size))
        self.superparticles = sp

    def superparticle_charges(self):
        """The charges of the device's super-particle list, (total, A) in reference storage."""
        import ctypes

        count, total = ctypes.c_int(), ctypes.c_int64()
        self.grid._check(_lib.load().opx_superparticles_count(self.grid.handle, ctypes.byref(count), ctypes.byref(total)))
        a = 3 if self.grid.num_colors == 2 else 9
        out = np.zeros(total.value * a)
        self.grid._check(_lib.load().opx_superparticles_download_charges(self.grid.handle, _lib.dptr(out), out.size))
        return out.reshape(-1, a)

    def superparticle_descriptors(self) -> SuperParticleSet:
        """The device's super-particle list (descriptors and charges)."""
        import ctypes

        count, total = ctypes.c_int(), ctypes.c_int64()
        self.grid._check(_lib.load().opx_superparticles_count(self.grid.handle, ctypes.byref(count), ctypes.byref(total)))
        arrs = [np.zeros(count.value, np.int32) for _ in range(6)]
        self.grid._check(_lib.load().opx_superparticles_download_descriptors(self.grid.handle, count.value,
                                                                            *[_lib.iptr(x) for x in arrs]))
        return SuperParticleSet(*arrs, charges=self.superparticle_charges())

    def initialize(self):
        """physics/Simulation.java:318-352."""
        self.grid._check(_lib.load().opx_simulation_initialize(self.grid.handle))

    def step(self, n: int = 1):
        """physics/Simulation.java:373-435 (external currents excluded). Diagnostics
        are initialized and run before the first step and run after every step."""
        diags = self.settings.diagnostics
        if diags and self.totalSimulationSteps == 0 and not self._diagnostics_started:
            for d in diags:
                d.initialize(self)
            self.runDiagnostics()
            self._diagnostics_started = True
        if diags:
            for _ in range(int(n)):
                self._one_step(1)
                self.runDiagnostics()
        else:
            self._one_step(int(n))

    def _one_step(self, n: int):
        self.grid._check(_lib.load().opx_simulation_step(self.grid.handle, n))
        self.totalSimulationSteps = _lib.load().opx_simulation_steps_done(self.grid.handle)

    def runDiagnostics(self):
        """physics/Simulation.java:454-460."""
        for d in self.settings.diagnostics:
            d.calculate(self.grid, self.totalSimulationSteps)

    def continues(self) -> bool:
        """physics/Simulation.java:441-443: steps <= iterations (iterations + 1 steps in total)."""
        return self.totalSimulationSteps <= self.iterations

    def run(self):
        while self.continues():
            self.step()

    def measure(self):
        return FieldMeasurements().measure(self.grid)
