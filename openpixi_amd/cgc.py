"""CGC initial condition on the device, beam along x0: the host-side mirror of
``CGCInitialCondition`` (physics/initial/CGC/CGCInitialCondition.java:55-93) with an
``MVModel`` charge density (MVModel.java:101-157), ``LightConePoissonSolverImprovedFull``
(LightConePoissonSolverImprovedFull.java:34-263), ``WilsonLineObservables``, and
``LightConeNGPSuperParticleCreator`` (LightConeNGPSuperParticleCreator.java:76-307) or
``LightConeNGPParticleCreator`` (LightConeNGPParticleCreator.java:82-297).

The classes keep the reference's names, constructor arguments and call order; the work
runs in ``libopx.so`` (``opx_cgc_*``), which keeps the charge density and the Gauss
violation on the device between the calls. The MV model's noise is drawn on the host
from ``JavaRandom``, the generator ``java.util.Random`` documents, so a seeded run draws
the same unit Gaussians, in the same order, as the reference.
"""
from __future__ import annotations

import ctypes
import math
import os
from decimal import ROUND_HALF_UP, Decimal

import numpy as np

from . import _lib

_MULT = 0x5DEECE66D
_ADDEND = 0xB
_MASK = (1 << 48) - 1


# Coefficients and split ln 2 of the freely distributable fdlibm e_log.c (Sun Microsystems, 1993),
# the algorithm java.lang.StrictMath.log specifies.
_LG = (6.666666666666735130e-01, 3.999999999940941908e-01, 2.857142874366239149e-01, 2.222219843214978396e-01,
       1.818357216161805012e-01, 1.531383769920937332e-01, 1.479819860511658591e-01)
_LN2_HI, _LN2_LO = 6.93147180369123816490e-01, 1.90821492927058770002e-10


def strict_log(x) -> np.ndarray:
    """StrictMath.log for positive finite x (fdlibm e_log.c): x = 2^k (1 + f) with 1 + f in
    [sqrt(2)/2, sqrt(2)), log(1 + f) = f - s (f - R(s^2)) with s = f / (2 + f) and the degree-7
    polynomial R, k ln 2 added in two parts. Within 1 ulp, and not always the correctly rounded
    value, which is why nextGaussian needs it and not the platform log."""
    x = np.array(x, np.float64, ndmin=1)
    k = np.zeros(x.shape, np.int64)
    sub = x < 2.0 ** -1022
    if np.any(sub):
        x[sub] *= 2.0 ** 54
        k[sub] -= 54
    bits = x.view(np.uint64)
    hx = (bits >> np.uint64(32)).astype(np.int64)
    k += (hx >> 20) - 1023
    hx &= 0x000fffff
    i = (hx + 0x95f64) & 0x100000
    hi = (hx | (i ^ 0x3ff00000)).astype(np.uint64)
    x = ((hi << np.uint64(32)) | (bits & np.uint64(0xffffffff))).view(np.float64)
    k += i >> 20
    f = x - 1.0
    dk = k.astype(np.float64)
    s = f / (2.0 + f)
    z = s * s
    w = z * z
    t1 = w * (_LG[1] + w * (_LG[3] + w * _LG[5]))
    t2 = z * (_LG[0] + w * (_LG[2] + w * (_LG[4] + w * _LG[6])))
    R = t2 + t1
    hfsq = 0.5 * f * f
    upper = ((hx - 0x6147a) | (0x6b851 - hx)) > 0
    with np.errstate(invalid="ignore", divide="ignore"):
        a = np.where(k == 0, f - (hfsq - s * (hfsq + R)),
                     dk * _LN2_HI - ((hfsq - (s * (hfsq + R) + dk * _LN2_LO)) - f))
        b = np.where(k == 0, f - s * (f - R), dk * _LN2_HI - ((s * (f - R) - dk * _LN2_LO) - f))
        out = np.where(upper, a, b)
        # |f| < 2^-20
        Rt = f * f * (0.5 - 0.33333333333333333 * f)
        t = np.where(f == 0, np.where(k == 0, 0.0, dk * _LN2_HI + dk * _LN2_LO),
                     np.where(k == 0, f - Rt, dk * _LN2_HI - ((Rt - dk * _LN2_LO) - f)))
    return np.where((0x000fffff & (2 + hx)) < 3, t, out)


class JavaRandom:
    """java.util.Random as its API specification defines it: a 48-bit linear congruential
    generator (seed' = seed * 0x5DEECE66D + 0xB mod 2^48, next(bits) the top `bits` bits),
    nextDouble from 26 + 27 bits, nextGaussian by the polar method of Box, Muller and
    Marsaglia with the second value of each accepted pair kept for the next call, through
    StrictMath.log and StrictMath.sqrt."""

    def __init__(self, seed=None):
        if seed is None:
            seed = int.from_bytes(os.urandom(8), "little")
        self.setSeed(seed)

    def setSeed(self, seed: int):
        self._seed = (int(seed) ^ _MULT) & _MASK
        self._have_next = False
        self._next_gaussian = 0.0

    def next(self, bits: int) -> int:
        self._seed = (self._seed * _MULT + _ADDEND) & _MASK
        r = self._seed >> (48 - bits)
        if r >= 1 << 31:          # Java's int is 32-bit two's complement
            r -= 1 << 32
        return r

    def nextInt(self) -> int:
        return self.next(32)

    def nextDouble(self) -> float:
        return ((self.next(26) << 27) + self.next(27)) * 2.0 ** -53

    def nextGaussian(self) -> float:
        if self._have_next:
            self._have_next = False
            return self._next_gaussian
        while True:
            v1 = 2 * self.nextDouble() - 1
            v2 = 2 * self.nextDouble() - 1
            s = v1 * v1 + v2 * v2
            if s < 1 and s != 0:
                break
        m = math.sqrt(-2 * float(strict_log(s)[0]) / s)
        self._next_gaussian = v2 * m
        self._have_next = True
        return v1 * m

    def _states(self, k: int) -> np.ndarray:
        """The next k states after the current one, seed_j = A_j seed + C_j (mod 2^48), with
        (A_j, C_j) built by composing the affine step with itself; uint64 arithmetic wraps
        mod 2^64, which 2^48 divides."""
        a = np.empty(k, np.uint64)
        c = np.empty(k, np.uint64)
        a[0] = _MULT
        c[0] = _ADDEND
        m = 1
        while m < k:
            t = min(m, k - m)
            am, cm = a[m - 1], c[m - 1]
            a[m:m + t] = a[:t] * am
            c[m:m + t] = a[:t] * cm + c[:t]
            m += t
        return (a * np.uint64(self._seed) + c) & np.uint64(_MASK)

    def gaussians(self, count: int) -> np.ndarray:
        """`count` successive nextGaussian() values, vectorized; leaves the generator where
        the scalar calls would."""
        count = int(count)
        out = np.empty(count, np.float64)
        pos = 0
        if count and self._have_next:
            out[0] = self._next_gaussian
            self._have_next = False
            pos = 1
        with np.errstate(over="ignore"):
            while pos < count:
                need = (count - pos + 1) // 2
                pairs = int(need / 0.78) + 64
                st = self._states(4 * pairs)
                hi = st[0::2] >> np.uint64(22)
                lo = st[1::2] >> np.uint64(21)
                d = ((hi << np.uint64(27)) + lo).astype(np.float64) * 2.0 ** -53
                v1 = 2 * d[0::2] - 1
                v2 = 2 * d[1::2] - 1
                s = v1 * v1 + v2 * v2
                idx = np.nonzero((s < 1) & (s != 0))[0]
                if len(idx) >= need:
                    idx = idx[:need]
                    used = int(idx[-1]) + 1
                else:
                    used = pairs
                sa = s[idx]
                m = np.sqrt(-2 * strict_log(sa) / sa)
                g = np.empty(2 * len(idx))
                g[0::2] = v1[idx] * m
                g[1::2] = v2[idx] * m
                take = min(len(g), count - pos)
                out[pos:pos + take] = g[:take]
                if take < len(g):
                    self._next_gaussian = float(g[take])
                    self._have_next = True
                pos += take
                self._seed = int(st[4 * used - 1])
        return out


def _grid(s):
    return getattr(s, "grid", s)


def _components(nc: int) -> int:
    return nc * nc - 1 if nc > 1 else 1


def _alg_size(nc: int) -> int:
    return 3 if nc == 2 else 9


def _check_beam(direction):
    if direction != 0:
        raise NotImplementedError("the device CGC initial condition supports the beam along x0 only")


class MVModel:
    """physics/initial/CGC/MVModel.java: MV color charges with longitudinal randomness,
    regulated in Fourier space (hard transverse UV cutoff, Gaussian longitudinal
    coherence, transverse IR mass)."""

    def __init__(self, direction, orientation, location, longitudinalWidth, mu, useSeed, seed,
                 ultravioletCutoffTransverse, longitudinalCoherenceLength, infraredCoefficient):
        _check_beam(direction)
        self.direction = int(direction)
        self.orientation = int(orientation)
        self.location = float(location)
        self.longitudinalWidth = float(longitudinalWidth)
        self.mu = float(mu)
        self.useSeed = bool(useSeed)
        self.seed = int(seed)
        self.ultravioletCutoffTransverse = float(ultravioletCutoffTransverse)
        self.longitudinalCoherenceLength = float(longitudinalCoherenceLength)
        self.infraredCoefficient = float(infraredCoefficient)
        self._grid = None

    def noise(self, grid) -> np.ndarray:
        """The unit Gaussians MVModel.initialize draws (MVModel.java:115-128): components outer,
        cells inner, from java.util.Random (seeded when useSeed)."""
        rand = JavaRandom(self.seed if self.useSeed else None)
        return rand.gaussians(_components(grid.getNumberOfColors()) * grid.getTotalNumberOfCells())

    def initialize(self, s, noise=None):
        """MVModel.initialize. The noise is drawn on the device from java.util.Random(seed), or
        from a random seed without useSeed (as `new Random()`); `noise` replaces it when given."""
        g = _grid(s)
        L = _lib.load()
        args = (self.location, self.longitudinalWidth, self.mu, self.ultravioletCutoffTransverse,
                self.longitudinalCoherenceLength, self.infraredCoefficient)
        if noise is None:
            seed = self.seed if self.useSeed else int.from_bytes(os.urandom(6), "little")
            g._check(L.opx_cgc_mv_charge_seeded(g.handle, *args, seed, None, 0))
        else:
            z = np.ascontiguousarray(noise, np.float64).reshape(-1)
            g._check(L.opx_cgc_mv_charge(g.handle, *args, _lib.dptr(z), z.size, None, 0))
        self._grid = g

    def getChargeDensity(self) -> np.ndarray:
        """(cells, A) in reference algebra storage."""
        return _download(self._grid, _lib.CGC_CHARGE)

    def getInfo(self) -> str:
        """MVModel.getInfo (MVModel.java:175-185)."""
        f = java_format_f
        return (f"MV, mu: {f(self.mu)}, w: {f(self.longitudinalWidth)}, UVT: {f(self.ultravioletCutoffTransverse)}, "
                f"c: {f(self.longitudinalCoherenceLength)}, m: {f(self.infraredCoefficient)}")

    def getDirection(self):
        return self.direction

    def getOrientation(self):
        return self.orientation

    def clear(self):
        if self._grid is not None:
            self._grid._check(_lib.load().opx_cgc_clear(self._grid.handle))


class ChargeDensity:
    """An IInitialChargeDensity given as an array (cells, A) in reference algebra storage."""

    def __init__(self, rho, direction=0, orientation=1):
        _check_beam(direction)
        self.rho = np.ascontiguousarray(rho, np.float64)
        self.direction = int(direction)
        self.orientation = int(orientation)
        self._grid = None

    def initialize(self, s):
        g = _grid(s)
        a = self.rho.reshape(-1)
        g._check(_lib.load().opx_cgc_set_charge(g.handle, _lib.dptr(a), a.size))
        self._grid = g

    def getChargeDensity(self):
        return self.rho.reshape(-1, _alg_size(self._grid.getNumberOfColors()))

    def getInfo(self) -> str:
        return "charge density"

    def getDirection(self):
        return self.direction

    def getOrientation(self):
        return self.orientation

    def clear(self):
        if self._grid is not None:
            self._grid._check(_lib.load().opx_cgc_clear(self._grid.handle))


def _download(g, which):
    out = np.zeros(g.getTotalNumberOfCells() * _alg_size(g.getNumberOfColors()))
    g._check(_lib.load().opx_cgc_download(g.handle, which, _lib.dptr(out), out.size))
    return out.reshape(g.getTotalNumberOfCells(), -1)


def java_format_f(x: float, precision: int = 6) -> str:
    """String.format("%f") of a double (java.util.Formatter): the digits of Double.toString
    rounded half up to `precision` decimals."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    q = Decimal(repr(x)).quantize(Decimal(1).scaleb(-precision), rounding=ROUND_HALF_UP)
    return format(q, "f")


class WilsonLineObservables:
    """physics/initial/CGC/WilsonLineObservables.java: tadpole and dipole of the Wilson lines the
    light-cone solve leaves on the transverse plane behind the nucleus, appended to text files
    in the reference's format (FileFunctions.getFile paths)."""

    def __init__(self, s, solver, density):
        self.grid = _grid(s)
        self.solver = solver
        self.density = density

    def tadpole(self) -> float:
        v = ctypes.c_double()
        self.grid._check(_lib.load().opx_cgc_wilson_tadpole(self.grid.handle, ctypes.byref(v)))
        return v.value

    def dipole(self):
        """(bin distances bin * ds, mean tr(V_x V_y^+) per bin, pair counts)."""
        L = _lib.load()
        nb, ds = ctypes.c_int(), ctypes.c_double()
        self.grid._check(L.opx_cgc_wilson_dipole(self.grid.handle, 0, ctypes.byref(nb), ctypes.byref(ds), None, None))
        corr = np.zeros(max(nb.value, 1))
        counts = np.zeros(max(nb.value, 1), np.int64)
        self.grid._check(L.opx_cgc_wilson_dipole(self.grid.handle, len(corr), None, None, _lib.dptr(corr),
                                                 counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        n = nb.value
        return np.arange(n) * ds.value, corr[:n], counts[:n]

    def computeTadpole(self, filename: str):
        from .diagnostics import java_sci, output_path

        with open(output_path(filename), "a") as f:
            f.write(self.density.getInfo() + ", V: " + java_sci(self.tadpole()) + "\n")

    def computeDipole(self, filename: str):
        from .diagnostics import java_sci, output_path

        d, corr, _ = self.dipole()
        with open(output_path(filename), "a") as f:
            f.write(self.density.getInfo() + "\n")
            f.write("d\ttr(V V^t)\n")
            for b in range(len(corr)):
                f.write(java_sci(d[b]) + "\t" + java_sci(corr[b]) + "\n")


class LightConePoissonSolverImprovedFull:
    """physics/initial/CGC/LightConePoissonSolverImprovedFull.java: per-sheet transverse
    Poisson equation, path-ordered Wilson lines at t = -at/2 and +at/2, the grid's
    transverse links rotated by them and E from the temporal plaquette."""

    def __init__(self):
        self._grid = None

    def initialize(self, s):
        self._grid = _grid(s)

    def solve(self, initialChargeDensity):
        _check_beam(initialChargeDensity.getDirection())
        g = self._grid
        g._check(_lib.load().opx_cgc_lightcone_solve(g.handle, int(initialChargeDensity.getOrientation()), None, 0))

    def getGaussViolation(self) -> np.ndarray:
        return _download(self._grid, _lib.CGC_GAUSS)


class LightConeNGPSuperParticleCreator:
    """physics/initial/CGC/LightConeNGPSuperParticleCreator.java: super particles over the
    block of planes that carries the solver's Gauss violation, refined to 4th order along
    the beam, added to the simulation's particle list. The Gauss violation is the one the
    last light-cone solve left on the device."""

    def __init__(self):
        self.particlesPerCell = 1
        self.zStart = None
        self.blockWidth = None
        self.added = 0

    def setGaussConstraint(self, gaussConstraint=None):
        """Kept for the reference's call sequence; the device holds the solver's violation."""

    def initialize(self, s, direction, orientation):
        _check_beam(direction)
        g = _grid(s)
        threads = int(getattr(getattr(s, "settings", None), "number_of_threads", 1) or 1)
        self.particlesPerCell = int(math.floor(g.getLatticeSpacing(0) / g.getTemporalSpacing() + 0.5))
        added, zs, bw = (ctypes.c_int() for _ in range(3))
        g._check(_lib.load().opx_cgc_create_superparticles(g.handle, int(orientation), threads, ctypes.byref(added),
                                                           ctypes.byref(zs), ctypes.byref(bw)))
        self.added, self.zStart, self.blockWidth = added.value, zs.value, bw.value


class LightConeNGPParticleCreator:
    """physics/initial/CGC/LightConeNGPParticleCreator.java: round(a0/at) NGP particles per cell of
    the block of planes that carries the solver's Gauss violation, their charges refined along
    the beam, added to the simulation's per-particle list (the non-optimized CGC types)."""

    def __init__(self):
        self.particlesPerCell = 1
        self.zStart = None
        self.zEnd = None
        self.added = 0

    def setGaussConstraint(self, gaussConstraint=None):
        """Kept for the reference's call sequence; the device holds the solver's violation."""

    def initialize(self, s, direction, orientation):
        _check_beam(direction)
        g = _grid(s)
        self.particlesPerCell = int(math.floor(g.getLatticeSpacing(0) / g.getTemporalSpacing() + 0.5))
        added = ctypes.c_int64()
        zs, ze = ctypes.c_int(), ctypes.c_int()
        g._check(_lib.load().opx_cgc_create_particles(g.handle, int(orientation), ctypes.byref(added),
                                                      ctypes.byref(zs), ctypes.byref(ze)))
        self.added, self.zStart, self.zEnd = added.value, zs.value, ze.value


class CGCInitialCondition:
    """physics/initial/CGC/CGCInitialCondition.java:55-93 for the optimized CPIC types
    (super particles)."""

    def __init__(self):
        self.initialChargeDensity = None
        self.solver = LightConePoissonSolverImprovedFull()
        self.initialParticleCreator = None
        self.computeTadpole = False
        self.tadpoleFilename = "tadpole.txt"
        self.computeDipole = False
        self.dipoleFilename = "dipole.txt"

    def setInitialChargeDensity(self, rho):
        self.initialChargeDensity = rho

    def setPoissonSolver(self, solver):
        if not isinstance(solver, LightConePoissonSolverImprovedFull):
            raise NotImplementedError("only LightConePoissonSolverImprovedFull is built on the device")
        self.solver = solver

    def applyInitialCondition(self, s):
        sim_type = getattr(getattr(s, "settings", None), "simulation_type", "TemporalOptimizedCGCNGP")
        creators = {"TemporalCGCNGP": LightConeNGPParticleCreator,
                    "TemporalImplicitCGCNGP": LightConeNGPParticleCreator,
                    "TemporalOptimizedCGCNGP": LightConeNGPSuperParticleCreator,
                    "TemporalOptimizedImplicitCGCNGP": LightConeNGPSuperParticleCreator}
        if sim_type not in creators:
            # the reference throws a RuntimeException here (CGCInitialCondition.java:77-89)
            raise RuntimeError(f"CGCInitialCondition: simulation type {sim_type} not supported!")
        self.initialChargeDensity.initialize(s)
        direction = self.initialChargeDensity.getDirection()
        orientation = self.initialChargeDensity.getOrientation()
        self.solver.initialize(s)
        self.solver.solve(self.initialChargeDensity)
        wilson = WilsonLineObservables(s, self.solver, self.initialChargeDensity)
        if self.computeTadpole:
            wilson.computeTadpole(self.tadpoleFilename)
        if self.computeDipole:
            wilson.computeDipole(self.dipoleFilename)
        self.initialParticleCreator = creators[sim_type]()
        self.initialParticleCreator.setGaussConstraint()
        self.initialParticleCreator.initialize(s, direction, orientation)
        self.initialChargeDensity.clear()


_POISSON_SOLVERS = {"improved full": LightConePoissonSolverImprovedFull}


def from_yaml_cgc(y: dict) -> list:
    """YamlCGC.applyTo (ui/util/yaml/initial/YamlCGC.java): one CGCInitialCondition per MVModel
    entry (YamlMVModel.getInitialCondition, ui/util/yaml/initial/YamlMVModel.java:53-68)."""
    name = y.get("poissonSolver")
    key = str(name).lower().strip() if name is not None else None
    if key not in _POISSON_SOLVERS:
        raise NotImplementedError(f"Poisson solver {name!r}: only 'improved full' is built on the device")
    for other in ("MVModelCoherent", "NucleusCoherent", "Nucleus", "NucleusThick"):
        if y.get(other):
            raise NotImplementedError(f"{other} charge densities are not built")
    out = []
    for m in y.get("MVModel") or []:
        seed = m.get("randomSeed")
        rho = MVModel(int(m["direction"]), int(m["orientation"]), float(m["longitudinalLocation"]),
                      float(m["longitudinalWidth"]), float(m["mu"]), seed is not None, int(seed or 0),
                      float(m.get("ultravioletCutoffTransverse", 4.44)),
                      float(m.get("longitudinalCoherenceLength", 0.0)), float(m.get("infraredCoefficient", 0.0)))
        ic = CGCInitialCondition()
        ic.setInitialChargeDensity(rho)
        ic.setPoissonSolver(_POISSON_SOLVERS[key]())
        ic.computeTadpole = bool(y.get("computeTadpole", False))
        ic.computeDipole = bool(y.get("computeDipole", False))
        ic.tadpoleFilename = str(y.get("tadpoleFilename", "tadpole.txt"))
        ic.dipoleFilename = str(y.get("dipoleFilename", "dipole.txt"))
        out.append(ic)
    return out
