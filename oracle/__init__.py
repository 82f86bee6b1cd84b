"""FP64 CPU oracle for OpenPixi's FastTYMSolver path (ctypes wrapper).

TEST INFRASTRUCTURE ONLY: the checker against which the HIP product path is
compared. Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package. The product package ``openpixi_amd`` never
imports it.

The C sources (``opx_oracle.c``) restate the reference's Java algorithm; each
function cites the reference file:line it follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libopx_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_ubyte)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


class _Grid(ctypes.Structure):
    _fields_ = [
        ("nc", ctypes.c_int),
        ("dims", ctypes.c_int * 3),
        ("as_", ctypes.c_double * 3),
        ("at", ctypes.c_double),
        ("g", ctypes.c_double),
        ("n", ctypes.c_int64),
        ("cum", ctypes.c_int64 * 4),
        ("gsz", ctypes.c_int),
        ("asz", ctypes.c_int),
        ("U", _dp),
        ("Unext", _dp),
        ("E", _dp),
        ("J", _dp),
        ("rho", _dp),
        ("active", _u8p),
        ("evaluatable", _u8p),
        ("nthreads", ctypes.c_int),
    ]


class _SP(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_int),
        ("orientation", _ip),
        ("num", _ip),
        ("index_offset", _ip),
        ("per_plane", _ip),
        ("shift", _ip),
        ("ppc", _ip),
        ("q_offset", _i64p),
        ("q", _dp),
        ("total_charges", ctypes.c_int64),
    ]


class _P(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_int),
        ("direction", _ip),
        ("pos0", _dp),
        ("pos1", _dp),
        ("vel", _dp),
        ("Q0", _dp),
        ("Q1", _dp),
        ("Ulink", _dp),
        ("update_charge", _u8p),
    ]


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        gp = ctypes.POINTER(_Grid)
        spp = ctypes.POINTER(_SP)
        pp = ctypes.POINTER(_P)
        sig = {
            "orc_group_mult": (None, [ctypes.c_int, _dp, _dp, _dp]),
            "orc_group_adj": (None, [ctypes.c_int, _dp, _dp]),
            "orc_group_proj": (None, [ctypes.c_int, _dp, _dp]),
            "orc_alg_get_link": (None, [ctypes.c_int, _dp, _dp]),
            "orc_alg_act": (None, [ctypes.c_int, _dp, _dp, _dp]),
            "orc_alg_square": (ctypes.c_double, [ctypes.c_int, _dp]),
            "orc_alg_get": (ctypes.c_double, [ctypes.c_int, _dp, ctypes.c_int]),
            "orc_alg_from_coeffs": (None, [ctypes.c_int, _dp, _dp]),
            "orc_links_from_coeffs": (None, [ctypes.c_int, ctypes.c_int64, _dp, _dp]),
            "orc_algs_from_coeffs": (None, [ctypes.c_int, ctypes.c_int64, _dp, _dp]),
            "orc_grid_create": (gp, [ctypes.c_int, _ip, _dp, ctypes.c_double, ctypes.c_double]),
            "orc_grid_destroy": (None, [gp]),
            "orc_grid_set_threads": (None, [gp, ctypes.c_int]),
            "orc_shift": (ctypes.c_int64, [gp, ctypes.c_int64, ctypes.c_int, ctypes.c_int]),
            "orc_set_active_region": (None, [gp, _ip, _ip]),
            "orc_refine_list": (None, [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int]),
            "orc_restore_violation": (ctypes.c_double, [gp, _dp]),
            "orc_restore_correct": (None, [gp, _dp, ctypes.c_double]),
            "orc_gauss_restore": (ctypes.c_int, [gp, ctypes.c_double, ctypes.c_int, ctypes.c_double, _ip, _dp, _dp,
                                                 _dp]),
            "orc_unitarity_flags": (ctypes.c_int64, [gp, ctypes.c_double, _u8p]),
            "orc_set_evaluation_region": (None, [gp, _ip, _ip]),
            "orc_staple_sum": (None, [gp, ctypes.c_int64, ctypes.c_int, _dp]),
            "orc_step": (None, [gp, ctypes.c_double]),
            "orc_step_links": (None, [gp, ctypes.c_double]),
            "orc_store_fields": (None, [gp]),
            "orc_reset_current": (None, [gp]),
            "orc_reset_charge": (None, [gp]),
            "orc_plaquette": (None, [gp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, _dp]),
            "orc_get_B": (None, [gp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _dp]),
            "orc_gauss_constraint": (None, [gp, ctypes.c_int64, _dp]),
            "orc_gauss_constraint_squared": (ctypes.c_double, [gp, ctypes.c_int64]),
            "orc_measure": (None, [gp, _dp]),
            "orc_projected_energy_density": (None, [gp, ctypes.c_int, _dp, _dp, _dp, _dp]),
            "orc_projected_poynting": (ctypes.c_int, [gp, ctypes.c_int, _dp, _dp]),
            "orc_implicit_step": (None, [gp, ctypes.c_double, ctypes.c_int, ctypes.c_double]),
            "orc_regulate_charge_gaussian": (None, [_ip, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_double, _dp]),
            "orc_poisson_2d": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp]),
            "orc_mv_charge": (None, [gp] + [ctypes.c_double] * 6 + [_dp, _dp]),
            "orc_wilson_segment": (None, [gp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp]),
            "orc_lightcone_solve": (None, [gp, _dp, ctypes.c_int, _dp]),
            "orc_lightcone_solve_vt": (None, [gp, _dp, ctypes.c_int, _dp, _dp]),
            "orc_lightcone_particles": (None, [gp, _dp] + [ctypes.c_int] * 4 + [_dp] * 4),
            "orc_wilson_tadpole": (ctypes.c_double, [ctypes.c_int, ctypes.c_long, _dp]),
            "orc_wilson_dipole": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         _dp, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_long)]),
            "orc_lightcone_block": (None, [gp, _dp, _ip, _ip]),
            "orc_lightcone_superparticles": (None, [gp, _dp] + [ctypes.c_int] * 5 + [_ip] * 6 + [_dp]),
            "orc_group_log": (None, [ctypes.c_int, _dp, _dp]),
            "orc_group_pow": (None, [ctypes.c_int, _dp, ctypes.c_double, _dp]),
            "orc_gauge_transform": (None, [gp, _dp]),
            "orc_sp_current_offset": (ctypes.c_int, [ctypes.c_int] * 6),
            "orc_sp_ngp_offset": (ctypes.c_int, [ctypes.c_int] * 6),
            "orc_sp_needs_update": (ctypes.c_int, [ctypes.c_int] * 4),
            "orc_sp_interpolate_to_grid": (None, [gp, spp, ctypes.c_int]),
            "orc_sp_interpolate_chargedensity": (None, [gp, spp, ctypes.c_int]),
            "orc_nearest_grid_point": (ctypes.c_int, [ctypes.c_double, ctypes.c_double]),
            "orc_sim_initialize": (None, [gp, spp, pp, ctypes.c_double]),
            "orc_sim_step": (None, [gp, spp, pp, ctypes.c_double, _ip]),
            "orc_time_steps": (ctypes.c_double, [gp, ctypes.c_int, ctypes.c_int]),
            "orc_adj_exp_vjp": (None, [_dp, _dp, _dp]),
            "orc_adj_force_vjp": (None, [gp, _dp, _dp, ctypes.c_double, _dp]),
            "orc_adj_bsq": (ctypes.c_double, [gp, _dp, ctypes.c_int, _dp]),
            "orc_adj_loss": (ctypes.c_double, [gp, _dp, _dp, _dp, ctypes.c_int]),
            "orc_adjoint_gradient": (ctypes.c_int, [gp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def gsize(nc: int) -> int:
    return 4 if nc == 2 else 18


def asize(nc: int) -> int:
    return 3 if nc == 2 else 9


# ---- element-level helpers (operate on 1-D float64 arrays) ----

def group_mult(nc, a, b):
    o = np.zeros(gsize(nc))
    lib().orc_group_mult(nc, _ptr(np.ascontiguousarray(a, np.float64)), _ptr(np.ascontiguousarray(b, np.float64)), _ptr(o))
    return o


def group_adj(nc, a):
    o = np.zeros(gsize(nc))
    lib().orc_group_adj(nc, _ptr(np.ascontiguousarray(a, np.float64)), _ptr(o))
    return o


def group_proj(nc, a):
    o = np.zeros(asize(nc))
    lib().orc_group_proj(nc, _ptr(np.ascontiguousarray(a, np.float64)), _ptr(o))
    return o


def alg_get_link(nc, v):
    o = np.zeros(gsize(nc))
    lib().orc_alg_get_link(nc, _ptr(np.ascontiguousarray(v, np.float64)), _ptr(o))
    return o


def alg_act(nc, v, g):
    o = np.zeros(asize(nc))
    lib().orc_alg_act(nc, _ptr(np.ascontiguousarray(v, np.float64)), _ptr(np.ascontiguousarray(g, np.float64)), _ptr(o))
    return o


def group_log(nc, u):
    """GroupElement.getAlgebraElement: H (reference algebra storage) with U = getLink(H)."""
    o = np.zeros(asize(nc))
    lib().orc_group_log(nc, _ptr(np.ascontiguousarray(u, np.float64)), _ptr(o))
    return o


def group_pow(nc, u, x):
    """GroupElement.pow(x) = getLink(x getAlgebraElement())."""
    o = np.zeros(gsize(nc))
    lib().orc_group_pow(nc, _ptr(np.ascontiguousarray(u, np.float64)), float(x), _ptr(o))
    return o


def wilson_tadpole(nc, VT):
    """WilsonLineObservables.computeTadpole's value: mean Re tr V over the transverse plane."""
    VT = np.ascontiguousarray(VT, np.float64)
    return float(lib().orc_wilson_tadpole(nc, VT.shape[0], _ptr(VT)))


def wilson_dipole(nc, n1, n2, a1, a2, VT, nbins):
    """WilsonLineObservables.computeDipole's bins: (mean tr(V_x V_y^+) per bin, pair counts)."""
    VT = np.ascontiguousarray(VT, np.float64)
    corr = np.zeros(nbins)
    counts = np.zeros(nbins, np.int64)
    lib().orc_wilson_dipole(nc, n1, n2, a1, a2, _ptr(VT), nbins, _ptr(corr),
                            counts.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
    return corr, counts


def regulate_charge_gaussian(dims, uvt, long_width, ir, aT, aL, rho):
    """FourierFunctions.regulateChargeDensityGaussian, beam along x0, on a (cells,) array."""
    r = np.ascontiguousarray(rho, np.float64).copy()
    lib().orc_regulate_charge_gaussian((ctypes.c_int * 3)(*dims), uvt, long_width, ir, aT, aL, _ptr(r))
    return r


def poisson_2d(n1, n2, aT, rho):
    """FourierFunctions.solvePoisson2D on one (n1 * n2,) sheet."""
    r = np.ascontiguousarray(rho, np.float64)
    phi = np.zeros(n1 * n2)
    lib().orc_poisson_2d(n1, n2, aT, _ptr(r), _ptr(phi))
    return phi


def alg_square(nc, v):
    return lib().orc_alg_square(nc, _ptr(np.ascontiguousarray(v, np.float64)))


def alg_coeffs(nc, v):
    """Generator coefficients of an algebra element in reference storage."""
    v = np.ascontiguousarray(v, np.float64)
    return np.array([lib().orc_alg_get(nc, _ptr(v), j) for j in range(nc * nc - 1)])


def alg_from_coeffs(nc, c):
    o = np.zeros(asize(nc))
    lib().orc_alg_from_coeffs(nc, _ptr(np.ascontiguousarray(c, np.float64)), _ptr(o))
    return o


def refine_list(q, ppc, skip, order, sweeps):
    """The light-cone creators' refinement sweeps on one periodic list q[nn][A] (returns a copy)."""
    q = np.array(q, np.float64, order="C", copy=True)
    nn, A = q.shape
    lib().orc_refine_list(_ptr(q.reshape(-1)), nn, A, int(ppc), int(skip), int(order), int(sweeps))
    return q


def links_from_coeffs(nc, coeffs):
    """Group elements exp(i c_a t_a) for an array (..., nc*nc-1) of coefficients."""
    c = np.ascontiguousarray(coeffs, np.float64)
    shape = c.shape[:-1]
    cnt = int(np.prod(shape)) if shape else 1
    out = np.zeros(shape + (gsize(nc),))
    lib().orc_links_from_coeffs(nc, cnt, _ptr(c.reshape(-1)), _ptr(out.reshape(-1)))
    return out


def algs_from_coeffs(nc, coeffs):
    """Reference algebra storage for an array (..., nc*nc-1) of coefficients."""
    c = np.ascontiguousarray(coeffs, np.float64)
    shape = c.shape[:-1]
    cnt = int(np.prod(shape)) if shape else 1
    out = np.zeros(shape + (asize(nc),))
    lib().orc_algs_from_coeffs(nc, cnt, _ptr(c.reshape(-1)), _ptr(out.reshape(-1)))
    return out


class OracleGrid:
    """A reference-layout lattice (AoS per cell) advanced by the FP64 oracle."""

    def __init__(self, nc, dims, spacing=(1.0, 1.0, 1.0), at=0.25, g=1.0, threads=1):
        L = lib()
        self.nc = nc
        self.dims = tuple(int(d) for d in dims)
        self.n = int(np.prod(self.dims))
        d = (ctypes.c_int * 3)(*self.dims)
        s = (ctypes.c_double * 3)(*[float(x) for x in spacing])
        self._g = L.orc_grid_create(nc, d, s, float(at), float(g))
        if not self._g:
            raise ValueError("unsupported number of colours")
        self.at = float(at)
        self.spacing = tuple(float(x) for x in spacing)
        self.coupling = float(g)
        self.steps_done = 0
        self._sp = None
        self._sp_keep = None
        self._p = None
        self._p_keep = None
        if threads > 1:
            L.orc_grid_set_threads(self._g, int(threads))

    def __del__(self):
        try:
            if getattr(self, "_g", None):
                lib().orc_grid_destroy(self._g)
                self._g = None
        except Exception:
            pass

    # field views (refresh after store_fields, which swaps pointers)
    def _view(self, name, per):
        ptr = getattr(self._g.contents, name)
        return np.ctypeslib.as_array(ptr, shape=(self.n * per,))

    @property
    def U(self):
        return self._view("U", 3 * gsize(self.nc)).reshape(self.n, 3, gsize(self.nc))

    @property
    def Unext(self):
        return self._view("Unext", 3 * gsize(self.nc)).reshape(self.n, 3, gsize(self.nc))

    @property
    def E(self):
        return self._view("E", 3 * asize(self.nc)).reshape(self.n, 3, asize(self.nc))

    @property
    def J(self):
        return self._view("J", 3 * asize(self.nc)).reshape(self.n, 3, asize(self.nc))

    @property
    def rho(self):
        return self._view("rho", asize(self.nc)).reshape(self.n, asize(self.nc))

    def set_active_region(self, lo, hi):
        lib().orc_set_active_region(self._g, (ctypes.c_int * 3)(*lo), (ctypes.c_int * 3)(*hi))

    def set_evaluation_region(self, lo, hi):
        lib().orc_set_evaluation_region(self._g, (ctypes.c_int * 3)(*lo), (ctypes.c_int * 3)(*hi))

    def shift(self, index, d, o):
        return lib().orc_shift(self._g, index, d, o)

    def staple_sum(self, index, d):
        o = np.zeros(gsize(self.nc))
        lib().orc_staple_sum(self._g, index, d, _ptr(o))
        return o

    def step(self, at=None):
        lib().orc_step(self._g, self.at if at is None else at)

    def step_links(self, at=None):
        lib().orc_step_links(self._g, self.at if at is None else at)

    def store_fields(self):
        lib().orc_store_fields(self._g)

    def reset_current(self):
        lib().orc_reset_current(self._g)

    def reset_charge(self):
        lib().orc_reset_charge(self._g)

    def plaquette(self, index, d1, d2, o1, o2, t):
        o = np.zeros(gsize(self.nc))
        lib().orc_plaquette(self._g, index, d1, d2, o1, o2, t, _ptr(o))
        return o

    def get_B(self, index, d, t):
        o = np.zeros(asize(self.nc))
        lib().orc_get_B(self._g, index, d, t, _ptr(o))
        return o

    def gauss_constraint(self, index):
        o = np.zeros(asize(self.nc))
        lib().orc_gauss_constraint(self._g, index, _ptr(o))
        return o

    def gauss_constraint_squared(self, index):
        return lib().orc_gauss_constraint_squared(self._g, index)

    def measure(self):
        """[E2_0,E2_1,E2_2,B2_0,B2_1,B2_2,gauss_mean_sq,total_charge_squared]."""
        o = np.zeros(8)
        lib().orc_measure(self._g, _ptr(o))
        return o

    def projected_energy_density(self, direction):
        n = self.dims[direction]
        arrs = [np.zeros(n) for _ in range(4)]
        lib().orc_projected_energy_density(self._g, direction, *[_ptr(a) for a in arrs])
        return arrs  # eT_el, eT_mag, eL_el, eL_mag

    def implicit_step(self, at=None, iterations=0, damping=0.0):
        """ImplicitTYMSolver.step (physics/fields/ImplicitTYMSolver.java:35-58)."""
        lib().orc_implicit_step(self._g, self.at if at is None else at, int(iterations), float(damping))

    # ---- CGC initial condition, beam along x0 (opx_oracle_cgc.c) ----
    def mv_charge(self, location, width, mu, uvt, long_width, ir, noise):
        """MVModel.initialize with unit Gaussian noise[(components, cells)]; returns rho (cells, A)."""
        noise = np.ascontiguousarray(noise, np.float64).reshape(-1)
        assert noise.size == (self.nc * self.nc - 1) * self.n
        rho = np.zeros((self.n, asize(self.nc)))
        lib().orc_mv_charge(self._g, location, width, mu, uvt, long_width, ir, _ptr(noise), _ptr(rho))
        return rho

    def lightcone_solve(self, rho, orientation):
        """LightConePoissonSolverImprovedFull.solve: sets U, Unext, E; returns the Gauss violation (cells, A).
        The Wilson lines on the plane behind the nucleus (getV) are kept in self.VT."""
        rho = np.ascontiguousarray(rho, np.float64)
        gauss = np.zeros((self.n, asize(self.nc)))
        self.VT = np.zeros((self.dims[1] * self.dims[2], gsize(self.nc)))
        lib().orc_lightcone_solve_vt(self._g, _ptr(rho), int(orientation), _ptr(gauss), _ptr(self.VT))
        return gauss

    def lightcone_superparticles(self, gauss, orientation, subdivisions):
        """LightConeNGPSuperParticleCreator.initializeParticles: (descriptor dict, charges list per super particle)."""
        gauss = np.ascontiguousarray(gauss, np.float64)
        zs, bw = ctypes.c_int(), ctypes.c_int()
        lib().orc_lightcone_block(self._g, _ptr(gauss), ctypes.byref(zs), ctypes.byref(bw))
        ppc = int(round(self.spacing[0] / self.at))
        sub = min(int(subdivisions), bw.value)
        cnt = sub * ppc
        arrs = {k: np.zeros(cnt, np.int32) for k in ("orientation", "num", "index_offset", "per_plane", "shift", "ppc")}
        T = self.dims[1] * self.dims[2]
        q = np.zeros((bw.value * T * ppc, asize(self.nc)))
        lib().orc_lightcone_superparticles(self._g, _ptr(gauss), int(orientation), zs.value, bw.value, sub, ppc,
                                           *[arrs[k].ctypes.data_as(_ip) for k in ("orientation", "num", "index_offset",
                                                                                    "per_plane", "shift", "ppc")],
                                           _ptr(q))
        charges, pos = [], 0
        for s_ in range(cnt):
            charges.append(q[pos:pos + arrs["num"][s_]])
            pos += arrs["num"][s_]
        return arrs, charges, (zs.value, bw.value)

    def lightcone_particles(self, gauss, orientation):
        """LightConeNGPParticleCreator.initializeParticles: (direction, pos0, pos1, vel, charges) of the
        per-particle CGC particles in list order, and the block (z_start, z_end)."""
        gauss = np.ascontiguousarray(gauss, np.float64)
        zs, bw = ctypes.c_int(), ctypes.c_int()
        lib().orc_lightcone_block(self._g, _ptr(gauss), ctypes.byref(zs), ctypes.byref(bw))
        ze = zs.value + bw.value
        ppc = int(np.floor(self.spacing[0] / self.at + 0.5))
        cnt = (ze - zs.value + 1) * self.dims[1] * self.dims[2] * ppc
        p0, p1, v = np.zeros((cnt, 3)), np.zeros((cnt, 3)), np.zeros((cnt, 3))
        q = np.zeros((cnt, asize(self.nc)))
        lib().orc_lightcone_particles(self._g, _ptr(gauss), int(orientation), zs.value, ze, ppc, _ptr(p0), _ptr(p1),
                                      _ptr(v), _ptr(q))
        return np.zeros(cnt, np.int32), p0, p1, v, q, (zs.value, ze)

    def projected_poynting(self, direction):
        """Longitudinal Poynting vector per slice (averaged, time averaged); SU(2) only."""
        n = self.dims[direction]
        avg, tavg = np.zeros(n), np.zeros(n)
        if lib().orc_projected_poynting(self._g, direction, _ptr(avg), _ptr(tavg)) != 0:
            raise RuntimeError("SU(3) code has not been tested yet.")
        return avg, tavg

    # ---- GaussConstraintRestoration, UnitarityTester (opx_oracle_gauss.c) ----
    def restore_violation(self):
        """(mean violation, gv as n group elements) of GaussConstraintRestoration.computeGaussViolation."""
        gv = np.zeros((self.n, gsize(self.nc)))
        v = lib().orc_restore_violation(self._g, _ptr(gv.reshape(-1)))
        return v, gv

    def restore_correct(self, gv, gamma):
        lib().orc_restore_correct(self._g, _ptr(np.ascontiguousarray(gv, np.float64).reshape(-1)), float(gamma))

    def gauss_restore(self, gamma, max_iterations, absolute_value):
        """iterateRestorationAlgorithm: (kept corrections, reason, before, after, history)."""
        reason = np.zeros(1, np.int32)
        b, a = np.zeros(1), np.zeros(1)
        hist = np.full(max(int(max_iterations), 0) + 1, np.nan)
        it = lib().orc_gauss_restore(self._g, float(gamma), int(max_iterations), float(absolute_value),
                                     reason.ctypes.data_as(_ip), _ptr(b), _ptr(a), _ptr(hist))
        if it < 0:
            raise RuntimeError("Use getGridStep(int direction).")
        return it, int(reason[0]), float(b[0]), float(a[0]), hist[: it + 1]

    def unitarity_flags(self, tol=1e-8):
        flags = np.zeros(self.n * 3, np.uint8)
        cnt = lib().orc_unitarity_flags(self._g, float(tol), flags.ctypes.data_as(_u8p))
        return int(cnt), flags.reshape(self.n, 3)

    def gauge_transform(self, g_cells):
        g_cells = np.ascontiguousarray(g_cells, np.float64).reshape(-1)
        assert g_cells.size == self.n * gsize(self.nc)
        lib().orc_gauge_transform(self._g, _ptr(g_cells))

    # ---- super particles ----
    def set_superparticles(self, orientation, num, index_offset, per_plane, shift, ppc, charges):
        """charges: list of (num_p, A) float64 arrays in reference algebra storage."""
        cnt = len(orientation)
        arrs = {k: np.ascontiguousarray(v, np.int32) for k, v in dict(
            orientation=orientation, num=num, index_offset=index_offset, per_plane=per_plane,
            shift=shift, ppc=ppc).items()}
        qoff = np.zeros(cnt, np.int64)
        acc = 0
        for i in range(cnt):
            qoff[i] = acc
            acc += int(num[i])
        q = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float64).reshape(-1) for c in charges])
                                 if cnt else np.zeros(0))
        sp = _SP()
        sp.count = cnt
        for k in ("orientation", "num", "index_offset", "per_plane", "shift", "ppc"):
            setattr(sp, k, arrs[k].ctypes.data_as(_ip))
        sp.q_offset = qoff.ctypes.data_as(_i64p)
        sp.q = q.ctypes.data_as(_dp)
        sp.total_charges = acc
        self._sp = sp
        self._sp_keep = (arrs, qoff, q)

    def superparticle_charges(self):
        return self._sp_keep[2].reshape(-1, asize(self.nc))

    def sp_interpolate_to_grid(self, t):
        lib().orc_sp_interpolate_to_grid(self._g, ctypes.byref(self._sp), t)

    def sp_interpolate_chargedensity(self, t):
        lib().orc_sp_interpolate_chargedensity(self._g, ctypes.byref(self._sp), t)

    # ---- per-particle NGP ----
    def set_particles(self, direction, pos0, vel, Q0):
        cnt = len(direction)
        A, G = asize(self.nc), gsize(self.nc)
        keep = dict(
            direction=np.ascontiguousarray(direction, np.int32),
            pos0=np.ascontiguousarray(pos0, np.float64).reshape(-1).copy(),
            pos1=np.zeros(cnt * 3),
            vel=np.ascontiguousarray(vel, np.float64).reshape(-1).copy(),
            Q0=np.ascontiguousarray(Q0, np.float64).reshape(-1).copy(),
            Q1=np.zeros(cnt * A),
            Ulink=np.zeros(cnt * G),
            update_charge=np.zeros(cnt, np.uint8),
        )
        p = _P()
        p.count = cnt
        p.direction = keep["direction"].ctypes.data_as(_ip)
        for k in ("pos0", "pos1", "vel", "Q0", "Q1", "Ulink"):
            setattr(p, k, keep[k].ctypes.data_as(_dp))
        p.update_charge = keep["update_charge"].ctypes.data_as(_u8p)
        self._p = p
        self._p_keep = keep

    def particle_state(self):
        """Current (pos0, pos1, Q0, Q1) as stored (pointer swaps are value swaps here)."""
        k = self._p_keep
        A = asize(self.nc)
        return (k["pos0"].reshape(-1, 3), k["pos1"].reshape(-1, 3), k["Q0"].reshape(-1, A), k["Q1"].reshape(-1, A))

    # ---- driver ----
    def initialize(self):
        lib().orc_sim_initialize(self._g, ctypes.byref(self._sp) if self._sp else None,
                                 ctypes.byref(self._p) if self._p else None, self.at)

    def sim_step(self, n=1):
        c = ctypes.c_int(self.steps_done)
        for _ in range(n):
            lib().orc_sim_step(self._g, ctypes.byref(self._sp) if self._sp else None,
                               ctypes.byref(self._p) if self._p else None, self.at, ctypes.byref(c))
        self.steps_done = c.value

    def time_steps(self, steps, warmup=0):
        return lib().orc_time_steps(self._g, steps, warmup)

    def set_threads(self, n):
        lib().orc_grid_set_threads(self._g, int(n))

    # ---- SU(2) adjoint (opx_oracle_adjoint.c) ----
    def adj_force_vjp(self, U, Ebar, scale=1.0):
        """(dF/dU)^T Ebar * scale for links U (n,3,4) and cotangent Ebar (n,3,3); F = proj(U S)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        Eb = np.ascontiguousarray(Ebar, dtype=np.float64)
        out = np.zeros((self.n, 3, 4))
        lib().orc_adj_force_vjp(self._g, _ptr(U), _ptr(Eb), float(scale), _ptr(out))
        return out

    def adj_bsq(self, U, direction, grad=True):
        """Magnetic part of the loss for links U and (optionally) its gradient."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        out = np.zeros((self.n, 3, 4))
        v = lib().orc_adj_bsq(self._g, _ptr(U), int(direction), _ptr(out) if grad else None)
        return (v, out) if grad else v

    def adj_loss(self, direction):
        """Loss of the current state: E, U (= U^{N-1}) and Unext (= U^N)."""
        U = np.ascontiguousarray(self.U)
        Un = np.ascontiguousarray(self.Unext)
        E = np.ascontiguousarray(self.E)
        return lib().orc_adj_loss(self._g, _ptr(U), _ptr(Un), _ptr(E), int(direction))

    def adjoint_gradient(self, steps, direction):
        """(loss, dL/dE0) for the initial state in U (links) and E; leaves the final state."""
        loss = ctypes.c_double()
        grad = np.zeros((self.n, 3, 3))
        rc = lib().orc_adjoint_gradient(self._g, int(steps), int(direction), ctypes.byref(loss), _ptr(grad))
        if rc != 0:
            raise ValueError(f"orc_adjoint_gradient failed ({rc})")
        return loss.value, grad


def adj_exp_vjp(w, Mbar):
    o = np.zeros(3)
    lib().orc_adj_exp_vjp(_ptr(np.ascontiguousarray(w, dtype=np.float64)),
                          _ptr(np.ascontiguousarray(Mbar, dtype=np.float64)), _ptr(o))
    return o


def sp_current_offset(o, off, pp, k, ppc, t):
    return lib().orc_sp_current_offset(o, off, pp, k, ppc, t)


def sp_ngp_offset(o, off, pp, k, ppc, t):
    return lib().orc_sp_ngp_offset(o, off, pp, k, ppc, t)


def sp_needs_update(o, k, ppc, t):
    return bool(lib().orc_sp_needs_update(o, k, ppc, t))


def nearest_grid_point(pos, a):
    return lib().orc_nearest_grid_point(pos, a)
