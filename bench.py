#!/usr/bin/env python3
"""Benchmark: SU(3) link-updates/s of the temporal-gauge leapfrog on a 128^3
periodic lattice, FP32 links/E with FP64 reunitarization every 10 steps versus
plain FP64, both on one MI355X, compared at equal Gauss residual.

Workload (the SU(3) 128^3 configuration): links exp(i A_a t_a) with
A_a ~ Gaussian(0, 0.1), E coefficients ~ Gaussian(0, 0.05) (seeded Philox) then
projected onto Gauss's law (conjugate gradients, double precision),
a = g = 1, dt = 0.25, the initial half step U(-dt/2) -> U(+dt/2), then 500
leapfrog steps. A step is storeFields + the fused E-field/link-update kernel;
every 10th step of the FP32 form also reprojects the links onto SU(3) in double
precision. Inputs are resident in HBM before timing.

Prints one JSON line. `value` = link-updates/s of the FP32+reunit form; each
form reports its Gauss residual and energy before/after, and the FP32 form its
relative drift from the FP64 form for every observable.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DIMS = (128, 128, 128)
NC = 3
DT = 0.25
REUNIT_EVERY = 10
# algorithmic HBM bytes per link update of the fused kernel (read U, read E, write E, write Unext)
# in the stored formats: FP32 SU(3) links hold their first two rows (12 reals), FP64 links all 18
BYTES_PER_LINK = {"fp32": 48 + 32 + 32 + 48, "fp64": 144 + 64 + 64 + 144}
# HBM bytes per launch of the fused kernel forms on the 128^3 lattice, from separate profiling
# runs (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, over `bench.py --probe`; KiB per dispatch).
# FETCH_SIZE is scaled by 2.0 and WRITE_SIZE by 1.0: the factors that make the counters read the
# known bytes of a 1 GiB device-to-device copy. These are recorded figures, not measured by this
# run: the JSON line carries them as roofline.traffic with roofline.traffic_source naming the
# profile summary. Keyed by (precision, kernel form, tile, lattice): (FETCH_SIZE KiB, WRITE_SIZE KiB, file).
_PROFILE = "profiles/k_cell_update_su3_128cube.md"
PROFILED_TRAFFIC = {("fp32", "cell", 0, DIMS): (265956.0, 491521.0, _PROFILE),
                    ("fp32", "cell", 1, DIMS): (266075.0, 491520.0, _PROFILE),
                    ("fp32", "cell", 2, DIMS): (280926.0, 491520.0, _PROFILE),
                    ("fp32", "link", 0, DIMS): (263227.0, 491520.0, _PROFILE),
                    ("fp64", "cell", 0, DIMS): (1476429.0, 1363132.0, _PROFILE),
                    ("fp64", "cell", 1, DIMS): (1169774.0, 1355513.0, _PROFILE),
                    ("fp64", "cell", 2, DIMS): (1079910.0, 1348559.0, _PROFILE),
                    ("fp64", "link", 0, DIMS): (700078.0, 1298745.0, _PROFILE)}
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
# VALU instructions per launch (rocprofv3 --pmc SQ_INSTS_VALU, same probe runs): with the launch
# time this run measures, the share of the launch the VALU needs to issue them at the 2.4 GHz peak
# clock (4 cycles per wave64 instruction on each of 1024 SIMDs).
PROFILED_VALU = {("fp32", "cell", 0, DIMS): (121012224.0, _PROFILE),
                 ("fp32", "cell", 1, DIMS): (121012224.0, _PROFILE),
                 ("fp32", "cell", 2, DIMS): (121012224.0, _PROFILE),
                 ("fp32", "link", 0, DIMS): (144932352.0, _PROFILE),
                 ("fp64", "cell", 0, DIMS): (166137634.0, _PROFILE),
                 ("fp64", "cell", 1, DIMS): (166426987.0, _PROFILE),
                 ("fp64", "cell", 2, DIMS): (166428236.0, _PROFILE),
                 ("fp64", "link", 0, DIMS): (163118117.0, _PROFILE)}
SIMDS, PEAK_CLOCK_GHZ = 1024, 2.4


def profiled_traffic(precision, kernel_config, dims):
    """(bytes per launch, source description) of a recorded counter run, or (None, None)."""
    rec = PROFILED_TRAFFIC.get((precision, kernel_config["kernel"], kernel_config["tile"], tuple(dims)))
    if rec is None:
        return None, None
    fetch, write, src = rec
    return ((fetch * FETCH_SCALE + write * WRITE_SCALE) * 1024,
            f"{src}: rocprofv3 --pmc FETCH_SIZE {fetch:.0f} KiB x {FETCH_SCALE} + WRITE_SIZE {write:.0f} KiB x "
            f"{WRITE_SCALE} per dispatch (scales from a 1 GiB device copy); recorded, not measured by this run")


def valu_issue(precision, kernel_config, dims, kernel_ms):
    """VALU issue time of the recorded instruction count at peak clock, against this run's launch."""
    rec = PROFILED_VALU.get((precision, kernel_config["kernel"], kernel_config["tile"], tuple(dims)))
    if rec is None:
        return None
    insts, src = rec
    ms = insts * 4 / SIMDS / (PEAK_CLOCK_GHZ * 1e9) * 1e3
    return {"instructions": insts, "issue_ms_at_peak_clock": ms, "frac_of_launch": ms / kernel_ms,
            "source": f"{src}: rocprofv3 --pmc SQ_INSTS_VALU per dispatch; recorded, not measured by this run"}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=500)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--kernel-reps", type=int, default=50, help="back-to-back launches of the fused kernel after the timed region (kernel_ms_back_to_back)")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--dims", type=int, nargs=3, default=list(DIMS))
    p.add_argument("--forms", default="fp32,fp64")
    p.add_argument("--no-autotune", action="store_true")
    p.add_argument("--no-cpic", action="store_true", help="skip the SU(2) CPIC collision line")
    p.add_argument("--cpic-steps", type=int, default=1000)
    p.add_argument("--no-batch", action="store_true", help="skip the batched-events line")
    p.add_argument("--batch-events", type=int, default=64)
    p.add_argument("--batch-steps", type=int, default=300)
    p.add_argument("--no-adjoint", action="store_true", help="skip the adjoint line")
    p.add_argument("--no-implicit", action="store_true", help="skip the implicit-solver line")
    p.add_argument("--implicit-steps", type=int, default=20)
    p.add_argument("--no-cgc", action="store_true", help="skip the CGC initial-condition line")
    p.add_argument("--adjoint-steps", type=int, default=200)
    p.add_argument("--fd-directions", type=int, default=256)
    p.add_argument("--probe-kernels", default="auto",
                   help="kernel configurations for --probe: comma list of link | stream:TILE:CHUNKS | auto")
    p.add_argument("--probe-colors", type=int, default=NC, choices=(2, 3),
                   help="gauge group of the --probe lattices: SU(2) or SU(3)")
    p.add_argument("--probe", action="store_true",
                   help="counter-collection run: a few fused-kernel launches per form plus a 1 GiB device copy "
                        "(the calibration kernel for FETCH_SIZE/WRITE_SIZE); prints no JSON line")
    return p.parse_args()


def parse_kernels(spec):
    """"link,stream:1:16" -> [("link", None, None), ("stream", 1, 16)]; "auto" keeps the default."""
    out = []
    for item in spec.split(","):
        parts = item.split(":")
        out.append((parts[0], int(parts[1]) if len(parts) > 1 else None, int(parts[2]) if len(parts) > 2 else None))
    return out


def probe(dims, forms, kernels, reps=5, nc=NC):
    """Launches for hardware-counter runs: `reps` fused updates per form and kernel
    configuration (SU(nc) links), then a 1 GiB float32 device-to-device copy of known size."""
    import torch
    from openpixi_amd import _lib

    L = _lib.load()
    for f in forms:
        g = make_grid(f, dims, nc=nc)
        for k in kernels:
            g.setKernel(*k)
            for _ in range(reps):
                g._check(L.opx_store_fields(g.handle))
                g._check(L.opx_step(g.handle, DT))
            g.sync()
            print(f"# probe {f} {g.kernelConfig()}: {reps} fused launches", file=sys.stderr, flush=True)
        g.close()
    a = torch.empty(2 ** 28, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"# probe copy: {reps} x 1 GiB read + 1 GiB write", file=sys.stderr, flush=True)


def make_grid(precision, dims, seed=0, nc=NC):
    from openpixi_amd import Grid, inputs

    n = int(np.prod(dims))
    g = Grid(nc, dims, 1.0, DT, 1.0, precision)
    A = inputs.random_links_coefficients(n, nc, 0.1, seed)
    g.setLinksFromCoefficients(A)
    del A
    Ec = inputs.random_efield_coefficients(n, nc, 0.05, seed)
    g.setEFromCoefficients(Ec)
    del Ec
    return g


def ready_events(stream, count):
    """`count` timing events, each recorded once on `stream` here. The first record of a new event
    creates it, which can hold the host for milliseconds; between the synchronization that opens a
    timed region and its first launch that leaves the GPU idle, and after an idle millisecond the
    GPU's clock dips over its next tens of milliseconds of load (up to 20 % slower launches)."""
    import torch

    evs = [torch.cuda.Event(enable_timing=True) for _ in range(count)]
    for e in evs:
        e.record(stream)
    return evs


def time_form(precision, dims, steps, warmup, kernel_reps, stream, autotune=True):
    import torch
    from openpixi_amd import FieldMeasurements, _lib

    L = _lib.load()
    g = make_grid(precision, dims)
    g._check(L.opx_set_stream(g.handle, ctypes_void(stream.cuda_stream)))
    cg_iters, gauss_raw, gauss_proj = g.projectGauss(tol=1e-10, max_iter=4000)
    reunit = REUNIT_EVERY if precision == "fp32" else 0
    # initial half step U(-dt/2) -> U(+dt/2), as Simulation.initialize
    g._check(L.opx_step_links(g.handle, DT))
    m0 = FieldMeasurements().measure(g)
    ev0, ev1 = ready_events(stream, 2)
    # the autotuner times its candidates on copies of the fields; it runs last before the warm-up
    # steps, so that the GPU is under load from then until the timed region
    tune_ms = g.autotune(reps=3) if autotune else None
    kcfg = g.kernelConfig()
    fused = kcfg["kernel"] == "cell"
    per_launch = not fused and reunit != 0
    step_flat = ready_events(stream, 2 * steps) if per_launch else []
    step_ev = list(zip(step_flat[0::2], step_flat[1::2]))
    g._check(L.opx_leapfrog(g.handle, warmup, reunit))
    torch.cuda.synchronize()
    # The timed region: `steps` calls of opx_leapfrog(1 step), a HIP event pair on the kernels'
    # stream around the whole region. A step is one launch of the fused kernel (the cell form
    # reprojects the links inside the kernel on every 10th step), so the region's average is the
    # kernel's average launch duration, gaps between launches included. An event recorded between
    # two launches holds the second back until the first has finished and its timestamp is written
    # (about 10 us per step in a kernel trace), so per-launch events are recorded only for a form
    # that runs the reprojection as a pass of its own, whose steps then leave out of the average.
    # Its events exist before the warm-up (ready_events), so nothing but the launches follows the
    # synchronization.
    t0 = time.perf_counter()
    ev0.record(stream)
    for i in range(steps):
        if per_launch:
            step_ev[i][0].record(stream)
        g._check(L.opx_leapfrog(g.handle, 1, reunit))
        if per_launch:
            step_ev[i][1].record(stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1)
    if per_launch:
        per_step = [a.elapsed_time(b) for a, b in step_ev]
        counted = [t for i, t in enumerate(per_step) if (warmup + i + 1) % reunit != 0]
        kms = sum(counted) / len(counted)
    else:
        kms = ms / steps
    m1 = FieldMeasurements().measure(g)
    # the fused update kernel alone, back to back on the same stream, after the timed region;
    # then the same launches one by one between event pairs, for their spread
    k0 = torch.cuda.Event(enable_timing=True)
    k1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    k0.record(stream)
    for _ in range(kernel_reps):
        g._check(L.opx_store_fields(g.handle))
        g._check(L.opx_step(g.handle, DT))
    k1.record(stream)
    torch.cuda.synchronize()
    kms_b2b = k0.elapsed_time(k1) / kernel_reps
    one = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(kernel_reps)]
    for a, b in one:
        g._check(L.opx_store_fields(g.handle))
        a.record(stream)
        g._check(L.opx_step(g.handle, DT))
        b.record(stream)
    torch.cuda.synchronize()
    single = [a.elapsed_time(b) for a, b in one]
    n = int(np.prod(dims))
    links = 3 * n
    energy0 = float(m0[0:6].sum())
    energy1 = float(m1[0:6].sum())
    res = {
        "ms_per_step": ms / steps,
        "wall_ms_per_step": 1e3 * wall / steps,
        "link_updates_per_s": links * steps / (ms * 1e-3),
        "kernel_ms": kms,
        "kernel_ms_source": ("HIP events around each launch in the timed region, reprojection steps left out"
                             if per_launch else "HIP events around the timed region / steps"),
        "kernel_ms_single_launches": {"count": len(single), "mean": sum(single) / len(single),
                                      "min": min(single), "max": max(single)},
        "kernel_gbs": links * BYTES_PER_LINK[precision] / (kms * 1e-3) / 1e9,
        "kernel_ms_back_to_back": kms_b2b,
        "gauss_projection": {"cg_iterations": cg_iters, "before": gauss_raw, "after": gauss_proj},
        "gauss_residual_start": float(m0[6]),
        "gauss_residual_end": float(m1[6]),
        "energy_start": energy0,
        "energy_end": energy1,
        "observables_end": [float(x) for x in m1[:7]],
        "kernel_config": kcfg,
        "autotune_ms": tune_ms,
    }
    g.close()
    return res


def time_cpic(steps, stream, autotune=True):
    """SU(2) CPIC collision, 512 x 64 x 64, FP32 links/E/J/charges with FP64
    reunitarization every 10 steps; two colour-neutral nuclei (Gaussian
    longitudinal profile of width 4 cells, transverse rho_a ~ Gaussian(0, 1)),
    64 particles per transverse cell each (16 cells x ppc = 4), moving at v = c
    in +x0 and -x0; U = 1 and E projected onto Gauss's law for the initial
    charges; active region x0 in [1, N0-2] and evaluation region x0 in [2, N0-3]
    (input/CGC/MV model/MVModel_new_optimized.yaml). A step is Simulation.step:
    storeFields, the fused update with the current, the charge density, and the
    current deposit with parallel transport."""
    import torch
    from openpixi_amd import FieldMeasurements, Grid, _lib, workloads

    L = _lib.load()
    dims, ppc = (512, 64, 64), 4
    g = Grid(2, dims, 1.0, 1.0 / ppc, 1.0, "fp32")
    g._check(L.opx_set_stream(g.handle, ctypes_void(stream.cuda_stream)))
    cols, q = workloads.cpic_collision(dims, ppc)
    g.setSuperParticles(cols["orientation"], cols["number"], cols["index_offset"], cols["per_plane"],
                        cols["sub_lattice_shift"], cols["particles_per_cell"], q)
    # active x0 in [1, N0-2], evaluated x0 in [2, N0-3], as the reference's CGC run settings
    g.setActiveRegion([1, 0, 0], [dims[0] - 2, dims[1] - 1, dims[2] - 1])
    g.setEvaluationRegion([2, 0, 0], [dims[0] - 3, dims[1] - 1, dims[2] - 1])
    g.resetCharge()
    g.interpolateChargedensity(0)
    cg_iters, gb, ga = g.projectGauss(tol=1e-8, max_iter=5000)
    g.simulationInitialize()
    e0, e1 = ready_events(stream, 2)
    tune_ms = g.autotune(reps=3) if autotune else None
    m0 = FieldMeasurements().measure(g)
    torch.cuda.synchronize()
    done = (steps // REUNIT_EVERY) * REUNIT_EVERY
    # every 10th Simulation step reprojects the new links inside its update (OPX_OPT_REUNIT_EVERY)
    g.setReunitarizeEvery(REUNIT_EVERY)
    e0.record(stream)
    g.simulationStep(done)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    m1 = FieldMeasurements().measure(g)
    non_finite, non_unitary = g.checkHealth(1e-4)
    n = int(np.prod(dims))
    res = {
        "workload": "SU(2) CPIC collision 512x64x64, FP32 + FP64 reunitarization every 10 steps, two nuclei of "
                    "16 x 4 x 4096 NGP particles, v = c, dt/a = 0.25",
        "steps": done,
        "ms_per_step": ms / done,
        "link_updates_per_s": 3 * n * done / (ms * 1e-3),
        "particles": int(len(q)),
        "gauss_projection": {"cg_iterations": cg_iters, "before": gb, "after": ga},
        "gauss_residual_start": float(m0[6]),
        "gauss_residual_end": float(m1[6]),
        "energy_start": float(m0[:6].sum()),
        "energy_end": float(m1[:6].sum()),
        "charge_norm_start": float(m0[7]),
        "charge_norm_end": float(m1[7]),
        "health": {"non_finite": non_finite, "non_unitary_1e-4": non_unitary},
        "kernel_config": g.kernelConfig(),
        "autotune_ms": tune_ms,
    }
    g.close()
    return res


BATCH_SIDES = (32, 48, 64, 80, 96)


def batch_event_sides(count, seed=0):
    from openpixi_amd import inputs

    return [int(s) for s in inputs.rng(seed, inputs.STREAM_SIZES).choice(BATCH_SIDES, size=count)]


def make_batch(precision, sides, stream, autotune):
    from openpixi_amd import _lib, inputs
    from openpixi_amd.batch import Batch

    b = Batch(2, [(s, s, s) for s in sides], 1.0, DT, 1.0, precision)
    b.set_stream(stream.cuda_stream)
    for e, s in enumerate(sides):
        b.event(e).setLinksFromCoefficients(inputs.random_links_coefficients(s ** 3, 2, 0.1, seed=1000 + e))
    tune_ms = b.autotune(reps=3) if autotune else None
    # E = 0; the initial half step U(-dt/2) -> U(+dt/2) is the identity on the links then
    b.step_links(DT)
    return b, tune_ms


def time_batch(count, steps, warmup, stream, autotune=True):
    """Batched events: `count` independent SU(2) lattices with cubic sides drawn
    uniformly (seeded Philox) from {32, 48, 64, 80, 96}; each starts as the SU(2)
    32^3 parity case: links exp(i A_a T_a) with A_a ~ Gaussian(0, 0.1) per link,
    E = 0, dt/a = 0.25. The FP32 batch reprojects its links onto SU(2) in double
    precision every 10 steps; every step is one persistent kernel over all
    events, replayed from a captured hipGraph. The FP64 batch runs the same steps
    as the per-event reference the FP32 observables are compared with."""
    import torch

    sides = batch_event_sides(count)
    cells = sum(s ** 3 for s in sides)
    res = {
        "workload": f"{count} SU(2) events, cubic sides from {{32,48,64,80,96}}, FP32 + FP64 reunitarization every "
                    f"10 steps, {steps} steps, one persistent kernel per step in a captured hipGraph",
        "events": count,
        "sides": sides,
        "cells": cells,
        "steps": steps,
    }
    obs = {}
    for precision in ("fp32", "fp64"):
        b, tune_ms = make_batch(precision, sides, stream, autotune)
        reunit = REUNIT_EVERY if precision == "fp32" else 0
        m0 = b.measure()
        e0, e1 = ready_events(stream, 2)
        b.leapfrog(warmup, reunit, graph=True)
        torch.cuda.synchronize()
        e0.record(stream)
        b.leapfrog(steps, reunit, graph=True)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        m1 = b.measure()
        info = b.info()
        obs[precision] = (m0, m1)
        res[precision] = {
            "ms_per_step": ms / steps,
            "events_per_s": count / (ms * 1e-3),
            "link_updates_per_s": 3 * cells * steps / (ms * 1e-3),
            "grid_workgroups": info["grid"],
            "form": b.form(),
            "autotune_ms": tune_ms,
            "max_gauss_residual_end": float(m1[:, 6].max()),
        }
        b.close()
        print(f"# batch {precision}: {json.dumps(res[precision])}", file=sys.stderr, flush=True)
    (a0, a1), (d0, d1) = obs["fp32"], obs["fp64"]
    ea, ed = a1[:, :6].sum(axis=1), d1[:, :6].sum(axis=1)
    per_obs = np.abs(a1[:, :6] - d1[:, :6]) / np.abs(d1[:, :6])
    res["value"] = res["fp32"]["events_per_s"]
    res["unit"] = "events/s"
    res["fp32_vs_fp64_per_event"] = {
        "energy_rel_max": float((np.abs(ea - ed) / np.abs(ed)).max()),
        "energy_rel_median": float(np.median(np.abs(ea - ed) / np.abs(ed))),
        "observable_rel_max": float(per_obs.max()),
        "gauss_abs_max": float(np.abs(a1[:, 6] - d1[:, 6]).max()),
    }
    # E^2 + B^2 read at staggered time levels (E at t, links at t + dt/2) moves by O(dt^2 omega^2)
    # for these rough (UV-dominated) initial links; this is the FP64 run's own change, not FP32 error
    res["fp64_energy_change_rel_max"] = float((np.abs(ed - d0[:, :6].sum(axis=1)) /
                                               np.abs(d0[:, :6].sum(axis=1))).max())
    return res


def time_adjoint(steps, directions, stream):
    """Adjoint: SU(2) 64^3, links exp(i A_a T_a) with A_a ~ Gaussian(0, 0.1) and
    E ~ Gaussian(0, 0.05) per algebra component (seeded Philox), dt/a = 0.25,
    200 steps of the FP32 form with FP64 reunitarization every 10 steps. The loss
    is the lattice average of e_T - e_L (longitudinal direction x0) of the final
    state; its gradient with respect to the initial E comes from the hand-written
    reverse sweep with checkpoints every 20 steps. Checked against central
    finite differences of the FP64 forward map along `directions` random
    directions, and against the FP64 adjoint."""
    import torch
    from openpixi_amd import Grid, _lib, inputs
    from openpixi_amd.adjoint import loss_and_gradient

    dims, K = (64, 64, 64), 20
    n = int(np.prod(dims))
    A = inputs.random_links_coefficients(n, 2, 0.1, seed=7)
    Ec = inputs.random_efield_coefficients(n, 2, 0.05, seed=7)
    g32 = Grid(2, dims, 1.0, DT, 1.0, "fp32")
    g64 = Grid(2, dims, 1.0, DT, 1.0, "fp64")
    for g in (g32, g64):
        g._check(_lib.load().opx_set_stream(g.handle, ctypes_void(stream.cuda_stream)))
    # reference-layout initial state from the coefficients (links through the library's getLink)
    g64.setLinksFromCoefficients(A)
    g64.setEFromCoefficients(Ec)
    U0, E0 = g64.getU().copy(), g64.getE().copy()

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    loss_and_gradient(g32, U0, E0, steps, REUNIT_EVERY, K, 0)   # warm-up (allocates the checkpoints)
    (l32, gr32), t_grad = timed(lambda: loss_and_gradient(g32, U0, E0, steps, REUNIT_EVERY, K, 0))
    (_, _), t_fwd = timed(lambda: loss_and_gradient(g32, U0, E0, steps, REUNIT_EVERY, K, 0, gradient=False))
    (l64, gr64), t_grad64 = timed(lambda: loss_and_gradient(g64, U0, E0, steps, 0, K, 0))
    rng = np.random.default_rng(2024)
    h = 1e-5
    fd = np.empty(directions)
    dot32 = np.empty(directions)
    dot64 = np.empty(directions)
    t0 = time.perf_counter()
    for k in range(directions):
        v = rng.normal(size=E0.shape)
        lp, _ = loss_and_gradient(g64, U0, E0 + h * v, steps, 0, K, 0, gradient=False)
        lm, _ = loss_and_gradient(g64, U0, E0 - h * v, steps, 0, K, 0, gradient=False)
        fd[k] = (lp - lm) / (2 * h)
        dot32[k] = np.sum(gr32 * v)
        dot64[k] = np.sum(gr64 * v)
    t_fd = time.perf_counter() - t0
    rms = float(np.sqrt(np.mean(fd ** 2)))
    res = {
        "workload": "SU(2) 64^3 adjoint, 200 steps FP32 + FP64 reunitarization every 10, checkpoints every 20, "
                    "loss = lattice average of e_T - e_L (x0 longitudinal), gradient w.r.t. initial E",
        "steps": steps,
        "checkpoint_every": K,
        "loss_fp32": l32,
        "loss_fp64": l64,
        "ms_forward": 1e3 * t_fwd,
        "ms_gradient": 1e3 * t_grad,
        "gradient_over_forward": t_grad / t_fwd,
        "ms_gradient_fp64": 1e3 * t_grad64,
        "fd_directions": directions,
        "fd_step": h,
        "fd_seconds": t_fd,
        "fp32_vs_fd_rel_l2": float(np.linalg.norm(dot32 - fd) / np.linalg.norm(fd)),
        "fp32_vs_fd_max_over_rms": float(np.max(np.abs(dot32 - fd)) / rms),
        "fp64_vs_fd_rel_l2": float(np.linalg.norm(dot64 - fd) / np.linalg.norm(fd)),
        "fp32_vs_fp64_adjoint_rel_l2": float(np.linalg.norm(gr32 - gr64) / np.linalg.norm(gr64)),
    }
    g32.close()
    g64.close()
    return res


def time_implicit(steps, iterations, damping, stream):
    """ImplicitTYMSolver on the SU(3) 128^3 lattice (the solver of the implicit CGC simulation
    types): FP32 + FP64 reunitarization every 10 steps, and FP64; link-updates/s of the
    implicit step and the energy change of FP32 against FP64."""
    import torch
    from openpixi_amd import FieldMeasurements, _lib

    L = _lib.load()
    out = {"workload": f"SU(3) {DIMS[0]}x{DIMS[1]}x{DIMS[2]} ImplicitTYMSolver, {iterations} iterations, damping "
                       f"{damping}, beam direction x0, dt/a = {DT}", "steps": steps, "iterations": iterations,
           "damping": damping}
    n = int(np.prod(DIMS))
    energies = {}
    for precision in ("fp32", "fp64"):
        g = make_grid(precision, DIMS)
        g._check(L.opx_set_stream(g.handle, ctypes_void(stream.cuda_stream)))
        g.projectGauss(tol=1e-10, max_iter=4000)
        reunit = REUNIT_EVERY if precision == "fp32" else 0
        g._check(L.opx_step_links(g.handle, DT))
        g._check(L.opx_set_field_solver(g.handle, _lib.SOLVER_IMPLICIT, iterations, damping))
        ev0, ev1 = ready_events(stream, 2)
        g._check(L.opx_leapfrog(g.handle, 2, reunit))
        torch.cuda.synchronize()
        ev0.record(stream)
        g._check(L.opx_leapfrog(g.handle, steps, reunit))
        ev1.record(stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / steps
        m = FieldMeasurements().measure(g)
        energies[precision] = float(m[0:6].sum())
        out[precision] = {"ms_per_step": ms, "link_updates_per_s": 3 * n / (ms * 1e-3),
                          "gauss_residual_end": float(m[6])}
        g.close()
    out["fp32_energy_vs_fp64"] = abs(energies["fp32"] - energies["fp64"]) / abs(energies["fp64"])
    return out


# Two MV nuclei of input/CGC/MV model/MVModel_new_optimized.yaml: (orientation, location as a
# fraction of the lattice, seed); width 2, mu 0.2, UV cutoff 2, coherence 4, IR 0.2, g = 2, dt/a = 0.5
CGC_NUCLEI = ((1, 0.25, 5), (-1, 0.75, 6))
CGC_PARAMS = dict(width=2.0, mu=0.2, uvt=2.0, coherence=4.0, ir=0.2)


def cgc_run(dims, precision, threads, stream=None, oracle=False):
    """The reference's CGC initialization (two nuclei, each MVModel -> LightConePoissonSolverImprovedFull
    -> LightConeNGPSuperParticleCreator) on `dims`, coupling 2 and dt/a = 0.5 as in the YAML file.
    Returns per-stage seconds (wall clock; every device call synchronizes) and counts."""
    from openpixi_amd.cgc import JavaRandom

    n = int(np.prod(dims))
    p = CGC_PARAMS
    st = {"mv_s": 0.0, "solve_s": 0.0, "particles_s": 0.0}
    if oracle:
        import oracle as O

        og = O.OracleGrid(2, dims, (1.0, 1.0, 1.0), 0.5, 2.0)
        charges = 0
        for orient, frac, seed in CGC_NUCLEI:
            t = time.perf_counter()
            noise = JavaRandom(seed).gaussians(3 * n)   # numpy, vectorized
            rho = og.mv_charge(frac * dims[0], p["width"], p["mu"], p["uvt"], p["coherence"], p["ir"], noise)
            t2 = time.perf_counter()
            gauss = og.lightcone_solve(rho, orient)
            t3 = time.perf_counter()
            arrs, _, _ = og.lightcone_superparticles(gauss, orient, threads)
            t4 = time.perf_counter()
            charges += int(arrs["num"].sum())
            for k, a, b in (("mv_s", t, t2), ("solve_s", t2, t3), ("particles_s", t3, t4)):
                st[k] += b - a
        st["charges"] = charges
        return st
    from openpixi_amd import Grid, _lib
    from openpixi_amd.cgc import LightConeNGPSuperParticleCreator, LightConePoissonSolverImprovedFull, MVModel

    g = Grid(2, dims, 1.0, 0.5, 2.0, precision)
    if stream is not None:
        g._check(_lib.load().opx_set_stream(g.handle, ctypes_void(stream.cuda_stream)))

    class _Sim:
        grid = g

        class settings:
            number_of_threads = threads

    blocks = []
    for orient, frac, seed in CGC_NUCLEI:
        t = time.perf_counter()
        m = MVModel(0, orient, frac * dims[0], p["width"], p["mu"], True, seed, p["uvt"], p["coherence"], p["ir"])
        m.initialize(g)   # java.util.Random(seed) noise drawn on the device
        t2 = time.perf_counter()
        solver = LightConePoissonSolverImprovedFull()
        solver.initialize(g)
        solver.solve(m)
        t3 = time.perf_counter()
        c = LightConeNGPSuperParticleCreator()
        c.initialize(_Sim, 0, orient)
        t4 = time.perf_counter()
        blocks.append((c.zStart, c.blockWidth))
        for k, a, b in (("mv_s", t, t2), ("solve_s", t2, t3), ("particles_s", t3, t4)):
            st[k] += b - a
    import ctypes

    cnt, tot = ctypes.c_int(), ctypes.c_int64()
    g._check(_lib.load().opx_superparticles_count(g.handle, ctypes.byref(cnt), ctypes.byref(tot)))
    st.update({"super_particles": cnt.value, "charges": tot.value, "blocks": blocks})
    # the charge density the particles deposit against the fields' Gauss law (block edges aside)
    g.simulationInitialize()
    from openpixi_amd import FieldMeasurements

    m0 = FieldMeasurements().measure(g)
    st["gauss_residual"] = float(m0[6])
    st["charge_norm"] = float(m0[7])
    g.close()
    return st


def time_cgc(stream, cpu=True):
    """CGC initial condition of the reference's input/CGC/MV model/MVModel_new_optimized.yaml
    (SU(2), 128 x 32 x 32, two MV nuclei, 8 subdivisions) and of the CPIC collision lattice
    512 x 64 x 64, FP32 fields: wall-clock seconds per stage on the device (the MV stage
    includes the java.util.Random noise, drawn on the device), and the FP64 oracle on the YAML
    lattice beside it (noise drawn by numpy)."""
    import torch

    cgc_run((16, 8, 8), "fp32", 8, stream)   # loads the kernels
    torch.cuda.synchronize()
    out = {"workload": "CGC initial condition (MVModel x 2, LightConePoissonSolverImprovedFull, "
                       "LightConeNGPSuperParticleCreator), SU(2), FP32 fields, coupling 2, dt/a = 0.5"}
    for name, dims in (("yaml_128x32x32", (128, 32, 32)), ("cpic_512x64x64", (512, 64, 64))):
        r = cgc_run(dims, "fp32", 8, stream)
        r["device_s"] = r["mv_s"] + r["solve_s"] + r["particles_s"]
        out[name] = r
        print(f"# cgc {name}: {json.dumps(r)}", file=sys.stderr, flush=True)
    if cpu:
        r = cgc_run((128, 32, 32), "fp64", 8, oracle=True)
        r["device_s"] = r["mv_s"] + r["solve_s"] + r["particles_s"]
        r["cores"] = 1
        r["kind"] = "port"
        out["cpu_oracle_yaml_128x32x32"] = r
        out["speedup_yaml_128x32x32"] = r["device_s"] / out["yaml_128x32x32"]["device_s"]
    return out


def ctypes_void(p):
    import ctypes

    return ctypes.c_void_p(int(p))


def cpu_threads():
    """Host threads for the CPU baseline: the cores this process may run on, capped by
    OMP_NUM_THREADS when it is set (the host's share for one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(budget_s, dims=DIMS):
    """The FP64 oracle (C restatement of FastTYMSolver, -O3 -ffp-contract=off, OpenMP over
    cells) on the measured workload's lattice (SU(3) 128^3, the same seeded links and E before
    the Gauss projection, which leaves the per-step work unchanged): as many steps as fit the
    budget on every host thread this process may use, then the same on one thread."""
    import oracle as O
    from openpixi_amd import inputs

    n = int(np.prod(dims))
    cores = cpu_threads()
    og = O.OracleGrid(NC, dims, (1.0, 1.0, 1.0), DT, 1.0, threads=cores)
    og.U[:] = O.links_from_coeffs(NC, inputs.random_links_coefficients(n, NC, 0.1, 0))
    og.E[:] = O.algs_from_coeffs(NC, inputs.random_efield_coefficients(n, NC, 0.05, 0))
    t1 = og.time_steps(1, 1)
    steps = max(1, int(budget_s / max(t1, 1e-6)))
    t = og.time_steps(steps, 0)
    # one thread: a third of the budget, at least one step
    og.set_threads(1)
    s1 = og.time_steps(1, 0)
    steps1 = max(0, int(budget_s / 3 / max(s1, 1e-6)) - 1)
    t_one = s1 + (og.time_steps(steps1, 0) if steps1 > 0 else 0.0)
    steps1 += 1
    lat = "x".join(str(d) for d in dims)
    return {
        "value": 3 * n * steps / t,
        "unit": "link-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"FP64 oracle (C, gcc -O3 -ffp-contract=off, OpenMP {cores} threads), SU(3) {lat}, "
                  f"{steps} steps after 1 warm-up step, {t:.1f} s",
        "single_core": {"value": 3 * n * steps1 / t_one, "unit": "link-updates/s", "cores": 1,
                        "sample": f"the same oracle and lattice on one thread, {steps1} steps, {t_one:.1f} s"},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this workload has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
    import torch

    dims = tuple(args.dims)
    forms = args.forms.split(",")
    if args.probe:
        probe(dims, forms, parse_kernels(args.probe_kernels), nc=args.probe_colors)
        return
    stream = torch.cuda.Stream()
    # no cyclic garbage collection during the run: a collection pause between a timed region's
    # opening synchronization and its first launch would leave the GPU idle there (see
    # ready_events); the bench's large arrays are freed by reference counting
    gc.disable()
    results = {}
    for f in forms:
        results[f] = time_form(f, dims, args.steps, args.warmup, args.kernel_reps, stream, not args.no_autotune)
        print(f"# {f}: {json.dumps(results[f])}", file=sys.stderr, flush=True)
    cpic = None
    if not args.no_cpic:
        cpic = time_cpic(args.cpic_steps, stream, not args.no_autotune)
        print(f"# cpic: {json.dumps(cpic)}", file=sys.stderr, flush=True)
    batch = None
    if not args.no_batch:
        batch = time_batch(args.batch_events, args.batch_steps, 20, stream, not args.no_autotune)
        print(f"# batch: {json.dumps(batch)}", file=sys.stderr, flush=True)
    implicit = None
    if not args.no_implicit:
        # the settings of input/CGC/MV model/MVModel_implicit_solver.yaml: 3 iterations, damping 0
        implicit = time_implicit(args.implicit_steps, 3, 0.0, stream)
        print(f"# implicit: {json.dumps(implicit)}", file=sys.stderr, flush=True)
    cgc = None
    if not args.no_cgc:
        cgc = time_cgc(stream, cpu=not args.no_cpu_baseline)
        print(f"# cgc: {json.dumps(cgc)}", file=sys.stderr, flush=True)
    adjoint = None
    if not args.no_adjoint:
        adjoint = time_adjoint(args.adjoint_steps, args.fd_directions, stream)
        print(f"# adjoint: {json.dumps(adjoint)}", file=sys.stderr, flush=True)
    main_form = "fp32" if "fp32" in results else forms[0]
    r = results[main_form]
    out = {
        "metric": "SU(3) link-updates/s at 128^3: FP32+FP64-reunit vs FP64 at equal Gauss residual",
        "value": r["link_updates_per_s"],
        "unit": "link-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if main_form == "fp32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"SU(3) pure Yang-Mills leapfrog, periodic {dims[0]}x{dims[1]}x{dims[2]}, dt/a=0.25, "
                        f"{'FP32 links/E + FP64 reunitarization every 10 steps' if main_form == 'fp32' else 'FP64'}",
            "lattice": list(dims),
            "colors": NC,
            "reunitarize_every": REUNIT_EVERY if main_form == "fp32" else 0,
        },
        "forms": results,
    }
    if "fp32" in results and "fp64" in results:
        out["fp32_over_fp64"] = results["fp32"]["link_updates_per_s"] / results["fp64"]["link_updates_per_s"]
        names = ["E2_0", "E2_1", "E2_2", "B2_0", "B2_1", "B2_2", "gauss"]
        a, b = results["fp32"]["observables_end"], results["fp64"]["observables_end"]
        out["fp32_drift_vs_fp64"] = {nm: (abs(x - y) / abs(y) if nm != "gauss" else abs(x - y))
                                     for nm, x, y in zip(names, a, b)}
        out["fp32_drift_vs_fp64"]["energy"] = abs(results["fp32"]["energy_end"] - results["fp64"]["energy_end"]) / \
            abs(results["fp64"]["energy_end"])
    achieved = r["kernel_gbs"]
    out["roofline"] = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": profiled_traffic(main_form, r["kernel_config"], dims)[0],
        "traffic_source": profiled_traffic(main_form, r["kernel_config"], dims)[1],
        "kernel": r["kernel_config"],
        "kernel_ms": r["kernel_ms"],
        "valu_issue": valu_issue(main_form, r["kernel_config"], dims, r["kernel_ms"]),
        "bytes_per_link": BYTES_PER_LINK[main_form],
    }
    if cpic is not None:
        out["cpic"] = cpic
    if batch is not None:
        out["batch"] = batch
    if adjoint is not None:
        out["adjoint"] = adjoint
    if implicit is not None:
        out["implicit"] = implicit
    if cgc is not None:
        out["cgc_init"] = cgc
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds, dims)
    out["host"] = platform.node()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
