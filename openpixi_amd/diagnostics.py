"""Diagnostics of the FastTYMSolver path: the reference's BulkQuantitiesInTime and
ProjectedEnergyDensity output files, computed by the device reductions and written in
the reference's file formats, and the UnitarityTester and GaussConstraintRestoration
diagnostics run on the device with the reference's console messages.

Number formats follow the Java library calls the reference uses:

* ``java_sci(x)`` is ``new DecimalFormat("0.################E0").format(x)``
  (diagnostics/FileFunctions.java:8,65-75; BulkQuantitiesInTime.java:107): one integer
  digit, up to 16 fraction digits without trailing zeros, exponent with a minus sign only
  when negative (``1E0``, ``-2.5E-3``, ``0E0``). The digits are the shortest decimal that
  reads back as the same double; 17 significant digits never need rounding.
* ``java_double_to_string(x)`` is ``Double.toString(x)`` (the time line of
  ProjectedEnergyDensity.java:87-88): plain decimal with at least one fraction digit for
  1e-3 <= |x| < 1e7 (``0.25``, ``100.0``), else ``d.dddE<exp>`` (``1.0E-4``, ``1.0E7``).

Both agree with the JDK for every finite double whose shortest decimal has at least two
digits or is exact, which covers every value these files hold. They differ only for
subnormal numbers, where the JDK may print a second digit closer to the binary value.
The decimal separator is '.', as in the JDK's root locale.
"""
from __future__ import annotations

import math
import os
import sys
from decimal import Decimal
from typing import List

import numpy as np

from .grid import FieldMeasurements, Grid


def _shortest_digits(x: float):
    """(negative, digit string without trailing zeros, decimal exponent E) with
    |x| = d1.d2d3... x 10^E, from the shortest round-trip decimal of x."""
    sign, digits, exp = Decimal(repr(float(x))).as_tuple()
    ds = "".join(str(d) for d in digits).lstrip("0")   # |x| = int(ds) * 10^exp
    return bool(sign), ds.rstrip("0"), exp + len(ds) - 1


def java_sci(x: float) -> str:
    """DecimalFormat("0.################E0").format(x)."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "-∞" if x < 0 else "∞"
    if x == 0.0:
        return "-0E0" if math.copysign(1.0, x) < 0 else "0E0"
    neg, ds, e = _shortest_digits(x)
    mant = ds[0] + ("." + ds[1:] if len(ds) > 1 else "")
    return ("-" if neg else "") + mant + "E" + str(e)


def java_double_to_string(x: float) -> str:
    """Double.toString(x)."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "-Infinity" if x < 0 else "Infinity"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    neg, ds, e = _shortest_digits(x)
    sign = "-" if neg else ""
    if 1e-3 <= abs(x) < 1e7:
        if e >= 0:
            ip = (ds + "0" * max(0, e + 1 - len(ds)))[: e + 1]
            fp = ds[e + 1:] or "0"
        else:
            ip = "0"
            fp = "0" * (-e - 1) + ds
        return sign + ip + "." + fp
    return sign + ds[0] + "." + (ds[1:] or "0") + "E" + str(e)


# ---------------------------------------------------------------- FileFunctions
def output_path(path: str) -> str:
    """FileFunctions.getFile (diagnostics/FileFunctions.java:17-36): a relative path is
    placed under 'output/'; missing folders are created."""
    new_path = path if path.startswith("/") else os.path.join("output", path)
    folder = os.path.dirname(new_path)
    if folder and not os.path.exists(folder):
        os.makedirs(folder, exist_ok=True)
    return new_path


def clear_file(path: str) -> None:
    """FileFunctions.clearFile (diagnostics/FileFunctions.java:43-56)."""
    new_path = path if path.startswith("/") else os.path.join("output", path)
    if os.path.exists(new_path):
        os.remove(new_path)


def tsv_string(values) -> str:
    """FileFunctions.generateTSVString (diagnostics/FileFunctions.java:63-75)."""
    return "\t".join(java_sci(v) for v in np.asarray(values, float).reshape(-1))


def step_interval(time_interval: float, time_step: float) -> int:
    """(int) Math.max(Math.round(timeInterval / timeStep), 1); Math.round rounds half up."""
    return int(max(math.floor(time_interval / time_step + 0.5), 1))


# ---------------------------------------------------------------- diagnostics
class BulkQuantitiesInTime:
    """diagnostics/methods/BulkQuantitiesInTime.java:17-121: every stepInterval steps, the
    lattice-averaged E^2, B^2, the diagonal pressures and the Gauss and charge measures."""

    HEADER = "#time \t E^2 \t B^2 \t P_x \t P_y \t P_z \t G \t rho^2"

    def __init__(self, path: str, time_interval: float, supress_output: bool = False):
        self.path = path
        self.time_interval = time_interval
        self.supress_output = supress_output
        self.eSquared = self.bSquared = self.px = self.py = self.pz = 0.0
        self.gaussViolation = self.totalChargeSquared = 0.0

    def initialize(self, sim) -> None:
        self.sim = sim
        self.stepInterval = step_interval(self.time_interval, sim.tstep)
        if not self.supress_output:
            clear_file(self.path)
            with open(output_path(self.path), "a") as f:
                f.write(self.HEADER + "\n")

    def calculate(self, grid: Grid, steps: int) -> None:
        if steps % self.stepInterval != 0:
            return
        m = FieldMeasurements().measure(grid)
        es, bs = m[0:3], m[3:6]
        self.eSquared = es[0] + es[1] + es[2]
        self.bSquared = bs[0] + bs[1] + bs[2]
        self.px = -es[0] + es[1] + es[2] - bs[0] + bs[1] + bs[2]
        self.py = +es[0] - es[1] + es[2] + bs[0] - bs[1] + bs[2]
        self.pz = +es[0] + es[1] - es[2] + bs[0] + bs[1] - bs[2]
        self.gaussViolation = float(m[6])
        self.totalChargeSquared = float(m[7])
        if not self.supress_output:
            row = [steps * self.sim.tstep, self.eSquared, self.bSquared, self.px, self.py, self.pz,
                   self.gaussViolation, self.totalChargeSquared]
            with open(output_path(self.path), "a") as f:
                f.write("\t".join(java_sci(v) for v in row) + "\n")


class ProjectedEnergyDensity:
    """diagnostics/methods/ProjectedEnergyDensity.java:43-110: every stepInterval steps, the
    time, the four energy-density components per slice along `direction` (transverse
    electric, transverse magnetic, longitudinal electric, longitudinal magnetic) and the
    longitudinal Poynting vector per slice (averaged, time averaged), one line each."""

    def __init__(self, path: str, time_interval: float, direction: int, computeEnergyDensity: bool = True,
                 computePoyntingVector: bool = True):
        self.path = path
        self.time_interval = time_interval
        self.direction = direction
        self.computeEnergyDensity = computeEnergyDensity
        self.computePoyntingVector = computePoyntingVector

    def initialize(self, sim) -> None:
        self.sim = sim
        self.stepInterval = step_interval(self.time_interval, sim.tstep)
        clear_file(self.path)

    def calculate(self, grid: Grid, steps: int) -> None:
        if steps % self.stepInterval != 0:
            return
        fm = FieldMeasurements()
        lines: List[str] = [java_double_to_string(steps * grid.time_step)]
        if self.computeEnergyDensity:
            lines += [tsv_string(a) for a in fm.projectedEnergyDensity(grid, self.direction)]
        if self.computePoyntingVector:
            lines += [tsv_string(a) for a in fm.projectedPoynting(grid, self.direction)]
        with open(output_path(self.path), "a") as f:
            f.write("\n".join(lines) + "\n")


class UnitarityTester:
    """diagnostics/methods/UnitarityTester.java:13-46: every stepInterval steps, one line
    "UnitarityTester: non-unitary link found @ i" per link (cell i, any direction) whose U or
    Unext has |Re tr| not below N + 1e-8."""

    TOLERANCE = 1e-8

    def __init__(self, time_interval: float, out=None):
        self.time_interval = time_interval
        self.out = out if out is not None else sys.stdout

    def initialize(self, sim) -> None:
        self.stepInterval = step_interval(self.time_interval, sim.tstep)

    def calculate(self, grid: Grid, steps: int) -> None:
        if steps % self.stepInterval != 0:
            return
        count, cells = grid.unitarityTest(self.TOLERANCE)
        for i in cells:
            print(f"UnitarityTester: non-unitary link found @ {int(i)}", file=self.out)


class GaussConstraintRestoration:
    """diagnostics/methods/GaussConstraintRestoration.java:17-112: every stepInterval steps from
    stepOffset on (or once at stepOffset with applyOnlyOnce), the restoration iteration on the
    device, with the reference's messages."""

    def __init__(self, time_interval: float, time_offset: float, gamma: float, max_iterations: int,
                 absolute_value: float, apply_only_once: bool, out=None):
        self.time_interval = time_interval
        self.time_offset = time_offset
        self.gamma = gamma
        self.max_iterations = max_iterations
        self.absolute_value = absolute_value
        self.apply_only_once = apply_only_once
        self.already_applied = False
        self.out = out if out is not None else sys.stdout
        self.last = None   # (kept, reason, before, after) of the last run

    def initialize(self, sim) -> None:
        self.stepInterval = step_interval(self.time_interval, sim.tstep)
        self.stepOffset = int(self.time_offset / sim.tstep)   # (int) truncates towards zero

    def calculate(self, grid: Grid, steps: int) -> None:
        if not self.apply_only_once:
            if (steps - self.stepOffset) % self.stepInterval == 0:
                self.iterate(grid)
        elif not self.already_applied and steps == self.stepOffset:
            self.iterate(grid)
            self.already_applied = True

    def iterate(self, grid: Grid) -> None:
        kept, reason, before, after = grid.restoreGaussConstraint(self.gamma, self.max_iterations,
                                                                  self.absolute_value)
        self.last = (kept, reason, before, after)
        if reason == 0:
            print(f"GaussConstraintRestoration: Reached absolute value at step #{kept}", file=self.out)
        elif reason == 1:
            print(f"GaussConstraintRestoration: Reached instability at step #{kept}", file=self.out)
        elif reason == 2:
            print(f"GaussConstraintRestoration: Reached maximum number of iterations at step #{kept - 1}",
                  file=self.out)


def from_yaml_output(section) -> list:
    """YamlOutput.applyTo (ui/util/yaml/YamlOutput.java:45-105) for the diagnostics of this path,
    in the reference's order (bulkQuantitiesInTime, unitarityTester, gaussConstraintRestoration,
    projectedEnergyDensity), with the YAML defaults of YamlUnitarityTester,
    YamlGaussConstraintRestoration and YamlProjectedEnergyDensity (computeEnergyDensity and
    computePoyntingVector true). Other output types are not part of this path."""
    out = []
    section = section or {}
    for o in section.get("bulkQuantitiesInTime") or []:
        out.append(BulkQuantitiesInTime(str(o["path"]), float(o["interval"])))
    # YamlUnitarityTester (interval has no default: an absent one is 0, every step)
    for o in section.get("unitarityTester") or []:
        out.append(UnitarityTester(float(o.get("interval", 0.0))))
    # YamlGaussConstraintRestoration defaults; absoluteValue's is the expression 10-8 = 2
    for o in section.get("gaussConstraintRestoration") or []:
        out.append(GaussConstraintRestoration(float(o.get("interval", 1.0)), float(o.get("offset", 0.0)),
                                              float(o.get("gamma", 0.25)), int(o.get("maxIterations", 100)),
                                              float(o.get("absoluteValue", 10 - 8)),
                                              bool(o.get("applyOnlyOnce", False))))
    for o in section.get("projectedEnergyDensity") or []:
        out.append(ProjectedEnergyDensity(str(o["path"]), float(o["interval"]), int(o.get("direction", 0)),
                                          bool(o.get("computeEnergyDensity", True)),
                                          bool(o.get("computePoyntingVector", True))))
    return out
