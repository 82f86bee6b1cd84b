"""Host side of the output writers: Java number formats, FileFunctions paths, the
diagnostics step interval and the YAML output section (no GPU)."""
import math
import os

import pytest

from openpixi_amd.diagnostics import (BulkQuantitiesInTime, ProjectedEnergyDensity, from_yaml_output,
                                      java_double_to_string, java_sci, output_path, step_interval, tsv_string)
from openpixi_amd.simulation import Settings

# new DecimalFormat("0.################E0").format(x) as the JDK prints it
SCI = [
    (0.0, "0E0"), (-0.0, "-0E0"), (1.0, "1E0"), (-2.5, "-2.5E0"), (12345.0, "1.2345E4"), (0.00012, "1.2E-4"),
    (1.0 / 3.0, "3.333333333333333E-1"), (2.0 / 3.0, "6.666666666666666E-1"), (0.1, "1E-1"),
    (0.1 + 0.2, "3.0000000000000004E-1"), (123.456, "1.23456E2"), (1e300, "1E300"), (1e-10, "1E-10"),
    (6.02214076e23, "6.02214076E23"), (-1.5e-7, "-1.5E-7"), (100.0, "1E2"), (1.7976931348623157e308,
                                                                              "1.7976931348623157E308"),
]

# Double.toString(x) as the JDK prints it
DTS = [
    (0.0, "0.0"), (-0.0, "-0.0"), (0.25, "0.25"), (100.0, "100.0"), (1e7, "1.0E7"), (9999999.0, "9999999.0"),
    (1e-4, "1.0E-4"), (0.001, "0.001"), (1234567.125, "1234567.125"), (12.5, "12.5"), (-0.75, "-0.75"),
    (1.5e-5, "1.5E-5"), (123456789.0, "1.23456789E8"), (0.1 + 0.2, "0.30000000000000004"), (1.0, "1.0"),
    (2.5e-3, "0.0025"), (float("nan"), "NaN"), (float("inf"), "Infinity"), (-float("inf"), "-Infinity"),
]


@pytest.mark.parametrize("x,want", SCI)
def test_decimal_format_scientific(x, want):
    assert java_sci(x) == want


@pytest.mark.parametrize("x,want", DTS)
def test_double_to_string(x, want):
    assert java_double_to_string(x) == want


def test_formats_read_back_exactly():
    import numpy as np

    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.normal(0, 1, 200) * 10.0 ** rng.integers(-30, 30, 200), [5e-324 * 3, 1e-308]])
    for x in xs:
        assert float(java_sci(x)) == x
        assert float(java_double_to_string(x)) == x


def test_tsv_and_step_interval():
    assert tsv_string([1.0, 0.5, -0.0]) == "1E0\t5E-1\t-0E0"
    # (int) Math.max(Math.round(timeInterval / timeStep), 1), Math.round = floor(x + 0.5)
    assert step_interval(0.6, 0.25) == 2
    assert step_interval(0.625, 0.25) == 3
    assert step_interval(0.0, 0.25) == 1
    assert step_interval(1.0, 0.1) == 10


def test_output_paths(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    p = output_path("runs/a/bulk.dat")
    assert p == os.path.join("output", "runs/a/bulk.dat")
    assert os.path.isdir(os.path.join(tmp_path, "output", "runs", "a"))
    absolute = str(tmp_path / "x" / "y.dat")
    assert output_path(absolute) == absolute


def test_yaml_output_section():
    s = Settings.from_yaml("""
simulationType: temporal yang-mills
timeStep: 0.25
duration: 2
numberOfColors: 2
gridCells: [4, 4, 4]
gridStep: 1
output:
  bulkQuantitiesInTime:
    - path: bulk.dat
      interval: 0.5
  projectedEnergyDensity:
    - path: ped.dat
      interval: 1.0
      direction: 2
      computePoyntingVector: false
""")
    kinds = [type(d) for d in s.diagnostics]
    assert kinds == [BulkQuantitiesInTime, ProjectedEnergyDensity]
    assert s.diagnostics[0].path == "bulk.dat" and s.diagnostics[0].time_interval == 0.5
    ped = s.diagnostics[1]
    assert ped.direction == 2 and ped.computeEnergyDensity and not ped.computePoyntingVector
    assert from_yaml_output(None) == []
