"""CPU checks of the boundary and the host layer (no GPU calls):

* libopx.so loads and exports every function include/opx.h declares, and the
  ctypes table binds exactly those functions;
* the product path has no CPU fallback: without the library every entry fails loudly;
* Settings.from_yaml reads the reference's YAML keys (ui/util/yaml/YamlSettings.java)
  and runs ceil(duration/timeStep) iterations (physics/Settings.java:374-379);
* region resolution counts negative coordinates from the end (YamlRegion).
"""
import ctypes
import os
import re

import pytest

from openpixi_amd import _lib
from openpixi_amd.simulation import Settings

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "opx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"\b(opx_\w+)\s*\(", src))


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert len(names) >= 40
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = sorted(n for n in names if not hasattr(lib, n))
    assert not missing, missing


def test_ctypes_table_binds_exactly_the_header():
    assert set(_lib.SIGNATURES) == declared_functions()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.OpxLibraryError):
        _lib.load()


YAML = """
simulationType: temporal optimized cgc ngp
gridStep: 1
couplingConstant: 2
numberOfDimensions: 3
numberOfColors: 2
gridCells: [128, 32, 32]
timeStep: 0.5
duration: 64
evaluationRegion:
  enabled: true
  point1: [2, 0, 0]
  point2: [-3, -1, -1]
activeRegion:
  enabled: true
  point1: [1, 0, 0]
  point2: [-2, -1, -1]
device:
  precision: fp32
  reunitarizeEvery: 10
panels:
  dividerLocation: 785
"""


def test_settings_from_reference_yaml():
    s = Settings.from_yaml(YAML)
    assert s.simulation_type == "TemporalOptimizedCGCNGP"
    assert s.grid_cells == [128, 32, 32]
    assert s.grid_steps == [1.0, 1.0, 1.0]
    assert s.coupling_constant == 2.0 and s.number_of_colors == 2
    assert s.time_step == 0.5 and s.iterations == 128
    assert s.active_region.resolved(s.grid_cells) == ([1, 0, 0], [126, 31, 31])
    assert s.evaluation_region.resolved(s.grid_cells) == ([2, 0, 0], [125, 31, 31])
    assert s.precision == "fp32" and s.reunitarize_every == 10


def test_settings_defaults_and_errors():
    s = Settings.from_yaml("timeStep: 0.3\nduration: 1.0\n")
    assert s.iterations == 4            # ceil(1.0 / 0.3)
    assert s.number_of_colors == 1      # YamlSettings default
    with pytest.raises(RuntimeError):
        Settings.from_yaml("simulationType: nonsense\n")
    with pytest.raises(RuntimeError):
        Settings.from_yaml("numberOfDimensions: 3\ngridCells: [4, 4]\n")
    s = Settings.from_yaml("numberOfDimensions: 3\ngridSteps: [1.0, 0.5, 2.0]\n")
    assert s.grid_steps == [1.0, 0.5, 2.0]


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys

    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], capture_output=True, text=True, timeout=120,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_bench_cpu_baseline_sample(monkeypatch):
    """The CPU leg of bench.py on a short budget (the bench runs it on SU(3) 128^3; a 16^3 lattice
    here): the oracle's step loop, with the thread count capped by OMP_NUM_THREADS, and the sample
    naming the lattice and the compiler flags."""
    import bench

    monkeypatch.setenv("OMP_NUM_THREADS", "2")
    assert bench.cpu_threads() <= 2
    r = bench.cpu_baseline(0.3, dims=(16, 16, 16))
    assert r["kind"] == "port" and r["unit"] == "link-updates/s" and r["cores"] <= 2
    assert r["value"] > 0 and r["single_core"]["value"] > 0 and r["single_core"]["cores"] == 1
    assert "16x16x16" in r["sample"] and "-O3 -ffp-contract=off" in r["sample"]
    assert bench.DIMS == (128, 128, 128)
