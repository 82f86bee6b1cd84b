"""MI355X (gfx950) real-time lattice Yang-Mills leapfrog with colored-particle coupling.

The product is the shared library ``libopx.so`` (C ABI in ``include/opx.h``)
built from ``csrc/`` by ``python -m openpixi_amd.build``. This package mirrors
OpenPixi's Grid / FastTYMSolver / ImplicitTYMSolver / FieldMeasurements / Simulation seams and
the CGC initial condition (``cgc``) over that ABI. Importing the package does not require the library; creating a Grid
does, and fails loudly when it is missing.
"""
from ._lib import LIB_PATH, OpxError, OpxLibraryError, load  # noqa: F401
from .grid import FastTYMSolver, FieldMeasurements, FieldSolver, Grid, ImplicitTYMSolver  # noqa: F401
from .cgc import (CGCInitialCondition, LightConeNGPSuperParticleCreator,  # noqa: F401
                  LightConePoissonSolverImprovedFull, MVModel)
from .simulation import ParticleSet, Settings, Simulation, SuperParticleSet  # noqa: F401

__all__ = [
    "Grid", "FieldSolver", "FastTYMSolver", "ImplicitTYMSolver", "FieldMeasurements", "Settings", "Simulation", "SuperParticleSet",
    "CGCInitialCondition", "MVModel", "LightConePoissonSolverImprovedFull", "LightConeNGPSuperParticleCreator",
    "OpxError", "OpxLibraryError", "load", "LIB_PATH",
]
