"""Synthetic workloads at the sizes the benchmark and tests use.

cpic_collision(): two SU(2) nuclei moving at v = c in +x0 and -x0 on an
N0 x N1 x N2 lattice, carried by CGC super particles in the layout
LightConeNGPSuperParticleCreator builds (physics/initial/CGC/LightConeNGPSuperParticleCreator.java:102-205):
a block of `block` longitudinal cells per nucleus, split into `subdivisions`
contiguous sub-blocks, one super particle per (sub-block, sub-lattice shift k),
ppc = round(a/dt) particles per cell, charge of particle (cell, k, x_perp)
= rho(ngp cell) / ppc with ngp = cell for k < ppc/2 and cell + 1 otherwise.
The colour charge density is a Gaussian longitudinal profile of width `width`
cells times a transverse field rho_a(x_perp) drawn i.i.d. Gaussian(0, 1) per
colour component (seeded Philox), with its transverse mean removed so that each
nucleus is colour neutral (the periodic Gauss law then has a solution). The reference builds rho from an MV-model
initial condition (FFT regulation, Poisson solve); this generator replaces that
step with a seeded closed-form density so that the oracle and the device see
identical inputs.
"""
from __future__ import annotations

import numpy as np

from . import inputs


def cpic_collision(dims=(512, 64, 64), ppc=4, centers=(128, 384), width=4.0, block=16, subdivisions=4, seed=0,
                   num_colors=2):
    n0, n1, n2 = dims
    pp = n1 * n2
    A = num_colors * num_colors - 1
    rng = inputs.rng(seed, inputs.STREAM_CHARGES)
    cols = dict(orientation=[], number=[], index_offset=[], per_plane=[], sub_lattice_shift=[],
                particles_per_cell=[])
    charges = []
    for orientation, c in zip((1, -1), centers):
        transverse = rng.normal(0.0, 1.0, size=(pp, A))          # rho_a(x_perp)
        transverse -= transverse.mean(axis=0)                    # colour-neutral nucleus (no zero mode)
        z_start = int(c - block // 2)
        ws = block // subdivisions
        for j in range(subdivisions):
            first = z_start + j * ws
            for k in range(ppc):
                cols["orientation"].append(orientation)
                cols["number"].append(ws * pp)
                cols["index_offset"].append(first * pp)
                cols["per_plane"].append(pp)
                cols["sub_lattice_shift"].append(k)
                cols["particles_per_cell"].append(ppc)
                z = np.arange(first, first + ws) + (0 if k < ppc // 2 else 1)
                prof = np.exp(-0.5 * ((z - c) / width) ** 2) / ppc
                coeff = (prof[:, None, None] * transverse[None, :, :]).reshape(ws * pp, A)
                charges.append(inputs.coeffs_to_storage(num_colors, coeff))
    cols = {k: np.asarray(v, np.int32) for k, v in cols.items()}
    return cols, np.concatenate(charges)
