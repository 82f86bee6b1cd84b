"""Reverse-mode derivative of the SU(2) leapfrog through the C ABI (opx_adjoint_gradient).

The loss is the lattice average of the transverse-minus-longitudinal energy
density after `steps` leapfrog steps (the ProjectedEnergyDensity components,
diagnostics/methods/ProjectedEnergyDensity.java:132-160); the gradient is with
respect to the initial electric field (3 algebra components per link, the
reference's SU(2) algebra storage), computed by the library's hand-written
adjoint kernels with checkpoints every `checkpoint_every` steps.

`LeapfrogLoss` exposes it as a torch.autograd.Function: forward runs the
forward map and the reverse sweep on the GPU, backward returns the stored
gradient scaled by the incoming cotangent.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .grid import Grid


def loss_and_gradient(grid: Grid, U0, E0, steps, reunit_every=0, checkpoint_every=20, direction=0, gradient=True):
    """Set the initial state (links U0, field E0; reference per-cell layout), run the
    forward map, and return (loss, dL/dE0 or None). The grid holds the final state afterwards."""
    grid.setU(np.ascontiguousarray(U0, dtype=np.float64))
    grid.setE(np.ascontiguousarray(E0, dtype=np.float64))
    loss = ctypes.c_double()
    n_values = grid.total_cells * 3 * 3
    grad = np.zeros((grid.total_cells, 3, 3)) if gradient else None
    rc = _lib.load().opx_adjoint_gradient(grid.handle, int(steps), int(reunit_every), int(checkpoint_every),
                                          int(direction), ctypes.byref(loss),
                                          _lib.dptr(grad) if gradient else None, n_values if gradient else 0)
    grid._check(rc)
    return loss.value, grad


def _torch():
    import torch

    return torch


def _make_function():
    torch = _torch()

    class LeapfrogLoss(torch.autograd.Function):
        """loss = LeapfrogLoss.apply(E0, grid, U0, steps, reunit_every, checkpoint_every, direction)"""

        @staticmethod
        def forward(ctx, E0, grid, U0, steps, reunit_every, checkpoint_every, direction):
            e = E0.detach().to("cpu", torch.float64).numpy().reshape(grid.total_cells, 3, 3)
            loss, grad = loss_and_gradient(grid, U0, e, steps, reunit_every, checkpoint_every, direction,
                                           gradient=E0.requires_grad)
            if grad is not None:
                ctx.save_for_backward(torch.from_numpy(grad.reshape(E0.shape)).to(E0.device, E0.dtype))
            return torch.tensor(loss, dtype=E0.dtype, device=E0.device)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g, None, None, None, None, None, None

    return LeapfrogLoss


_FN = None


def LeapfrogLoss():
    """The torch.autograd.Function class (created on first use so that importing
    this module does not import torch)."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN


def leapfrog_loss(E0, grid: Grid, U0, steps, reunit_every=0, checkpoint_every=20, direction=0):
    """Differentiable loss of the initial field E0 (a torch tensor of num_cells*3*3 values)."""
    return LeapfrogLoss().apply(E0, grid, U0, steps, reunit_every, checkpoint_every, direction)
