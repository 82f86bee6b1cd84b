"""SU(2) adjoint on the device (opx_adjoint_gradient) against the oracle's
adjoint, finite differences of the device forward map, torch's gradcheck, and
its own FP64 form."""
import numpy as np
import pytest

import oracle as O
from openpixi_amd import Grid, OpxError
from openpixi_amd.adjoint import leapfrog_loss, loss_and_gradient

from .helpers import make_state, rel_err

pytestmark = pytest.mark.gpu


def device(dims, spacing=(1.0, 1.0, 1.0), g=1.0, precision="fp64"):
    return Grid(2, dims, list(spacing), 0.25, g, precision)


@pytest.mark.parametrize("dims,spacing,g,steps,K,direction", [
    ((6, 5, 7), (1.0, 1.0, 1.0), 1.0, 9, 4, 0),
    ((4, 6, 5), (1.0, 0.8, 1.3), 1.3, 7, 3, 2),
    ((8, 8, 8), (1.0, 1.0, 1.0), 1.0, 12, 20, 1),
])
def test_adjoint_matches_oracle(gpu, dims, spacing, g, steps, K, direction):
    print(f"adjoint vs oracle dims={dims} spacing={spacing} steps={steps} K={K}", flush=True)
    U, E0 = make_state(2, dims, 0.4, 0.2, seed=11)
    og = O.OracleGrid(2, dims, spacing, 0.25, g)
    og.U[:] = U
    og.E[:] = E0
    lo, go = og.adjoint_gradient(steps, direction)
    dg = device(dims, spacing, g)
    ld, gd = loss_and_gradient(dg, U, E0, steps, 0, K, direction)
    assert abs(ld - lo) < 1e-12 * abs(lo)
    assert rel_err(gd, go) < 1e-11
    # the context holds the final forward state
    assert rel_err(dg.getE(), og.E) < 1e-12
    assert rel_err(dg.getUnext(), og.Unext) < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_checkpoint_interval_does_not_change_gradient(gpu, precision):
    dims = (6, 6, 6)
    U, E0 = make_state(2, dims, 0.3, 0.1, seed=12)
    grads = []
    for K in (1, 3, 7, 50):
        dg = device(dims, precision=precision)
        grads.append(loss_and_gradient(dg, U, E0, 14, 5 if precision == "fp32" else 0, K, 0)[1])
    for gk in grads[1:]:
        assert np.array_equal(gk, grads[0])


def test_gradient_matches_device_finite_differences(gpu):
    dims = (12, 12, 12)
    steps = 30
    U, E0 = make_state(2, dims, 0.1, 0.05, seed=13)
    dg = device(dims)
    _, grad = loss_and_gradient(dg, U, E0, steps, 0, 8, 0)
    rng = np.random.default_rng(0)
    h = 1e-5   # perturbation 2e-4 of |E| ~ 0.05: the h^2 term stays below 1e-8 relative
    for _ in range(8):
        v = rng.normal(size=E0.shape)
        lp, _ = loss_and_gradient(dg, U, E0 + h * v, steps, 0, 8, 0, gradient=False)
        lm, _ = loss_and_gradient(dg, U, E0 - h * v, steps, 0, 8, 0, gradient=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - np.sum(grad * v)) < 1e-7 * abs(fd), (fd, np.sum(grad * v))


def test_torch_gradcheck(gpu):
    import torch

    dims = (3, 3, 4)
    U, E0 = make_state(2, dims, 0.4, 0.2, seed=14)
    dg = device(dims)
    e = torch.tensor(E0.reshape(-1), dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda x: leapfrog_loss(x, dg, U, 4, 0, 2, 0), (e,), eps=1e-6, atol=1e-9,
                                    rtol=1e-6)
    loss = leapfrog_loss(e, dg, U, 4, 0, 2, 0)
    loss.backward()
    og = O.OracleGrid(2, dims, (1.0, 1.0, 1.0), 0.25, 1.0)
    og.U[:] = U
    og.E[:] = E0
    _, go = og.adjoint_gradient(4, 0)
    assert rel_err(e.grad.numpy().reshape(go.shape), go) < 1e-11


def test_fp32_adjoint_close_to_fp64(gpu):
    dims = (16, 16, 16)
    U, E0 = make_state(2, dims, 0.1, 0.05, seed=15)
    l64, g64 = loss_and_gradient(device(dims), U, E0, 40, 0, 20, 0)
    l32, g32 = loss_and_gradient(device(dims, precision="fp32"), U, E0, 40, 10, 20, 0)
    assert abs(l32 - l64) < 1e-5 * abs(l64)
    assert np.linalg.norm(g32 - g64) < 1e-4 * np.linalg.norm(g64)


def test_unsupported_configurations_are_rejected(gpu):
    U3, E3 = make_state(3, (4, 4, 4), 0.1, 0.1, seed=1)
    g3 = Grid(3, (4, 4, 4), 1.0, 0.25, 1.0, "fp64")
    with pytest.raises(OpxError):
        loss_and_gradient(g3, U3, E3, 2)
    U, E0 = make_state(2, (4, 4, 4), 0.1, 0.1, seed=1)
    g2 = device((4, 4, 4))
    g2.setActiveRegion([1, 1, 1], [2, 2, 2])
    with pytest.raises(OpxError):
        loss_and_gradient(g2, U, E0, 2)
    with pytest.raises(OpxError):
        loss_and_gradient(device((4, 4, 4)), U, E0, 0)
