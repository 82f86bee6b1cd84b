"""opx_projected_energy_density_noalloc (the entry point that allocates no device memory; its
per-slice sums live in a static device buffer) against opx_projected_energy_density through the
C ABI: bit-identical results on two contexts in a row, the second smaller, every direction; a
direction with more slices than the static buffer holds is refused."""
import numpy as np
import pytest

from openpixi_amd import FastTYMSolver, _lib

from .helpers import device_grid, make_state

pytestmark = pytest.mark.gpu


def both_forms(grid, direction):
    L = _lib.load()
    n = grid.num_cells[direction]
    ref = [np.full(n, np.nan) for _ in range(4)]
    got = [np.full(n, np.nan) for _ in range(4)]
    grid._check(L.opx_projected_energy_density(grid.handle, direction, *[_lib.dptr(a) for a in ref]))
    grid._check(L.opx_projected_energy_density_noalloc(grid.handle, direction, *[_lib.dptr(a) for a in got]))
    return ref, got


@pytest.mark.parametrize("nc,dims,precision", [(3, (6, 8, 10), "fp64"), (2, (4, 3, 5), "fp32")],
                         ids=["su3-6x8x10-fp64", "su2-4x3x5-fp32"])
def test_noalloc_matches_allocating_entry_point(gpu, nc, dims, precision):
    print(f"no-alloc projected energy nc={nc} dims={dims} {precision}", flush=True)
    U, E = make_state(nc, dims, 0.3, 0.2, seed=43)
    g = device_grid(nc, dims, U, E, precision=precision)
    FastTYMSolver().stepLinks(g, 0.25)
    for d in range(3):
        ref, got = both_forms(g, d)
        for a, b in zip(ref, got):
            assert np.all(np.isfinite(a)) and np.array_equal(a, b), (d, a, b)
        assert np.max(np.abs(ref[0])) > 0 and np.max(np.abs(ref[1])) > 0
    # an evaluation region changes the sums the same way in both
    g.setEvaluationRegion((1, 0, 1), (dims[0] - 2, dims[1] - 1, dims[2] - 2))
    for d in range(3):
        ref, got = both_forms(g, d)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)
    g.close()


def test_noalloc_refuses_more_slices_than_its_buffer(gpu):
    from openpixi_amd import OpxError

    L = _lib.load()
    dims = (4100, 2, 2)
    U, E = make_state(2, dims, 0.1, 0.05, seed=44)
    g = device_grid(2, dims, U, E, precision="fp32")
    out = [np.zeros(dims[0]) for _ in range(4)]
    with pytest.raises(OpxError, match="4096"):
        g._check(L.opx_projected_energy_density_noalloc(g.handle, 0, *[_lib.dptr(a) for a in out]))
    # the short directions of the same context still work and match
    ref, got = both_forms(g, 1)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    g.close()
