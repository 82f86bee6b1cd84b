"""Batched events (one persistent kernel per step, optional captured hipGraph)
against per-event oracle runs and against eager single-event stepping."""
import numpy as np
import pytest

from openpixi_amd import FastTYMSolver
from openpixi_amd.batch import Batch

from .helpers import make_state, oracle_grid, rel_err

pytestmark = pytest.mark.gpu

EVENT_DIMS = [(6, 6, 6), (8, 5, 7), (4, 9, 6), (10, 4, 5), (5, 5, 11)]


def filled_batch(nc, precision, dims_list=EVENT_DIMS, sigma_a=0.1, sigma_e=0.05):
    b = Batch(nc, dims_list, 1.0, 0.25, 1.0, precision)
    states = []
    for e, dims in enumerate(dims_list):
        U, E = make_state(nc, dims, sigma_a, sigma_e, seed=100 + e)
        b.event(e).setU(U)
        b.event(e).setE(E)
        states.append((U, E))
    return b, states


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("nc", [2, 3])
def test_batch_matches_oracle_fp64(gpu, nc, graph):
    print(f"batch fp64 nc={nc} graph={graph}", flush=True)
    b, states = filled_batch(nc, "fp64")
    b.step_links(0.25)
    steps = 14
    b.leapfrog(steps, 0, graph)
    for e, dims in enumerate(EVENT_DIMS):
        og = oracle_grid(nc, dims, *states[e])
        og.step_links(0.25)
        for _ in range(steps):
            og.store_fields()
            og.step(0.25)
        assert rel_err(b.event(e).getU(), og.U) < 1e-12, e
        assert rel_err(b.event(e).getE(), og.E) < 1e-12, e
        m = b.measure()[e]
        mo = og.measure()
        assert np.allclose(m[:7], mo[:7], rtol=1e-11, atol=1e-300), e


@pytest.mark.parametrize("form", ["link", "cell"])
@pytest.mark.parametrize("nc", [2, 3])
def test_graph_replays_match_eager(gpu, nc, form):
    """Replays of the captured period (10 steps incl. reunitarization, which the cell form does
    inside its update) interleaved with eager work agree exactly with eager stepping after every
    replay."""
    print(f"graph replays nc={nc} form={form}", flush=True)
    bg, _ = filled_batch(nc, "fp32")
    be, _ = filled_batch(nc, "fp32")
    bg.set_form(form)
    be.set_form(form)
    bg.step_links(0.25)
    be.step_links(0.25)
    for r in range(4):
        bg.leapfrog(10, 10, graph=True)
        be.leapfrog(10, 10, graph=False)
        mg = bg.measure()   # eager work and synchronisation between replays
        me = be.measure()
        assert np.array_equal(mg, me), r
        for e in range(len(EVENT_DIMS)):
            assert np.array_equal(bg.event(e).getU(), be.event(e).getU()), (r, e)
            assert np.array_equal(bg.event(e).getE(), be.event(e).getE()), (r, e)


@pytest.mark.parametrize("form,kernel", [(None, "link"), ("cell", "cell")])
@pytest.mark.parametrize("nc", [2, 3])
def test_batch_matches_single_contexts(gpu, nc, form, kernel):
    """An event of the batch evolves as the same lattice in its own context, FP32 with the
    reprojection every 10 steps (inside the update in the cell forms)."""
    b, states = filled_batch(nc, "fp32")
    if form is not None:
        b.set_form(form)
    b.step_links(0.25)
    b.leapfrog(20, 10, graph=True)
    from openpixi_amd import Grid, _lib

    for e, dims in enumerate(EVENT_DIMS):
        g = Grid(nc, dims, 1.0, 0.25, 1.0, "fp32")
        g.setKernel(kernel)
        g.setU(states[e][0])
        g.setE(states[e][1])
        FastTYMSolver().stepLinks(g, 0.25)
        g._check(_lib.load().opx_leapfrog(g.handle, 20, 10))
        assert rel_err(b.event(e).getU(), g.getU()) < 1e-6, e
        assert rel_err(b.event(e).getE(), g.getE()) < 1e-5, e


def test_batch_autotune_and_uneven_sizes(gpu):
    dims_list = [(32, 32, 32), (48, 48, 48), (16, 24, 40)]
    b, states = filled_batch(2, "fp32", dims_list)
    ms = b.autotune(reps=2)
    assert ms > 0
    info = b.info()
    assert info["count"] == 3 and info["total_cells"] == sum(int(np.prod(d)) for d in dims_list)
    b.step_links(0.25)
    b.leapfrog(6, 0, graph=True)
    og = oracle_grid(2, dims_list[2], *states[2])
    og.step_links(0.25)
    for _ in range(6):
        og.store_fields()
        og.step(0.25)
    assert rel_err(b.event(2).getU(), og.U) < 2e-6


def test_stepping_an_event_view_is_rejected(gpu):
    from openpixi_amd import OpxError

    b, _ = filled_batch(2, "fp64", [(4, 4, 4), (5, 4, 3)])
    with pytest.raises(OpxError):
        FastTYMSolver().step(b.event(0), 0.25)
    # the batch kernels update every cell of every event: an event has no active region
    with pytest.raises(OpxError, match="active region"):
        b.event(1).setActiveRegion((1, 1, 0), (3, 2, 2))
    # the evaluation region only selects measured cells and stays allowed
    b.event(1).setEvaluationRegion((1, 1, 0), (3, 2, 2))


@pytest.mark.parametrize("nc,form", [(2, "tile"), (2, "tile-cached"), (2, "cell"), (3, "cell")])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_tile_form_matches_oracle(gpu, graph, nc, form):
    """The LDS-tile (SU(2)) and cell batch forms against per-event oracle runs."""
    dims_list = [(6, 6, 6), (8, 5, 7), (20, 17, 33), (1, 9, 6), (10, 18, 5)]
    if nc == 3:
        dims_list = [(6, 6, 6), (8, 5, 7), (1, 9, 6), (10, 18, 5)]
    b, states = filled_batch(nc, "fp64", dims_list)
    b.set_form(form)
    assert b.form() == form
    b.step_links(0.25)
    b.leapfrog(12, 0, graph)
    for e, dims in enumerate(dims_list):
        og = oracle_grid(nc, dims, *states[e])
        og.step_links(0.25)
        for _ in range(12):
            og.store_fields()
            og.step(0.25)
        assert rel_err(b.event(e).getU(), og.U) < 1e-12, e
        assert rel_err(b.event(e).getE(), og.E) < 1e-12, e


def test_batch_tile_graph_replays_match_eager(gpu):
    bg, _ = filled_batch(2, "fp32")
    be, _ = filled_batch(2, "fp32")
    for b in (bg, be):
        b.set_form("tile")
        b.step_links(0.25)
    for r in range(3):
        bg.leapfrog(10, 10, graph=True)
        be.leapfrog(10, 10, graph=False)
        assert np.array_equal(bg.measure(), be.measure()), r
        for e in range(len(EVENT_DIMS)):
            assert np.array_equal(bg.event(e).getU(), be.event(e).getU()), (r, e)
