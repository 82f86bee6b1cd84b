"""Batched independent events (lattices of uneven sizes) over the C ABI.

Each event is exposed as a Grid-like view (``Batch.event(e)``) for data
exchange and measurements; stepping goes through the batch, one persistent
kernel per step for all events, optionally replayed from a captured hipGraph.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import OpxError
from .grid import Grid


class _EventView(Grid):
    """A Grid bound to one event of a batch (memory owned by the batch)."""

    def __init__(self, batch, handle, num_colors, dims, spacing, time_step, coupling, precision):
        self._batch = batch
        self.num_colors = num_colors
        self.num_cells = list(dims)
        self.spacing = list(spacing)
        self.time_step = time_step
        self.coupling = coupling
        self.precision = precision
        self._h = ctypes.c_void_p(handle)
        self.total_cells = int(np.prod(dims))
        self.simulation_steps = 0

    def close(self):
        self._h = None  # owned by the batch

    def __del__(self):
        pass


class Batch:
    def __init__(self, num_colors, dims_list, spacing=1.0, time_step=0.25, coupling=1.0, precision="fp32"):
        L = _lib.load()
        if np.isscalar(spacing):
            spacing = [float(spacing)] * 3
        self.num_colors = int(num_colors)
        self.dims = [tuple(int(x) for x in d) for d in dims_list]
        self.count = len(self.dims)
        flat = (ctypes.c_int * (3 * self.count))(*[x for d in self.dims for x in d])
        h = ctypes.c_void_p()
        prec = {"fp64": _lib.OPX_FP64, "fp32": _lib.OPX_FP32}[precision]
        rc = L.opx_batch_create(self.num_colors, self.count, flat, _lib.dvec(spacing), float(time_step),
                                float(coupling), prec, ctypes.byref(h))
        if rc != 0:
            raise OpxError(rc, "opx_batch_create failed")
        self._h = h
        self.events = [_EventView(self, L.opx_batch_event(h, e), self.num_colors, self.dims[e], spacing,
                                  time_step, coupling, precision) for e in range(self.count)]

    def _check(self, rc):
        if rc != 0:
            msg = _lib.load().opx_batch_last_error(self._h)
            raise OpxError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().opx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def event(self, e) -> Grid:
        return self.events[e]

    def step_links(self, dt):
        self._check(_lib.load().opx_batch_step_links(self._h, float(dt)))

    def leapfrog(self, steps, reunit_every=0, graph=True):
        self._check(_lib.load().opx_batch_leapfrog(self._h, int(steps), int(reunit_every), 1 if graph else 0))

    def measure(self):
        out = np.zeros(self.count * 8)
        self._check(_lib.load().opx_batch_measure(self._h, _lib.dptr(out)))
        return out.reshape(self.count, 8)

    def sync(self):
        self._check(_lib.load().opx_batch_sync(self._h))

    def autotune(self, reps=3):
        ms = ctypes.c_double()
        self._check(_lib.load().opx_batch_autotune(self._h, int(reps), ctypes.byref(ms)))
        return ms.value

    def set_grid(self, workgroups):
        self._check(_lib.load().opx_batch_set_option(self._h, _lib.BATCH_OPT_GRID, int(workgroups)))

    FORMS = ["link", "tile", "tile-cached", "cell"]

    def set_form(self, form):
        """"link" (per-link persistent kernel), "tile" or "tile-cached" (LDS-tile kernel with
        non-temporal or cached stores of E and Unext; SU(2) only), "cell" (the three links of
        a cell per thread)."""
        self._check(_lib.load().opx_batch_set_option(self._h, _lib.BATCH_OPT_FORM, self.FORMS.index(form)))

    def form(self):
        v = ctypes.c_int()
        self._check(_lib.load().opx_batch_get_option(self._h, _lib.BATCH_OPT_FORM, ctypes.byref(v)))
        return self.FORMS[v.value]

    def info(self):
        c, g, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        t = ctypes.c_int64()
        self._check(_lib.load().opx_batch_get_info(self._h, ctypes.byref(c), ctypes.byref(t), ctypes.byref(g),
                                                   ctypes.byref(s)))
        return {"count": c.value, "total_cells": t.value, "grid": g.value, "steps": s.value, "form": self.form()}

    def set_stream(self, stream_handle):
        self._check(_lib.load().opx_batch_set_stream(self._h, ctypes.c_void_p(int(stream_handle))))
