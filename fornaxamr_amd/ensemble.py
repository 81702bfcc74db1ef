"""Ragged ensembles of independent hydro meshes on one GPU (C ABI fx_ensemble_*).

The reference runs one problem per process; an ensemble advances many
independent problems of different sizes and gammas together.  Schedules:

* ``mode=0``: every instance's kernels are launched separately (the
  per-instance baseline);
* ``mode=1``: each stage is one batched launch over a table of
  (instance, block) entries; ``persistent_wgs > 0`` uses that many persistent
  workgroups draining the table through atomic queues.

``autotune()`` times the candidate schedules on the live state and keeps the
fastest; ``set_graph(True)`` captures a whole step into a hipGraph.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .problems import Problem

_dp = C.POINTER(C.c_double)


class GpuEnsemble:
    def __init__(self, problems: list[Problem]):
        self.lib = _native.load()
        n = len(problems)
        descs = (_native.MeshDesc * n)()
        for d, p in zip(descs, problems):
            if p.mhd:
                raise ValueError("ensembles are hydro only")
            for i in range(3):
                d.n[i] = int(p.n[i])
            for i in range(6):
                d.xlim[i] = float(p.xlim[i])
                d.bc[i] = int(p.bc[i])
            d.gamma = float(p.gamma)
            d.mhd = 0
            d.eta = 0.0
        h = C.c_void_p()
        _native.check(self.lib.fx_ensemble_create(n, descs, C.byref(h)))
        self.h = h
        self.problems = problems
        # the state reset() and autotune() return to: the problems' initial
        # states, replaced per instance by set_state
        self.states = [None] * n
        for i, p in enumerate(problems):
            self.set_state(i, p.W)
        nb = C.c_int()
        nc = C.c_longlong()
        _native.check(self.lib.fx_ensemble_info(self.h, C.byref(nb), C.byref(nc)))
        self.nblocks = nb.value
        self.ncells = nc.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.fx_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        """Back to the reference states (both slots) and a fresh dt history:
        each instance's initial state, or the last one set_state gave it."""
        for i, W in enumerate(self.states):
            _native.check(self.lib.fx_ensemble_set_prim(self.h, i, W.ctypes.data_as(_dp)))

    def set_state(self, i, W):
        """State of instance i (both slots) and a fresh dt history for it; it
        also becomes the state reset() and autotune() return instance i to."""
        i = int(i)
        if not 0 <= i < len(self.problems):
            raise IndexError(f"instance {i} out of range")
        W = np.array(W, dtype=np.float64, order="C", copy=True)
        want = (6,) + tuple(self.problems[i].n)
        if W.shape != want:
            raise ValueError(f"instance {i}: W has shape {W.shape}, expected {want}")
        _native.check(self.lib.fx_ensemble_set_prim(self.h, i, W.ctypes.data_as(_dp)))
        self.states[i] = W

    def set_schedule(self, mode=1, persistent_wgs=0, xcd=1, table=0):
        _native.check(self.lib.fx_ensemble_set_schedule(self.h, int(mode), int(persistent_wgs),
                                                        int(xcd), int(table)))

    def set_graph(self, on=True):
        _native.check(self.lib.fx_ensemble_set_graph(self.h, 1 if on else 0))

    def step_async(self, cfl=0.4, dt_factor=1.2):
        _native.check(self.lib.fx_ensemble_step_async(self.h, float(cfl), float(dt_factor)))

    def sync(self):
        _native.check(self.lib.fx_ensemble_sync(self.h))

    def last_dt(self):
        out = np.empty(len(self.problems))
        _native.check(self.lib.fx_ensemble_last_dt(self.h, out.ctypes.data_as(_dp)))
        return out

    def get_prim(self, i, slot=0):
        p = self.problems[i]
        W = np.empty((6,) + tuple(p.n))
        _native.check(self.lib.fx_ensemble_get_prim(self.h, int(i), int(slot), W.ctypes.data_as(_dp)))
        return W

    def autotune(self, cfl=0.4, dt_factor=1.2, reps=3):
        """Pick the fastest schedule; the state is reset afterwards (see reset)."""
        out = (C.c_int * 4)()
        ms = C.c_double()
        _native.check(self.lib.fx_ensemble_autotune(self.h, float(cfl), float(dt_factor), int(reps),
                                                    out, C.byref(ms)))
        self.reset()
        return {"mode": out[0], "persistent_wgs": out[1], "xcd": out[2], "table": out[3],
                "ms_per_step": ms.value}
