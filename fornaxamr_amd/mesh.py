"""Host-side mirror of the reference's RK2 driver over the gfx950 C ABI.

``GpuMesh`` plays the part of ``prj_mesh`` + ``prj_timeint_step`` for a
uniform single-level mesh (src/prj_timeint.c:3221-3283, src/main.c:510-545):
state lives on the device; ``stage(1|2, dt)`` is the body of
``prj_timeint_stage1/2`` and ``step()`` the main-loop step with the CFL dt and
its growth limit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .problems import Problem

_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def face_shapes(n):
    nx, ny, nz = n
    return [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]


class GpuMesh:
    def __init__(self, n, xlim, bc, gamma, mhd=False, eta=0.0, stream=None):
        self.lib = _native.load()
        d = _native.MeshDesc()
        for i in range(3):
            d.n[i] = int(n[i])
        for i in range(6):
            d.xlim[i] = float(xlim[i])
            d.bc[i] = int(bc[i])
        d.gamma = float(gamma)
        d.mhd = 1 if mhd else 0
        d.eta = float(eta)
        h = C.c_void_p()
        _native.check(self.lib.fx_mesh_create(C.byref(d), C.byref(h)))
        self.h = h
        self.n = tuple(int(v) for v in n)
        self.mhd = bool(mhd)
        self.nh = 9 if mhd else 6
        if stream is not None:
            _native.check(self.lib.fx_mesh_set_stream(self.h, C.c_void_p(stream)))

    @classmethod
    def from_problem(cls, p: Problem, **kw):
        m = cls(p.n, p.xlim, p.bc, p.gamma, mhd=p.mhd, eta=p.eta, **kw)
        m.set_state(p.W, p.bf)
        return m

    def close(self):
        if getattr(self, "h", None):
            self.lib.fx_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, W, bf=None):
        W = np.ascontiguousarray(W, dtype=np.float64)
        assert W.shape == (self.nh,) + self.n, W.shape
        _native.check(self.lib.fx_mesh_set_prim(self.h, _ptr(W)))
        if self.mhd:
            bf = [np.ascontiguousarray(b, dtype=np.float64) for b in bf]
            for b, s in zip(bf, face_shapes(self.n)):
                assert b.shape == s
            _native.check(self.lib.fx_mesh_set_bf(self.h, *[_ptr(b) for b in bf]))

    def get_prim(self, slot=0):
        W = np.empty((self.nh,) + self.n)
        _native.check(self.lib.fx_mesh_get_prim(self.h, int(slot), _ptr(W)))
        return W

    def get_bf(self, slot=0):
        bf = [np.empty(s) for s in face_shapes(self.n)]
        _native.check(self.lib.fx_mesh_get_bf(self.h, int(slot), *[_ptr(b) for b in bf]))
        return bf

    def calc_dt(self, cfl):
        dt = C.c_double()
        _native.check(self.lib.fx_calc_dt(self.h, float(cfl), C.byref(dt)))
        return dt.value

    def stage(self, s, dt):
        _native.check(self.lib.fx_stage(self.h, int(s), float(dt)))

    INTEGRATORS = {"rk2": 0, "essprk10_4": 1}

    def set_integrator(self, name):
        """Time integrator of step/step_async: "rk2" (default) or "essprk10_4"
        (Ketcheson's eSSPRK(10,4), the reference's default; hydro and MHD)."""
        _native.check(self.lib.fx_mesh_set_integrator(self.h, self.INTEGRATORS[name]))

    def set_mhd_schedule(self, ct_split=0, side_stream=-1):
        """MHD launch schedule: CT workgroups per block (1, 2, 4; 0 automatic)
        and the ghost-layer face kernel beside the flux kernel (1), in order
        (0) or automatic (-1).  Every schedule gives the same bits."""
        _native.check(self.lib.fx_mesh_set_mhd_schedule(self.h, int(ct_split), int(side_stream)))

    RIEMANN = {"hllc": 0, "hlle": 1}

    def set_riemann(self, name):
        """Riemann solver of a hydro mesh: "hllc" (default, the reference's
        prj_riemann_hllc) or "hlle" (HLL over the same wave-speed bounds)."""
        _native.check(self.lib.fx_mesh_set_riemann(self.h, self.RIEMANN[name]))

    def step(self, cfl, dt_factor=1.2):
        dt = C.c_double()
        _native.check(self.lib.fx_step(self.h, float(cfl), float(dt_factor), C.byref(dt)))
        return dt.value

    def step_dt(self, dt):
        """One step of the integrator with the caller's dt (synchronous)."""
        _native.check(self.lib.fx_step_dt(self.h, float(dt)))

    def step_async(self, cfl, dt_factor=1.2):
        _native.check(self.lib.fx_step_async(self.h, float(cfl), float(dt_factor)))

    def sync(self):
        _native.check(self.lib.fx_sync(self.h))

    def set_timing(self, on=True):
        _native.check(self.lib.fx_mesh_set_timing(self.h, 1 if on else 0))

    def kernel_time(self, stage=0):
        ms = C.c_double()
        n = C.c_int()
        _native.check(self.lib.fx_mesh_kernel_time(self.h, int(stage), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def view(self):
        """Device addresses and strides of the state (fx_mesh_view).  Taking a
        view drops the CFL maxima the last stage 2 recorded, so the next step
        reduces over the cells; after writing the state through a view taken
        before a later step, take the view again before stepping."""
        v = _native.MeshView()
        _native.check(self.lib.fx_mesh_view_get(self.h, C.byref(v)))
        return v
