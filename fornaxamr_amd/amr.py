"""Coarse-fine corrections of a refined mesh on device block arrays
(C ABI fx_cf_*).

The reference applies them every stage by walking blocks and neighbour slots
on the host: prj_riemann_flux_send (src/prj_riemann.c:1701-1768) overwrites
each coarse face flux behind a finer neighbour with the mean of the four fine
faces, and prj_mhd_emf_send (src/prj_mhd.c:967-1029) overwrites each coarse
edge EMF along a finer neighbour with the mean of the two fine edges.  Here
the walk happens once per topology (:class:`CoarseFinePlan`); each stage
replays device correction lists.

:func:`links_from_slots` is the host-side walk the reference's own slot
arrays need: it lists, in the reference's loop order (blocks by id, slots
0..55), every slot of a local block whose neighbour is coarser.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native


def links_from_slots(info, slots, rank: int = 0) -> list[_native.CfLink]:
    """Slots with a coarser neighbour of every local block.

    ``info[b] = (id, active, level, rank)`` and ``slots[b, n] = (id, rank,
    rel_level, type, send_loc_start[3], send_loc_end[3], recv_loc_start[3],
    recv_loc_end[3])`` as in ``prj_block`` / ``prj_neighbor``
    (src/prj_types.h:77-213).  A block is local when id >= 0, it is active and
    this rank owns it (src/prj_riemann.c:1710, src/prj_mhd.c:14-18)."""
    info = np.asarray(info)
    slots = np.asarray(slots)
    out = []
    for b in range(len(info)):
        if not (info[b, 0] >= 0 and info[b, 1] == 1 and info[b, 3] == rank):
            continue
        for n in range(slots.shape[1]):
            s = slots[b, n]
            if s[0] < 0 or s[2] >= 0:
                continue
            k = _native.CfLink()
            k.block = b
            k.neighbor = int(s[0])
            k.rel_level = int(s[2])
            k.type = int(s[3])
            k.neighbor_local = int(s[1] == rank)
            for d in range(3):
                k.send_start[d] = int(s[4 + d])
                k.recv_start[d] = int(s[10 + d])
                k.recv_end[d] = int(s[13 + d])
            out.append(k)
    return out


def _addr(t):
    return None if t is None else t.data_ptr()


class CoarseFinePlan:
    """Device correction lists of one mesh topology.

    ``geo[b] = (xmin[3], xmax[3], dx[3])`` for every block; ``local[b]`` marks
    the blocks this rank updates; ``flux[b]`` / ``emf[b]`` / ``fidelity[b]``
    are three device tensors (float64 flux: nvar x 8000 in the reference's
    VIDX layout; float64 emf and int32 fidelity: 8820 each, EDGE_IDX) or
    None.  The tensors must stay alive and in place while the plan is used."""

    def __init__(self, geo, local, links, nvar, flux=None, emf=None, fidelity=None):
        self.lib = _native.load()
        geo = np.asarray(geo, dtype=np.float64)
        nb = len(geo)
        blocks = (_native.CfBlock * nb)()
        for b in range(nb):
            blk = blocks[b]
            for d in range(3):
                blk.xmin[d] = geo[b, d]
                blk.xmax[d] = geo[b, 3 + d]
                blk.dx[d] = geo[b, 6 + d]
                blk.flux[d] = _addr(flux[b][d]) if flux is not None and flux[b] is not None else None
                blk.emf[d] = _addr(emf[b][d]) if emf is not None and emf[b] is not None else None
                blk.edge_fidelity[d] = (_addr(fidelity[b][d])
                                        if fidelity is not None and fidelity[b] is not None else None)
            blk.local = int(bool(local[b]))
        arr = (_native.CfLink * max(len(links), 1))(*links)
        h = C.c_void_p()
        _native.check(self.lib.fx_cf_plan_create(nb, blocks, len(links), arr, int(nvar),
                                                 C.byref(h)))
        self.h = h
        self._keep = (flux, emf, fidelity)

    def info(self) -> dict:
        v = (C.c_int * 4)()
        _native.check(self.lib.fx_cf_plan_info(self.h, v))
        return {"flux_records": v[0], "flux_levels": v[1], "emf_records": v[2],
                "emf_levels": v[3]}

    def flux_send(self, stream=None):
        """prj_riemann_flux_send (local phase), stream-ordered."""
        _native.check(self.lib.fx_cf_flux_send(self.h, _stream(stream)))

    def emf_send(self, status=None, stream=None):
        """prj_mhd_emf_send, stream-ordered; ``status`` is a zeroed device
        int32 tensor of 4 entries or None."""
        _native.check(self.lib.fx_cf_emf_send(self.h, _addr(status), _stream(stream)))

    def close(self):
        if getattr(self, "h", None):
            self.lib.fx_cf_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _stream(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return getattr(stream, "cuda_stream", stream)
