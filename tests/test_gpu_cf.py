"""Coarse-fine corrections on the GPU (fx_cf_*) against the oracle.

On the committed refined-mesh topologies (built by the reference's own AMR
code, tests/golden/cf_mesh_*.npz) the device prj_riemann_flux_send (6 and 9
flux components) and prj_mhd_emf_send (with edge fidelity) must give the
oracle's arrays bit for bit -- every entry of every block, untouched ones
included -- and hash to the reference's SHA-256 in
tests/golden/cf_corrections.json.  Further cases: replaying a plan is
idempotent, a non-finite restricted EMF is reported in the status word, a
synthetic three-level chain (a block that both receives and sends along the
same edge line, in either block order) needs two ordered launches and still
matches the sequential oracle, and a plan without storage refuses to run.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "cf_corrections.json")))


def load_mesh(name):
    return O.AmrMesh.load(os.path.join(HERE, "golden", f"cf_mesh_{name}.npz"))


def to_dev(arrays):
    import torch
    return [None if a is None else [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in a]
            for a in arrays]


def from_dev(tensors):
    return [None if t is None else [x.cpu().numpy() for x in t] for t in tensors]


def make_plan(mesh, nvar, flux=None, emf=None, fid=None):
    from fornaxamr_amd.amr import CoarseFinePlan, links_from_slots
    links = links_from_slots(mesh.info, mesh.slots)
    return CoarseFinePlan(mesh.geo, mesh.local(), links, nvar, flux=flux, emf=emf, fidelity=fid)


def assert_same(got, want, what):
    for b, (g, w) in enumerate(zip(got, want)):
        for d in range(3):
            assert np.array_equal(g[d], w[d]), (what, b, d)


@pytest.mark.parametrize("name", sorted(GOLD))
@pytest.mark.parametrize("nvar", [6, 9])
def test_flux_send_matches_oracle(gpu, name, nvar):
    import torch
    g = GOLD[name]
    mesh = load_mesh(name)
    key = f"flux{nvar}"
    host = O.cf_inputs(mesh, nvar, g["seeds"][key])
    dev = to_dev(host)
    plan = make_plan(mesh, nvar, flux=dev)
    info = plan.info()
    assert info["flux_records"] > 0 and info["flux_levels"] >= 1
    plan.flux_send()
    torch.cuda.synchronize()
    O.cf_flux_send(mesh, host, nvar)
    got = from_dev(dev)
    assert_same(got, host, key)
    assert O.cf_digest(got) == g[key + "_out"]
    # the corrected faces are means of fine faces the correction does not
    # touch: a second replay changes nothing
    plan.flux_send()
    torch.cuda.synchronize()
    assert_same(from_dev(dev), host, key + " replay")
    plan.close()


@pytest.mark.parametrize("name", sorted(GOLD))
def test_emf_send_matches_oracle(gpu, name):
    import torch
    g = GOLD[name]
    mesh = load_mesh(name)
    host = O.cf_emf_inputs(mesh, g["seeds"]["emf"])
    fid_h = [[np.full(O.RB_EDGES, 7, np.int32) for _ in range(3)] for _ in range(mesh.nblocks)]
    dev, fid_d = to_dev(host), to_dev(fid_h)
    status = torch.zeros(4, dtype=torch.int32, device="cuda")
    plan = make_plan(mesh, 9, emf=dev, fid=fid_d)
    info = plan.info()
    assert info["emf_records"] > 0 and info["emf_levels"] >= 1
    plan.emf_send(status)
    torch.cuda.synchronize()
    assert O.cf_emf_send(mesh, host, fid_h) == 0
    assert status.cpu().tolist() == [0, 0, 0, 0]
    got, got_fid = from_dev(dev), from_dev(fid_d)
    assert_same(got, host, "emf")
    loc = mesh.local()
    for b in np.flatnonzero(loc):
        for d in range(3):
            assert np.array_equal(got_fid[b][d], fid_h[b][d]), ("fidelity", b, d)
    assert O.cf_digest(got) == g["emf_out"]
    assert O.cf_digest(got_fid, loc) == g["fidelity_local_out"]
    plan.emf_send(status)
    torch.cuda.synchronize()
    assert_same(from_dev(dev), host, "emf replay")
    plan.close()


def test_emf_send_reports_non_finite(gpu):
    import torch
    mesh = load_mesh("two_level")
    host = O.cf_emf_inputs(mesh, 3)
    loc = mesh.local()
    f = next(b for b in np.flatnonzero(loc) if (mesh.slots[b, :, 2] < 0).any())
    for d in range(3):
        host[f][d][:] = np.inf
    dev = to_dev(host)
    status = torch.zeros(4, dtype=torch.int32, device="cuda")
    plan = make_plan(mesh, 9, emf=dev)
    plan.emf_send(status)
    torch.cuda.synchronize()
    st = status.cpu().tolist()
    assert st[0] == 1
    # the reported coarse block is one that f restricts onto
    coarse = {int(s[0]) for s in mesh.slots[f] if s[0] >= 0 and s[2] < 0}
    assert st[1] in coarse and 0 <= st[2] < 3 and 0 <= st[3] < O.RB_EDGES
    fid = [[np.zeros(O.RB_EDGES, np.int32) for _ in range(3)] for _ in range(mesh.nblocks)]
    assert O.cf_emf_send(mesh, host, fid) == 1
    plan.close()


# ---- a synthetic chain: level 2 -> level 1 -> level 0 along one edge line ----
def _chain(order):
    """Blocks C (level 0) [0,1]^3, B (level 1) [1,1.5]x[0,.5]x[0,.5] and F
    (level 2) [1,1.25]x[.5,.75]x[0,.25]: F restricts onto B's edges along
    x = 1, y = .5, which B restricts onto C.  ``order`` gives the block ids of
    C, B, F.  Slot data follow the rel_level = -1 rules of
    prj_neighbor_compute_geometry (src/prj_mesh.c:111-133)."""
    boxes = {"C": ((0, 0, 0), (1, 1, 1), 0), "B": ((1, 0, 0), (1.5, .5, .5), 1),
             "F": ((1, .5, 0), (1.25, .75, .25), 2)}
    nb = 3
    geo = np.zeros((nb, 9))
    info = np.zeros((nb, 4), np.int32)
    slots = np.full((nb, 56, 16), 0, np.int32)
    slots[:, :, 0] = -1
    ids = dict(zip("CBF", order))
    for k, (lo, hi, lev) in boxes.items():
        b = ids[k]
        dx = (1.0 / 16) / 2 ** lev
        geo[b] = [*lo, *hi, dx, dx, dx]
        info[b] = [b, 1, lev, 0]
    for fine, coarse in (("B", "C"), ("F", "B")):
        a, nbr = ids[fine], ids[coarse]
        rec = [nbr, 0, -1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0]
        for d in range(3):
            if abs(geo[a][3 + d] - geo[nbr][d]) < 1e-12:        # neighbour above
                ss, se, rs, re = 16 - 4, 16, -2, 0
            elif abs(geo[nbr][3 + d] - geo[a][d]) < 1e-12:      # neighbour below
                ss, se, rs, re = 0, 4, 16, 18
            else:
                above = geo[nbr][d] + geo[nbr][3 + d] > geo[a][d] + geo[a][3 + d]
                ss, se = 0, 16
                rs, re = (0, 8) if above else (8, 16)
            rec[4 + d], rec[7 + d], rec[10 + d], rec[13 + d] = ss, se, rs, re
        slots[a, 0] = rec
    return O.AmrMesh(geo, info, slots)


@pytest.mark.parametrize("order", [(0, 1, 2), (0, 2, 1), (2, 1, 0)])
def test_chained_restriction_keeps_sequential_order(gpu, order):
    import torch
    mesh = _chain(order)
    host = O.cf_emf_inputs(mesh, 11)
    fid_h = [[np.zeros(O.RB_EDGES, np.int32) for _ in range(3)] for _ in range(3)]
    dev, fid_d = to_dev(host), to_dev(fid_h)
    plan = make_plan(mesh, 9, emf=dev, fid=fid_d)
    assert plan.info()["emf_levels"] == 2
    status = torch.zeros(4, dtype=torch.int32, device="cuda")
    plan.emf_send(status)
    torch.cuda.synchronize()
    assert O.cf_emf_send(mesh, host, fid_h) == 0
    assert status.cpu().tolist() == [0, 0, 0, 0]
    assert_same(from_dev(dev), host, "chain emf")
    assert_same(from_dev(fid_d), fid_h, "chain fidelity")
    # flux: the chain's corrections touch different face arrays, one level
    fl = O.cf_inputs(mesh, 6, 12)
    fd = to_dev(fl)
    plan2 = make_plan(mesh, 6, flux=fd)
    assert plan2.info()["flux_levels"] == 1
    plan2.flux_send()
    torch.cuda.synchronize()
    O.cf_flux_send(mesh, fl, 6)
    assert_same(from_dev(fd), fl, "chain flux")
    plan.close()
    plan2.close()


def test_plan_errors(gpu):
    from fornaxamr_amd import _native
    from fornaxamr_amd.amr import CoarseFinePlan, links_from_slots
    mesh = load_mesh("two_level")
    links = links_from_slots(mesh.info, mesh.slots)
    # no storage: both parts unavailable, the send calls refuse
    plan = CoarseFinePlan(mesh.geo, mesh.local(), links, 6)
    with pytest.raises(_native.FxError, match="flux correction unavailable"):
        plan.flux_send()
    with pytest.raises(_native.FxError, match="EMF restriction unavailable"):
        plan.emf_send()
    plan.close()
    # slot ranges that would leave the block arrays
    bad = links_from_slots(mesh.info, mesh.slots)
    flux = to_dev(O.cf_inputs(mesh, 6, 1))
    for k in bad:
        if k.type == 1:
            for d in range(3):
                k.send_start[d] = 40
    plan = CoarseFinePlan(mesh.geo, mesh.local(), bad, 6, flux=flux)
    with pytest.raises(_native.FxError, match="outside"):
        plan.flux_send()
    plan.close()
    lib = _native.load()
    h = C.c_void_p()
    assert lib.fx_cf_plan_create(0, None, 0, None, 6, C.byref(h)) == -1
    assert lib.fx_cf_flux_send(None, None) == -1
    assert lib.fx_cf_emf_send(None, None, None) == -1


def test_plan_without_links(gpu):
    """A mesh with no coarser neighbours (uniform): empty lists, the send calls
    are no-ops, the fidelity reset still runs over the local blocks."""
    import torch
    from fornaxamr_amd.amr import CoarseFinePlan
    mesh = load_mesh("two_level")
    emf = to_dev(O.cf_emf_inputs(mesh, 5))
    fid = to_dev([[np.full(O.RB_EDGES, 7, np.int32) for _ in range(3)]
                  for _ in range(mesh.nblocks)])
    plan = CoarseFinePlan(mesh.geo, mesh.local(), [], 9, emf=emf, fidelity=fid)
    assert plan.info() == {"flux_records": 0, "flux_levels": 0, "emf_records": 0,
                           "emf_levels": 0}
    before = from_dev(emf)
    plan.flux_send()
    plan.emf_send()
    torch.cuda.synchronize()
    assert_same(from_dev(emf), before, "emf untouched")
    loc = mesh.local()
    got = from_dev(fid)
    for b in np.flatnonzero(loc):
        assert set(np.unique(got[b][0]).tolist()) == {0, 2}
    plan.close()
