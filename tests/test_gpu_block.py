"""Per-block entry points on the reference's own block layouts.

fx_flux_update_block (the drop-in form of one prj_flux_update call) and
fx_flux_div_block against the reference: bit-identical outputs, with every
entry the reference leaves untouched also left untouched (NaN pre-fill), for
hydro (HLLC) and MHD (HLLD, both stage slots).  The reference outputs come from
the serial build when it is present (oracle/_ref) and are always checked
against the committed SHA-256 fixture made from that build.  The batched
device form fx_flux_update_blocks must give the same bits for several blocks
in one launch.
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "block_flux.json")))
_dp = C.POINTER(C.c_double)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def _p(a):
    return a.ctypes.data_as(C.c_void_p).value


def _run_block(lib, mhd, slot, slab, eos, bf):
    from fornaxamr_amd import _native
    nh = 9 if mhd else 6
    io = _native.BlockIO()
    W = np.ascontiguousarray(slab.reshape(nh, -1))
    io.W = _p(W) + slot * O.RB_CELLS * 8
    io.var_stride = 2 * O.RB_CELLS
    io.eosvar = _p(eos)
    out = {"flux": [np.full(nh * O.RB_CELLS, np.nan) for _ in range(3)],
           "v_riemann": [np.full(3 * O.RB_CELLS, np.nan) for _ in range(3)]}
    if mhd:
        out["Bv1"] = [np.full(O.RB_CELLS, np.nan) for _ in range(3)]
        out["Bv2"] = [np.full(O.RB_CELLS, np.nan) for _ in range(3)]
    for d in range(3):
        io.flux[d] = _p(out["flux"][d])
        io.v_riemann[d] = _p(out["v_riemann"][d])
        if mhd:
            io.bf[d] = _p(bf[d])
            io.Bv1[d] = _p(out["Bv1"][d])
            io.Bv2[d] = _p(out["Bv2"][d])
    _native.check(lib.fx_flux_update_block(C.byref(io), mhd))
    return out


@pytest.mark.parametrize("case", sorted(GOLD))
def test_flux_update_block_matches_reference(gpu, case):
    g = GOLD[case]
    mhd, seed, slot = g["mhd"], g["seed"], g["slot"]
    slab, gam, bf = O.block_inputs(seed, mhd)
    eos = O.block_eosvar(slab, slot, gam)
    assert _sha(slab) == g["input_sha"] and _sha(eos) == g["eosvar_sha"]
    out = _run_block(gpu, mhd, slot, slab, eos, bf)
    for k, arrs in out.items():
        assert [_sha(a) for a in arrs] == g["out_sha"][k], (case, k)
    if O.ref_available():
        ref = O.ref_flux_update_raw(mhd, slot, slab, eos, bf)
        for k in out:
            for d in range(3):
                assert np.array_equal(out[k][d], ref[k][d], equal_nan=True), (case, k, d)
    # prj_flux_div on those fluxes
    nh = 9 if mhd else 6
    fd = np.empty((nh, 16, 16, 16))
    fl = (_dp * 3)(*[a.ctypes.data_as(_dp) for a in out["flux"]])
    area = np.array(g["area"])
    from fornaxamr_amd import _native
    _native.check(gpu.fx_flux_div_block(fl, area.ctypes.data_as(_dp), g["vol"], nh,
                                        fd.ctypes.data_as(_dp)))
    assert _sha(fd) == g["flux_div_sha"]


@pytest.mark.parametrize("shift", [0, 1], ids=["aligned", "offset8"])
def test_batched_blocks_equal_single_block(gpu, shift):
    """Several blocks in one launch give the single-block bits; output arrays
    one double off a 16-byte boundary (shift 1) take the kernel's per-lane
    store path instead of its paired 16-byte stores."""
    import torch
    from fornaxamr_amd import _native
    cases = [(0, 11, 0), (0, 12, 1), (0, 13, 1)]
    singles, blocks, keep = [], [], []
    dev = torch.device("cuda")
    for mhd, seed, slot in cases:
        slab, gam, bf = O.block_inputs(seed, mhd)
        eos = O.block_eosvar(slab, slot, gam)
        singles.append(_run_block(gpu, mhd, slot, slab, eos, bf))
        W = torch.from_numpy(np.ascontiguousarray(slab.reshape(6, -1))).to(dev)
        E = torch.from_numpy(eos).to(dev)
        F = [torch.full((6 * O.RB_CELLS + shift,), float("nan"), dtype=torch.float64,
                        device=dev)[shift:] for _ in range(3)]
        V = [torch.full((3 * O.RB_CELLS + shift,), float("nan"), dtype=torch.float64,
                        device=dev)[shift:] for _ in range(3)]
        assert all((t.data_ptr() % 16 == 0) == (shift == 0) for t in F + V)
        io = _native.BlockIO()
        io.W = W.data_ptr() + slot * O.RB_CELLS * 8
        io.var_stride = 2 * O.RB_CELLS
        io.eosvar = E.data_ptr()
        for d in range(3):
            io.flux[d] = F[d].data_ptr()
            io.v_riemann[d] = V[d].data_ptr()
        blocks.append(io)
        keep.append((W, E, F, V))
    arr = (_native.BlockIO * len(blocks))(*blocks)
    raw = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(dev)
    status = torch.zeros(8, dtype=torch.int32, device=dev)
    _native.check(gpu.fx_flux_update_blocks(len(blocks), C.c_void_p(raw.data_ptr()), 0,
                                            C.c_void_p(status.data_ptr()),
                                            C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    for b, (W, E, F, V) in enumerate(keep):
        for d in range(3):
            assert np.array_equal(F[d].cpu().numpy(), singles[b]["flux"][d], equal_nan=True)
            assert np.array_equal(V[d].cpu().numpy(), singles[b]["v_riemann"][d], equal_nan=True)


def test_bad_block_arguments(gpu):
    from fornaxamr_amd import _native
    io = _native.BlockIO()
    assert gpu.fx_flux_update_block(C.byref(io), 0) == -1
    assert gpu.fx_flux_update_blocks(1, None, 0, None, None) == -1
    assert gpu.fx_flux_update_blocks(1, C.c_void_p(8), 1, None, None) == -1


# ---- prj_mhd_bf2bc / prj_mhd_bf2bc_all (src/prj_mhd.c:555-625) ----------------
BF2BC = json.load(open(os.path.join(HERE, "golden", "bf2bc.json")))


def _bf2bc_host(lib, slot, slab, bf, done):
    from fornaxamr_amd import _native
    out = np.ascontiguousarray(slab, dtype=np.float64).copy()
    io = _native.Bf2bcIO()
    io.W = _p(out) + slot * O.RB_CELLS * 8
    io.var_stride = 2 * O.RB_CELLS
    bfa = [np.ascontiguousarray(b) for b in bf]
    for d in range(3):
        io.bf[d] = _p(bfa[d])
    dn = np.ascontiguousarray(done, dtype=np.int32) if done is not None else None
    io.done = _p(dn) if dn is not None else None
    rc = lib.fx_bf2bc_block(C.byref(io))
    return rc, out


@pytest.mark.parametrize("case", sorted(BF2BC))
def test_bf2bc_block_matches_reference(gpu, case):
    g = BF2BC[case]
    slab, _, bf = O.block_inputs(g["seed"], True)
    done = O.bf2bc_done_flags(g["seed"]) if g["use_done"] else None
    rc, out = _bf2bc_host(gpu, g["slot"], slab, bf, done)
    assert rc == 0
    assert _sha(out) == g["out_sha"]
    if O.ref_available():
        assert np.array_equal(out, O.ref_bf2bc_raw(g["slot"], slab, bf, done))


def test_bf2bc_blocks_device_batch(gpu):
    """fx_bf2bc_blocks over three device-resident blocks in one launch: the
    same bits as the per-block host calls."""
    import torch
    from fornaxamr_amd import _native
    names = sorted(BF2BC)
    slabs, bfs, dones, want = [], [], [], []
    for name in names:
        g = BF2BC[name]
        slab, _, bf = O.block_inputs(g["seed"], True)
        done = O.bf2bc_done_flags(g["seed"]) if g["use_done"] else None
        slabs.append(torch.from_numpy(slab.copy()).cuda())
        bfs.append([torch.from_numpy(np.ascontiguousarray(b)).cuda() for b in bf])
        dones.append(torch.from_numpy(done).cuda() if done is not None else None)
        want.append(g["out_sha"])
    descs = (_native.Bf2bcIO * len(names))()
    for b, name in enumerate(names):
        io = descs[b]
        io.W = slabs[b].data_ptr() + BF2BC[name]["slot"] * O.RB_CELLS * 8
        io.var_stride = 2 * O.RB_CELLS
        for d in range(3):
            io.bf[d] = bfs[b][d].data_ptr()
        io.done = dones[b].data_ptr() if dones[b] is not None else None
    raw = torch.frombuffer(bytearray(bytes(descs)), dtype=torch.uint8).cuda()
    status = torch.zeros(8, dtype=torch.int32, device="cuda")
    _native.check(gpu.fx_bf2bc_blocks(len(names), raw.data_ptr(), status.data_ptr(), None))
    torch.cuda.synchronize()
    assert status.cpu().tolist() == [0] * 8
    for b in range(len(names)):
        assert _sha(slabs[b].cpu().numpy()) == want[b]


def test_bf2bc_nonfinite_field_reports(gpu):
    """A non-finite face average comes back as the reference's abort message."""
    slab, _, bf = O.block_inputs(31, True)
    bf = [b.copy() for b in bf]
    bf[1][1234] = np.inf
    rc, _ = _bf2bc_host(gpu, 0, slab, bf, None)
    assert rc == -3
    assert b"non-finite cell-centered magnetic field" in gpu.fx_last_error()
