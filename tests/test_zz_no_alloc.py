"""fx_flux_update_block_noalloc (prj_flux_update on host-array blocks staged
through the library's fixed device array, no allocation) against
fx_flux_update_block on the same blocks: bit-identical outputs, with the
entries prj_flux_update leaves untouched left untouched (NaN pre-fill).  Two
calls in a row, the second on fewer blocks, then a hydro call; sizes outside
1..FX_NOALLOC_MAX_BLOCKS are refused."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu
MAX_BLOCKS = 8


def _p(a):
    return a.ctypes.data_as(C.c_void_p).value


def _block(mhd, seed, slot):
    """One block's inputs and NaN-filled outputs, and its descriptor."""
    from fornaxamr_amd import _native
    nh = 9 if mhd else 6
    slab, gam, bf = O.block_inputs(seed, mhd)
    eos = O.block_eosvar(slab, slot, gam)
    W = np.ascontiguousarray(slab.reshape(nh, -1))
    out = {"flux": [np.full(nh * O.RB_CELLS, np.nan) for _ in range(3)],
           "v_riemann": [np.full(3 * O.RB_CELLS, np.nan) for _ in range(3)]}
    if mhd:
        out["Bv1"] = [np.full(O.RB_CELLS, np.nan) for _ in range(3)]
        out["Bv2"] = [np.full(O.RB_CELLS, np.nan) for _ in range(3)]
    io = _native.BlockIO()
    io.W = _p(W) + slot * O.RB_CELLS * 8
    io.var_stride = 2 * O.RB_CELLS
    io.eosvar = _p(eos)
    for d in range(3):
        io.flux[d] = _p(out["flux"][d])
        io.v_riemann[d] = _p(out["v_riemann"][d])
        if mhd:
            io.bf[d] = _p(bf[d])
            io.Bv1[d] = _p(out["Bv1"][d])
            io.Bv2[d] = _p(out["Bv2"][d])
    return io, out, (W, eos, bf)


def _check_call(lib, mhd, seeds):
    from fornaxamr_amd import _native
    blocks = [_block(mhd, s, s % 2) for s in seeds]
    singles = [_block(mhd, s, s % 2) for s in seeds]
    arr = (_native.BlockIO * len(blocks))(*[b[0] for b in blocks])
    _native.check(lib.fx_flux_update_block_noalloc(len(blocks), arr, mhd, None))
    for io, out, _ in singles:
        _native.check(lib.fx_flux_update_block(C.byref(io), mhd))
    for (_, got, _), (_, ref, _) in zip(blocks, singles):
        for k in ref:
            for d in range(3):
                assert np.array_equal(got[k][d], ref[k][d], equal_nan=True), (mhd, k, d)
                assert not np.all(np.isnan(got[k][d]))


def test_noalloc_matches_block_entry(gpu):
    _check_call(gpu, 1, [41, 42, 43, 44, 45])
    _check_call(gpu, 1, [46, 47])              # fewer blocks: the array keeps the last call's
    _check_call(gpu, 0, [48, 49, 50])
    _check_call(gpu, 1, list(range(51, 51 + MAX_BLOCKS)))


def test_noalloc_refuses_sizes_beyond_the_array(gpu):
    from fornaxamr_amd import _native
    io, _, _ = _block(0, 60, 0)
    arr = (_native.BlockIO * (MAX_BLOCKS + 1))(*[io] * (MAX_BLOCKS + 1))
    assert gpu.fx_flux_update_block_noalloc(MAX_BLOCKS + 1, arr, 0, None) == -1
    assert b"outside 1..8" in gpu.fx_last_error()
    assert gpu.fx_flux_update_block_noalloc(0, arr, 0, None) == -1
    assert gpu.fx_flux_update_block_noalloc(1, None, 0, None) == -1
