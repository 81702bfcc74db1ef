"""The C restatement against the committed golden fixtures (reference outputs).

Fixtures come from tests/golden/make_golden.py (serial reference build): the
reference's own MC face slope, per-face HLLC/HLLD and per-edge EMF functions
on random states, and SHA-256 digests of whole RK2 trajectories.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import pyoracle as O
from helpers import sha

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
_dp = C.POINTER(C.c_double)


def _p(a):
    return np.ascontiguousarray(a).ctypes.data_as(_dp)


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_mc_slope_vectors():
    """The oracle's MC face slope against the reference's on random stencils
    and the limiter's special cases (signed zeros, subnormals, overflow when
    doubled, infinities, NaN): the same bits, signs of zeros included."""
    d = np.load(os.path.join(GOLD, "mc_vectors.npz"))
    lib = O.oracle_lib()
    got = np.array([lib.fxo_mc_slope(a, b, c) for a, b, c in zip(d["qm"], d["q0"], d["qp"])])
    assert np.array_equal(_bits(got), _bits(d["slope"]))
    if O.ref_available("hydro"):
        ref = np.empty_like(got)
        O.ref_lib(False).refh_mc_face_slope(len(got), _p(d["qm"]), _p(d["q0"]), _p(d["qp"]),
                                            _p(ref))
        assert np.array_equal(_bits(ref), _bits(d["slope"]))


def test_hllc_vectors():
    d = np.load(os.path.join(GOLD, "hllc_vectors.npz"))
    lib = O.oracle_lib()
    for i in range(d["WL"].shape[0]):
        f = np.zeros(6)
        v = np.zeros(3)
        lib.fxo_hllc(_p(d["WL"][i]), _p(d["WR"][i]), d["pL"][i], d["pR"][i], d["gL"][i],
                     d["gR"][i], _p(f), _p(v), d["du"][i], d["dv"][i], d["dw"][i])
        assert np.array_equal(f, d["flux"][i]), i
        assert np.array_equal(v, d["v_face"][i]), i


def test_hlld_vectors():
    d = np.load(os.path.join(GOLD, "hlld_vectors.npz"))
    lib = O.oracle_lib()
    for i in range(d["WL"].shape[0]):
        f = np.zeros(9)
        v = np.zeros(3)
        b1 = C.c_double()
        b2 = C.c_double()
        rc = lib.fxo_hlld(_p(d["WL"][i]), _p(d["WR"][i]), d["pL"][i], d["pR"][i], d["gL"][i],
                          d["gR"][i], d["bn"][i], _p(f), _p(v), C.byref(b1), C.byref(b2))
        assert rc == 0
        assert np.array_equal(f, d["flux"][i]), i
        assert np.array_equal(v, d["v_face"][i]), i
        assert (b1.value, b2.value) == tuple(d["bv"][i]), i


def test_emf_vectors():
    d = np.load(os.path.join(GOLD, "emf_vectors.npz"))
    lib = O.oracle_lib()
    for i in range(d["emf"].shape[0]):
        out = C.c_double()
        rc = lib.fxo_emf_upwind(_p(d["emf_face"][i]), _p(d["emf_cell"][i]), _p(d["v_norm"][i]),
                                C.byref(out))
        assert rc == 0 and out.value == d["emf"][i], i


def test_emf_block_vectors():
    """Resistive per-edge EMF (src/prj_mhd.c:627-718 with eta * curl B) of the
    oracle against the reference's own prj_mhd_emf_upwind on one block's face
    arrays: value and the stored edge entry, bit for bit."""
    d = np.load(os.path.join(GOLD, "emf_block_vectors.npz"))
    emf, st = O.oracle_emf_upwind_block(d["edge"], d["emf_face"], d["emf_cell"], d["v_norm"],
                                        [d["bfx"], d["bfy"], d["bfz"]], d["dx"], d["eta"])
    assert np.all(st == 0)
    assert np.array_equal(emf, d["emf"])
    assert np.array_equal(d["stored"], d["emf"])
    assert np.count_nonzero(d["eta"]) > 0.6 * d["eta"].size


def test_emf_block_resistive_term_pins_stencil():
    """A plausible slip in the curl stencil (shifted index, swapped width) must
    change the result: the oracle evaluated with the widths permuted differs
    from the reference vectors on every edge with eta != 0 (so the vectors do
    pin which face and which width enter)."""
    d = np.load(os.path.join(GOLD, "emf_block_vectors.npz"))
    bad_dx = d["dx"][[1, 2, 0]]
    emf, _ = O.oracle_emf_upwind_block(d["edge"], d["emf_face"], d["emf_cell"], d["v_norm"],
                                       [d["bfx"], d["bfy"], d["bfz"]], bad_dx, d["eta"])
    nz = d["eta"] != 0
    assert np.all(emf[nz] != d["emf"][nz])
    assert np.array_equal(emf[~nz], d["emf"][~nz])


with open(os.path.join(GOLD, "trajectories.json")) as _f:
    TRAJ = json.load(_f)

import sys  # noqa: E402
sys.path.insert(0, GOLD)
import make_golden as MG  # noqa: E402


@pytest.mark.parametrize("name", sorted(TRAJ))
def test_trajectory(name):
    rec = TRAJ[name]
    pr, steps = MG.make_problem(name)
    assert sha(pr.W) == rec["input_sha"], "initial-condition generator changed"
    sim = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=pr.mhd, eta=pr.eta)
    sim.set_state(pr.W, pr.bf)
    for s in range(steps):
        dt = sim.step(pr.cfl)
        assert dt == rec["dt"][s], (s, dt, rec["dt"][s])
        assert sha(sim.W0) == rec["W_sha"][s], f"step {s}"
        if pr.mhd:
            assert [sha(b) for b in sim.bf0] == rec["bf_sha"][s]


def test_block_flux_fixture_matches_reference_build():
    """The per-block fixture (SHA-256 of prj_flux_update / prj_flux_div outputs
    on seeded raw blocks) reproduces from the serial reference build."""
    import hashlib
    import json
    if not O.ref_available():
        pytest.skip("serial reference build not present")
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "block_flux.json")))

    def sha(a):
        return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()

    for name, g in gold.items():
        slab, gam, bf = O.block_inputs(g["seed"], g["mhd"])
        eos = O.block_eosvar(slab, g["slot"], gam)
        assert sha(slab) == g["input_sha"], name
        out = O.ref_flux_update_raw(g["mhd"], g["slot"], slab, eos, bf)
        assert {k: [sha(a) for a in v] for k, v in out.items()} == g["out_sha"], name
        fd, _, _ = O.ref_flux_div_raw(g["mhd"], out["flux"])
        assert sha(fd) == g["flux_div_sha"], name


def test_bf2bc_fixture_matches_reference_build():
    """The per-block bf2bc fixture (SHA-256 of the W slab prj_mhd_bf2bc(_all)
    leaves on seeded raw blocks) reproduces from the serial reference build,
    and the done flags do skip cells (the two calls differ)."""
    if not O.ref_available():
        pytest.skip("serial reference build not present")
    gold = json.load(open(os.path.join(GOLD, "bf2bc.json")))
    for name, g in gold.items():
        slab, _, bf = O.block_inputs(g["seed"], True)
        assert sha(slab) == g["input_sha"], name
        done = O.bf2bc_done_flags(g["seed"]) if g["use_done"] else None
        assert sha(O.ref_bf2bc_raw(g["slot"], slab, bf, done)) == g["out_sha"], name
    slab, _, bf = O.block_inputs(33, True)
    assert not np.array_equal(O.ref_bf2bc_raw(1, slab, bf, O.bf2bc_done_flags(33)),
                              O.ref_bf2bc_raw(1, slab, bf))
