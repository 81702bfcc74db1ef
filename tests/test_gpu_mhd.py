"""MHD RK2 stage with constrained transport on the GPU against the oracle.

Bar: bit-identical cell states, face fields and dt; |div B| per cell at
round-off relative to |B|/dx (target < 1e-13)."""
import numpy as np
import pytest

import pyoracle as O
from fornaxamr_amd import problems as P
from helpers import div_b, max_rel, ulp_diff

pytestmark = pytest.mark.gpu


def _pair(pr):
    """GPU mesh and oracle on the same state.  The oracle takes the CT
    resistivity mask's log correctly rounded, as the GPU does (the libm mode
    is pinned to the reference build in tests/test_oracle_reference.py)."""
    from fornaxamr_amd.mesh import GpuMesh
    gm = GpuMesh.from_problem(pr)
    om = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=True, eta=pr.eta, log_cr=True)
    om.set_state(pr.W, pr.bf)
    return gm, om


RESISTIVE = {
    # uniform mhd_eta: eta * curl B on every edge (src/prj_mhd.c:689-711)
    "uniform_eta": lambda: P.random_smooth(n=(32, 32, 32), seed=25, bc=(0,) * 6, mhd=True,
                                           eta=1.0e-3),
    # anomalous mask > 0 on a patch with rho > 1e13, Ye > 0.25 (src/prj_timeint.c:230-290)
    "mask_patch": lambda: P.dense_resistive(n=(32, 32, 32), seed=0),
    "mask_plus_uniform": lambda: P.dense_resistive(n=(32, 48, 32), seed=1, eta=2.0e-3),
}


@pytest.mark.parametrize("case", sorted(RESISTIVE))
def test_resistive_ct_steps(gpu, case):
    """Resistive CT, stage by stage, bit-identical to the oracle: cells, face
    fields and dt over three RK2 steps."""
    pr = RESISTIVE[case]()
    if case != "uniform_eta":
        assert P.eta_mask_cells(pr.W).sum() > 1000
    gm, om = _pair(pr)
    for s in range(3):
        dt = om.calc_dt(pr.cfl)
        assert gm.calc_dt(pr.cfl) == dt, s
        for st in (1, 2):
            gm.stage(st, dt)
            om.stage(st, dt)
            gm.sync()
            slot = 1 if st == 1 else 0
            want_W = om.W1 if st == 1 else om.W0
            want_bf = om.bf1 if st == 1 else om.bf0
            assert ulp_diff(gm.get_prim(slot), want_W) == 0, (s, st)
            for a, b in zip(gm.get_bf(slot), want_bf):
                assert np.array_equal(a, b), (s, st)


@pytest.mark.parametrize("ct_split", [1, 2, 4])
@pytest.mark.parametrize("side", [0, 1])
def test_schedules_match_oracle(gpu, ct_split, side):
    """Every CT split (one block's planes over 1, 2 or 4 workgroups, each with
    its own prologue) and the ghost-layer face kernel in order or on the side
    stream: cells, faces and dt bit-identical to the oracle over two steps,
    stage by stage, on a resistive mesh (the edges' curl B reads faces the CT
    updates)."""
    pr = P.random_smooth(n=(32, 48, 32), seed=44, bc=(0,) * 6, mhd=True, eta=1.0e-3)
    gm, om = _pair(pr)
    gm.set_mhd_schedule(ct_split, side)
    for s in range(2):
        dt = om.calc_dt(pr.cfl)
        assert gm.calc_dt(pr.cfl) == dt, s
        for st in (1, 2):
            gm.stage(st, dt)
            om.stage(st, dt)
            gm.sync()
            slot = 1 if st == 1 else 0
            assert ulp_diff(gm.get_prim(slot), om.W1 if st == 1 else om.W0) == 0, (s, st)
            for a, b in zip(gm.get_bf(slot), om.bf1 if st == 1 else om.bf0):
                assert np.array_equal(a, b), (s, st)
    # the step path with recorded CFL maxima per CT workgroup of this split
    for s in range(2):
        assert gm.step(pr.cfl) == om.step(pr.cfl), s
    assert ulp_diff(gm.get_prim(0), om.W0) == 0
    gm.close()


def test_large_mesh_default_schedule(gpu):
    """144^3 (729 blocks: more than fit the chip in one round, so the default
    schedule is CT split 1 with the ghost-layer faces in order, as at 256^3):
    two steps bit-identical to the oracle."""
    pr = P.magnetized_sedov(n=(144, 144, 144))
    gm, om = _pair(pr)
    for s in range(2):
        assert gm.step(pr.cfl) == om.step(pr.cfl), s
    assert ulp_diff(gm.get_prim(0), om.W0) == 0
    for a, b in zip(gm.get_bf(0), om.bf0):
        assert np.array_equal(a, b)
    gm.close()


def test_bench_size_256_steps_match_oracle(gpu):
    """The bench's MHD workload at its full size (256^3 magnetized Sedov, 4096
    blocks; default schedule: CT split 1, ghost-layer faces in order, their
    stencils on the ghost planes z = -1 / nz from the compact copies): two RK2
    steps, dt, cells and face fields bit-identical to the oracle."""
    pr = P.magnetized_sedov(n=(256, 256, 256))
    gm, om = _pair(pr)
    for s in range(2):
        assert gm.step(pr.cfl) == om.step(pr.cfl), s
    assert ulp_diff(gm.get_prim(0), om.W0) == 0
    for a, b in zip(gm.get_bf(0), om.bf0):
        assert np.array_equal(a, b)
    gm.close()


def test_schedule_arguments(gpu):
    from fornaxamr_amd import _native
    pr = P.random_smooth(n=(32, 32, 32), seed=45, bc=(0,) * 6, mhd=True)
    gm, _ = _pair(pr)
    for bad in ((3, 0), (8, 0), (1, 2)):
        with pytest.raises(_native.FxError):
            gm.set_mhd_schedule(*bad)
    gm.close()


def test_single_stage(gpu):
    pr = P.random_smooth(n=(32, 32, 32), seed=41, bc=(0,) * 6, mhd=True)
    gm, om = _pair(pr)
    dt = om.calc_dt(pr.cfl)
    assert gm.calc_dt(pr.cfl) == dt
    gm.stage(1, dt)
    om.stage(1, dt)
    assert max_rel(gm.get_prim(1), om.W1) <= 1e-12
    assert ulp_diff(gm.get_prim(1), om.W1) == 0
    for a, b in zip(gm.get_bf(1), om.bf1):
        assert np.array_equal(a, b)
    gm.stage(2, dt)
    om.stage(2, dt)
    assert ulp_diff(gm.get_prim(0), om.W0) == 0
    for a, b in zip(gm.get_bf(0), om.bf0):
        assert np.array_equal(a, b)


def test_magnetized_sedov_steps(gpu):
    pr = P.magnetized_sedov(n=(32, 48, 32))
    gm, om = _pair(pr)
    for s in range(5):
        assert gm.step(pr.cfl) == om.step(pr.cfl), s
    assert ulp_diff(gm.get_prim(0), om.W0) == 0
    for a, b in zip(gm.get_bf(0), om.bf0):
        assert np.array_equal(a, b)


def test_recorded_cfl_maximum(gpu):
    """The CT kernel of stage 2 records the next step's CFL maximum per
    workgroup (a block's planes split over four workgroups on this small
    mesh).  It must equal the reduction over the cells (forced by taking a
    view) and the oracle's, and a state set between steps must drop it."""
    pr = P.magnetized_sedov(n=(32, 48, 32))
    gm, om = _pair(pr)
    for s in range(3):
        assert gm.step(pr.cfl) == om.step(pr.cfl), s
    recorded = gm.calc_dt(pr.cfl)
    gm.view()
    assert recorded == gm.calc_dt(pr.cfl) == om.calc_dt(pr.cfl)
    W2 = om.W0.copy()
    W2[1:4] *= 2.0
    bf = [b.copy() for b in om.bf0]
    gm.set_state(W2, bf)
    om.set_state(W2, bf)
    for s in range(2):
        assert gm.step(pr.cfl) == om.step(pr.cfl), s
    assert ulp_diff(gm.get_prim(0), om.W0) == 0


def test_divb_roundoff_128(gpu):
    """C3: magnetized Sedov at 128^3 (512 blocks), 50 RK2 steps:
    |div B| dx / |B| < 1e-13 in every cell."""
    pr = P.magnetized_sedov(n=(128, 128, 128))
    from fornaxamr_amd.mesh import GpuMesh
    gm = GpuMesh.from_problem(pr)
    for _ in range(50):
        gm.step_async(pr.cfl)
    gm.sync()
    bf = gm.get_bf(0)
    dx = [(pr.xlim[2 * d + 1] - pr.xlim[2 * d]) / pr.n[d] for d in range(3)]
    W = gm.get_prim(0)
    bmag = np.sqrt(W[6] ** 2 + W[7] ** 2 + W[8] ** 2)
    rel = np.abs(div_b(bf, dx)) * dx[0] / np.maximum(bmag, 1e-300)
    assert float(rel.max()) < 1e-13


def test_error_word_reports_bad_state(gpu):
    """A negative density in the state surfaces as the reference's HLLD abort
    message through the C ABI, not as a silent NaN."""
    from fornaxamr_amd import _native
    pr = P.random_smooth(n=(32, 32, 32), seed=42, bc=(0,) * 6, mhd=True)
    pr.W[0, 5, 5, 5] = -1.0
    from fornaxamr_amd.mesh import GpuMesh
    gm = GpuMesh.from_problem(pr)
    with pytest.raises(_native.FxError, match="non-positive (density|gas pressure)"):
        gm.stage(1, 1e-4)
        gm.sync()
