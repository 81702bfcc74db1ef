"""The C restatement (oracle/fx_oracle.c) against the serial reference build.

The reference build (oracle/_ref) exists only where the reference sources are
(the development container); these tests skip elsewhere.  Everything is
compared bit for bit.
"""
import numpy as np
import pytest

import pyoracle as O
from fornaxamr_amd import problems as P

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def _run_both(pr, steps):
    ref = O.RefSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=pr.mhd, eta=pr.eta)
    ora = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=pr.mhd, eta=pr.eta)
    assert np.array_equal(ref.dx, ora.dx)
    ref.set_state(pr.W, pr.bf)
    ora.set_state(pr.W, pr.bf)
    for s in range(steps):
        dtr = ref.step(pr.cfl)
        dto = ora.step(pr.cfl)
        assert dtr == dto, (s, dtr, dto)
        assert np.array_equal(ref.get_prim(0), ora.W0), f"W differs after step {s}"
        if pr.mhd:
            for a, b in zip(ref.get_bf(0), ora.bf0):
                assert np.array_equal(a, b), f"face field differs after step {s}"
    ref.close()


@pytest.mark.parametrize("bc", [(3,) * 6, (0,) * 6, (1,) * 6, (3, 3, 0, 1, 1, 0)])
def test_hydro_boundaries(bc):
    _run_both(P.random_smooth(n=(32, 32, 32), seed=21, bc=bc, gamma=1.4), 3)


def test_hydro_anisotropic_blocks():
    _run_both(P.random_smooth(n=(48, 32, 64), seed=22, bc=(3, 3, 0, 0, 1, 1)), 2)


def test_sedov_c1_stages():
    """C1 scheme on the 4x4x4-block Sedov mesh, every stage compared."""
    pr = P.sedov(n=(64, 64, 64))
    ref = O.RefSim(pr.n, pr.xlim, pr.bc, pr.gamma)
    ora = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma)
    ref.set_state(pr.W)
    ora.set_state(pr.W)
    for s in range(3):
        dt = ref.calc_dt(pr.cfl)
        assert dt == ora.calc_dt(pr.cfl)
        for st in (1, 2):
            ref.stage(st, dt)
            ora.stage(st, dt)
            slot = 1 if st == 1 else 0
            assert np.array_equal(ref.get_prim(slot), ora.W1 if st == 1 else ora.W0)
    ref.close()


def test_mhd_random_outflow():
    _run_both(P.random_smooth(n=(32, 32, 32), seed=23, bc=(0,) * 6, mhd=True), 3)


def test_mhd_sedov():
    _run_both(P.magnetized_sedov(n=(32, 32, 48)), 3)


def test_mhd_emf_edges():
    pr = P.random_smooth(n=(32, 32, 32), seed=24, bc=(0,) * 6, mhd=True)
    ref = O.RefSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=True)
    ora = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=True)
    ref.set_state(pr.W, pr.bf)
    ora.set_state(pr.W, pr.bf)
    dt = ref.calc_dt(pr.cfl)
    ref.stage(1, dt)
    emf = [np.empty(s) for s in O.edge_shapes(pr.n)]
    ora.stage(1, dt, emf_out=emf)
    for a, b in zip(ref.get_emf(), emf):
        assert np.array_equal(a, b)
    ref.close()


@pytest.mark.parametrize("bc", [(3,) * 6, (0,) * 6, (3, 3, 0, 1, 1, 0)])
def test_essprk10_4_matches_reference_build(bc):
    """eSSPRK(10,4) steps (CFL dt x 6, ten substeps of dt/6, three convex
    blends) of the oracle against the reference built with
    TIME_INTEGRATION=eSSPRK10_4: bit-identical states and dt sequence."""
    if not O.ref_available("hydro_essprk10_4"):
        pytest.skip("serial reference build (eSSPRK10_4) not present")
    pr = P.random_smooth(n=(32, 32, 32), seed=71, bc=bc, gamma=1.4)
    ref = O.RefSim(pr.n, pr.xlim, pr.bc, pr.gamma, key="hydro_essprk10_4")
    ref.set_state(pr.W)
    om = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma)
    om.set_state(pr.W)
    for s in range(3):
        assert ref.step(pr.cfl) == om.step_essprk10_4(pr.cfl), s
        assert np.array_equal(ref.get_prim(0), om.W0), s


@pytest.mark.parametrize("case", ["random", "sedov", "resistive"])
def test_essprk10_4_mhd_matches_reference_build(case):
    """eSSPRK(10,4) on MHD meshes: the face fields are saved with the cells
    and blended before them, the cell field of a blend taken from the blended
    faces (src/prj_timeint.c:2287-2355, 2382-2417); oracle against the
    reference built with PRJ_MHD=1 and TIME_INTEGRATION=eSSPRK10_4."""
    if not O.ref_available("mhd_essprk10_4"):
        pytest.skip("serial reference build (MHD eSSPRK10_4) not present")
    if case == "random":
        pr = P.random_smooth(n=(32, 32, 32), seed=72, bc=(0,) * 6, mhd=True)
    elif case == "sedov":
        pr = P.magnetized_sedov(n=(32, 32, 48))
    else:
        pr = P.dense_resistive(n=(32, 32, 32), seed=73, eta=1.0e-3)
    ref = O.RefSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=True, eta=pr.eta, key="mhd_essprk10_4")
    ref.set_state(pr.W, pr.bf)
    om = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=True, eta=pr.eta)
    om.set_state(pr.W, pr.bf)
    for s in range(2):
        assert ref.step(pr.cfl) == om.step_essprk10_4(pr.cfl), s
        assert np.array_equal(ref.get_prim(0), om.W0), s
        for a, b in zip(ref.get_bf(0), om.bf0):
            assert np.array_equal(a, b), s
    ref.close()


# ---- resistive CT (src/prj_mhd.c:681-711, src/prj_timeint.c:228-290) -------

def _stage_differs(pr, pr0):
    """One stage of the oracle on pr and on pr0 (same state, no resistivity):
    the face fields must differ, or a resistive test would pin nothing."""
    out = []
    for p in (pr, pr0):
        o = O.OracleSim(p.n, p.xlim, p.bc, p.gamma, mhd=True, eta=p.eta)
        o.set_state(p.W, p.bf)
        o.stage(1, o.calc_dt(p.cfl))
        out.append([b.copy() for b in o.bf1])
    return max(float(np.max(np.abs(a - b))) for a, b in zip(*out))


def test_mhd_uniform_eta():
    """mhd_eta = 1e-3 everywhere: eta * curl B on every edge."""
    pr = P.random_smooth(n=(32, 32, 32), seed=25, bc=(0,) * 6, mhd=True, eta=1.0e-3)
    pr0 = P.random_smooth(n=(32, 32, 32), seed=25, bc=(0,) * 6, mhd=True)
    assert _stage_differs(pr, pr0) > 1e-9
    _run_both(pr, 3)


def test_mhd_eta_mask_patch():
    """Anomalous resistivity: rho > 1e13 and Ye > 0.25 on a patch of the box,
    so the per-cell mask 1e12 log(rho/1e13) log(Ye/0.25) is positive there."""
    pr = P.dense_resistive(n=(32, 32, 32), seed=0)
    assert P.eta_mask_cells(pr.W).sum() > 1000
    pr0 = P.dense_resistive(n=(32, 32, 32), seed=0)
    pr0.W[5] = 0.1           # no mask: Ye below the threshold everywhere
    assert _stage_differs(pr, pr0) > 0.0
    _run_both(pr, 3)


def test_mhd_eta_mask_plus_uniform():
    _run_both(P.dense_resistive(n=(32, 32, 48), seed=1, eta=2.0e-3), 2)


def test_mhd_emf_edges_resistive():
    pr = P.dense_resistive(n=(32, 32, 32), seed=2, eta=1.0e-3)
    ref = O.RefSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=True, eta=pr.eta)
    ora = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=True, eta=pr.eta)
    ref.set_state(pr.W, pr.bf)
    ora.set_state(pr.W, pr.bf)
    dt = ref.calc_dt(pr.cfl)
    ref.stage(1, dt)
    emf = [np.empty(s) for s in O.edge_shapes(pr.n)]
    ora.stage(1, dt, emf_out=emf)
    for a, b in zip(ref.get_emf(), emf):
        assert np.array_equal(a, b)
    ref.close()
