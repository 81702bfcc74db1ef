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
