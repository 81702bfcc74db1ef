"""eSSPRK(10,4) on the GPU (the reference's default integrator) against the
oracle, which tests/test_oracle_reference.py pins bit-exact to the reference
built with TIME_INTEGRATION=eSSPRK10_4: identical dt sequences and states."""
import numpy as np
import pytest

import pyoracle as O
from fornaxamr_amd import problems as P
from helpers import cons_totals

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bc", [(3,) * 6, (0,) * 6, (3, 3, 0, 1, 1, 0)])
def test_essprk_steps_match_oracle(gpu, bc):
    from fornaxamr_amd.mesh import GpuMesh
    pr = P.random_smooth(n=(32, 48, 32), seed=72, bc=bc, gamma=1.4)
    gm = GpuMesh.from_problem(pr)
    gm.set_integrator("essprk10_4")
    om = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma)
    om.set_state(pr.W)
    for s in range(4):
        assert gm.step(pr.cfl) == om.step_essprk10_4(pr.cfl), s
        assert np.array_equal(gm.get_prim(0), om.W0), s
    gm.close()


def test_essprk_sedov_and_conservation(gpu):
    from fornaxamr_amd.mesh import GpuMesh
    pr = P.sedov(n=(64, 64, 64))
    gm = GpuMesh.from_problem(pr)
    gm.set_integrator("essprk10_4")
    om = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma)
    om.set_state(pr.W)
    for s in range(5):
        assert gm.step(pr.cfl) == om.step_essprk10_4(pr.cfl), s
    assert np.array_equal(gm.get_prim(0), om.W0)
    gm.close()
    tb = P.turbulent_box(n=(64, 64, 64), seed=9)
    gm = GpuMesh.from_problem(tb)
    gm.set_integrator("essprk10_4")
    vol = (1.0 / 64) ** 3
    t0 = cons_totals(tb.W, vol)
    for _ in range(3):
        gm.step_async(tb.cfl)
    gm.sync()
    t1 = cons_totals(gm.get_prim(0), vol)
    scale = np.array([t0[0], t0[0], t0[0], t0[0], t0[4], t0[5]])
    assert np.all(np.abs(t1 - t0) / np.abs(scale) < 1e-12)
    gm.close()


def test_essprk_rejects_mhd(gpu):
    from fornaxamr_amd import _native
    from fornaxamr_amd.mesh import GpuMesh
    pr = P.random_smooth(n=(32, 32, 32), seed=3, bc=(0,) * 6, mhd=True)
    gm = GpuMesh.from_problem(pr)
    with pytest.raises(_native.FxError):
        gm.set_integrator("essprk10_4")
    gm.close()
