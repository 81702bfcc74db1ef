"""Randomised parity sweep: seeded configurations over the options the mesh
path offers -- mesh shapes of 16, 32 or 48 cells per axis (one-block-thick
axes included), per-axis boundaries (periodic where the axis has two blocks
or more, outflow, reflect, mixed),
gamma, RK2 or eSSPRK(10,4), HLLC or HLLE (hydro), hydro or MHD (outflow, the
only field boundary the reference implements; with and without a uniform
resistivity) -- each stepped twice on the GPU and by the oracle: every dt,
cell and face field bit-identical."""
import random

import numpy as np
import pytest

import pyoracle as O
from fornaxamr_amd import problems as P

pytestmark = pytest.mark.gpu


def _configs(count=32, seed=2026):
    rng = random.Random(seed)
    out = []
    for c in range(count):
        mhd = c % 3 == 2
        n = tuple(rng.choice((16, 32, 48)) for _ in range(3))
        if mhd:
            bc = (0,) * 6
        else:
            bc = []
            for d in range(3):
                # a periodic axis needs two blocks (the mesh refuses one, as the
                # reference's ghost fill finds no neighbour for it)
                kind = rng.choice(("periodic", "open")) if n[d] >= 32 else "open"
                if kind == "periodic":
                    bc += [3, 3]
                else:
                    bc += [rng.choice((0, 1)), rng.choice((0, 1))]
            bc = tuple(bc)
        out.append(dict(
            n=n, bc=bc, mhd=mhd, seed=rng.randrange(1 << 20),
            gamma=round(rng.uniform(1.2, 1.8), 6),
            integrator=rng.choice(("rk2", "essprk10_4")),
            riemann="hllc" if mhd else rng.choice(("hllc", "hlle")),
            eta=rng.choice((0.0, 1.0e-3)) if mhd else 0.0))
    return out


CONFIGS = _configs()


@pytest.mark.parametrize("cfg", CONFIGS, ids=[
    "{}{}x{}x{}-bc{}-{}-{}{}".format("mhd" if c["mhd"] else "hydro", *c["n"],
                                     "".join(map(str, c["bc"])), c["integrator"], c["riemann"],
                                     "-eta" if c["eta"] else "") for c in CONFIGS])
def test_random_configuration(gpu, cfg):
    from fornaxamr_amd.mesh import GpuMesh
    pr = P.random_smooth(n=cfg["n"], seed=cfg["seed"], bc=cfg["bc"], gamma=cfg["gamma"],
                         mhd=cfg["mhd"], eta=cfg["eta"])
    gm = GpuMesh.from_problem(pr)
    om = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=pr.mhd, eta=pr.eta, log_cr=True,
                     riemann=cfg["riemann"])
    om.set_state(pr.W, pr.bf)
    gm.set_integrator(cfg["integrator"])
    if not pr.mhd:
        gm.set_riemann(cfg["riemann"])
    for s in range(2):
        if cfg["integrator"] == "essprk10_4":
            dt_o = om.step_essprk10_4(pr.cfl)
        else:
            dt_o = om.step(pr.cfl)
        assert gm.step(pr.cfl) == dt_o, s
        assert np.array_equal(gm.get_prim(0), om.W0), s
        if pr.mhd:
            for a, b in zip(gm.get_bf(0), om.bf0):
                assert np.array_equal(a, b), s
    gm.close()
