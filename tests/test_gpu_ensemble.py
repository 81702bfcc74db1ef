"""Ragged ensembles of hydro meshes on the GPU.

Every schedule of the batched step (per-instance launches, one batched launch
per stage with and without the XCD remap and the interleaved table, persistent
workgroups on atomic queues, hipGraph capture) must give, for every instance,
bit for bit the states and dt sequence of the same problem advanced alone by
the single-mesh path, which the other GPU tests pin to the oracle; one
instance is also compared with the oracle directly.
"""
import numpy as np
import pytest

import pyoracle as O
from fornaxamr_amd import _native
from fornaxamr_amd import problems as P
from helpers import ulp_diff

pytestmark = pytest.mark.gpu
STEPS = 4
CFL = 0.4
FACTOR = 1.2


def _instances():
    return [
        P.random_smooth(n=(32, 32, 32), seed=41, bc=(3,) * 6, gamma=1.4),
        P.random_smooth(n=(48, 32, 64), seed=42, bc=(3, 3, 0, 0, 1, 1), gamma=1.6),
        P.turbulent_box(n=(64, 64, 64), seed=43, gamma=1.35),
        P.sedov(n=(32, 32, 32)),
        P.random_smooth(n=(16, 32, 48), seed=44, bc=(0, 0, 3, 3, 1, 0), gamma=1.5),
    ]


@pytest.fixture(scope="module")
def expected(gpu):
    """Per instance: dt of each step and the state after each step, from the
    single-mesh path."""
    from fornaxamr_amd.mesh import GpuMesh
    out = []
    for pr in _instances():
        m = GpuMesh.from_problem(pr)
        dts, Ws = [], []
        for _ in range(STEPS):
            dts.append(m.step(CFL, FACTOR))
            Ws.append(m.get_prim(0))
        m.close()
        out.append((dts, Ws))
    return out


def test_single_mesh_reference_matches_oracle(expected):
    pr = _instances()[1]
    om = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma)
    om.set_state(pr.W)
    dts, Ws = expected[1]
    for s in range(STEPS):
        assert om.step(CFL, FACTOR) == dts[s]
    assert ulp_diff(om.W0, Ws[-1]) == 0


SCHEDULES = [
    (0, 0, 1, 0),       # per-instance launches
    (1, 0, 1, 0),       # batched, XCD remap, instance-major table
    (1, 0, 0, 0),
    (1, 0, 1, 1),       # interleaved table
    (1, 7, 1, 0),       # persistent, few workgroups, per-XCD queues
    (1, 7, 0, 1),       # persistent, one queue
    (1, 256, 1, 1),
    (1, 600, 0, 0),     # more workgroups than table entries
]


@pytest.mark.parametrize("sched", SCHEDULES)
def test_schedule_matches_single_mesh(expected, sched):
    from fornaxamr_amd.ensemble import GpuEnsemble
    e = GpuEnsemble(_instances())
    e.set_schedule(*sched)
    for s in range(STEPS):
        e.step_async(CFL, FACTOR)
        dts = e.last_dt()
        for i, (edts, _) in enumerate(expected):
            assert dts[i] == edts[s], (sched, i, s)
    e.sync()
    for i, (_, Ws) in enumerate(expected):
        assert np.array_equal(e.get_prim(i), Ws[-1]), (sched, i)
    e.close()


@pytest.mark.parametrize("sched", [(1, 0, 1, 0), (1, 256, 1, 1), (0, 0, 1, 0)])
def test_graph_replays(expected, sched):
    """A captured step replayed STEPS times, with eager work (host copies of
    the state, an unrelated single-mesh step) between the replays; every
    replay is checked."""
    from fornaxamr_amd.ensemble import GpuEnsemble
    from fornaxamr_amd.mesh import GpuMesh
    e = GpuEnsemble(_instances())
    e.set_schedule(*sched)
    e.set_graph(True)
    other = GpuMesh.from_problem(P.random_smooth(n=(32, 32, 32), seed=9, bc=(3,) * 6))
    for s in range(STEPS):
        e.step_async(CFL, FACTOR)
        e.sync()
        dts = e.last_dt()
        for i, (edts, Ws) in enumerate(expected):
            assert dts[i] == edts[s], (sched, i, s)
            assert np.array_equal(e.get_prim(i), Ws[s]), (sched, i, s)
        other.step(0.4)
    e.close()
    other.close()


def test_graph_recaptures_on_new_cfl(expected):
    from fornaxamr_amd.ensemble import GpuEnsemble
    e = GpuEnsemble(_instances())
    e.set_schedule(1, 0, 1, 0)
    e.set_graph(True)
    e.step_async(0.2, FACTOR)       # captured with cfl 0.2
    e.reset()
    for s in range(STEPS):
        e.step_async(CFL, FACTOR)   # must not replay the cfl 0.2 graph
    e.sync()
    for i, (_, Ws) in enumerate(expected):
        assert np.array_equal(e.get_prim(i), Ws[-1]), i
    e.close()


@pytest.mark.parametrize("sched", [(0, 0, 1, 0), (1, 0, 1, 0)])
def test_state_set_between_steps(expected, sched):
    """Each step takes the CFL maxima its predecessor's stage 2 recorded; a
    state set between steps (here: back to the initial one, then one
    instance changed) must be reduced afresh."""
    from fornaxamr_amd.ensemble import GpuEnsemble
    e = GpuEnsemble(_instances())
    e.set_schedule(*sched)
    for _ in range(2):
        e.step_async(CFL, FACTOR)
    e.reset()
    for s in range(STEPS):
        e.step_async(CFL, FACTOR)
        dts = e.last_dt()
        for i, (edts, _) in enumerate(expected):
            assert dts[i] == edts[s], (sched, i, s)
    pr = _instances()[3]
    W2 = pr.W.copy()
    W2[1:4] += 0.5
    e.set_state(3, W2)
    om = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma)
    om.set_state(W2)
    e.step_async(CFL, FACTOR)
    assert e.last_dt()[3] == om.step(CFL, FACTOR)
    e.sync()
    assert ulp_diff(e.get_prim(3), om.W0) == 0
    e.close()


def test_autotune_then_step(expected):
    from fornaxamr_amd.ensemble import GpuEnsemble
    e = GpuEnsemble(_instances())
    e.set_graph(True)
    best = e.autotune(CFL, FACTOR, reps=2)
    assert best["mode"] == 1 and best["ms_per_step"] > 0.0
    for _ in range(STEPS):
        e.step_async(CFL, FACTOR)
    e.sync()
    for i, (_, Ws) in enumerate(expected):
        assert np.array_equal(e.get_prim(i), Ws[-1]), i
    e.close()


def test_bad_arguments(gpu):
    from fornaxamr_amd.ensemble import GpuEnsemble
    with pytest.raises(ValueError):
        GpuEnsemble([P.random_smooth(n=(32, 32, 32), seed=1, mhd=True)])
    e = GpuEnsemble(_instances()[:1])
    with pytest.raises(_native.FxError):
        e.set_schedule(2, 0, 0, 0)
    assert e.lib.fx_ensemble_get_prim(e.h, 1, 0, None) == -1      # FX_EINVAL
    assert e.lib.fx_ensemble_set_schedule(e.h, 1, -1, 0, 0) == -1
    e.close()


def test_set_state_checks_shape_and_is_kept_by_reset(gpu):
    """set_state refuses a state of the wrong shape (it would read past the
    host array), and reset()/autotune() return an instance to the state
    set_state gave it, not to the problem's initial one."""
    from fornaxamr_amd.ensemble import GpuEnsemble
    probs = _instances()[:2]
    e = GpuEnsemble(probs)
    with pytest.raises(ValueError):
        e.set_state(0, np.zeros((6, 16, 32, 32)))
    with pytest.raises(ValueError):
        e.set_state(1, probs[0].W)
    with pytest.raises(IndexError):
        e.set_state(2, probs[0].W)
    W2 = probs[1].W.copy()
    W2[1:4] *= 0.5
    e.set_state(1, W2)
    e.step_async(CFL, FACTOR)
    e.sync()
    e.reset()
    assert np.array_equal(e.get_prim(1, 0), W2)
    assert np.array_equal(e.get_prim(0, 0), probs[0].W)
    e.set_graph(True)
    e.autotune(CFL, FACTOR, reps=1)
    assert np.array_equal(e.get_prim(1, 0), W2)
    e.close()


def test_full_ragged_ensemble_48(gpu):
    """C4 at its full size: the 48-member draw (6249 blocks, members up to
    128^3), autotuned schedule, two graph-replayed steps.  Every instance must
    be bit-identical (state and dt) to its own single-mesh run, and conserve
    mass, momentum, energy and rho*Ye in its periodic box to round-off."""
    from fornaxamr_amd.ensemble import GpuEnsemble
    from fornaxamr_amd.mesh import GpuMesh
    from helpers import cons_totals
    probs = P.ragged_ensemble(48)
    e = GpuEnsemble(probs)
    assert e.nblocks == sum(p.nblocks for p in probs)
    assert max(p.n[0] for p in probs) == 128
    e.set_graph(True)
    best = e.autotune(CFL, FACTOR, reps=1)
    print("autotuned schedule", best, flush=True)
    dts = []
    for _ in range(2):
        e.step_async(CFL, FACTOR)
        dts.append(e.last_dt())
    e.sync()
    for i, pr in enumerate(probs):
        m = GpuMesh.from_problem(pr)
        for s in range(2):
            assert m.step(CFL, FACTOR) == dts[s][i], (i, s)
        W = e.get_prim(i)
        assert np.array_equal(W, m.get_prim(0)), i
        m.close()
        vol = 1.0 / pr.ncells
        t0, t1 = cons_totals(pr.W, vol), cons_totals(W, vol)
        scale = np.abs(t0) + 1e-300
        scale[1:4] = max(scale[0], float(np.sum(np.abs(pr.W[0] * pr.W[1:4]))) * vol)
        assert np.all(np.abs(t1 - t0) <= 1e-12 * scale), (i, t1 - t0)
    e.close()
