"""The reference with the reference-side shims linked to the GPU library.

oracle/_ref/libprj_ref_{hydro,mhd}[_essprk10_4]_fxflux.so are the serial
reference build (RK2 or the default eSSPRK(10,4)) with two shims of
fornaxamr_amd/shim/ compiled in:

  * prj_flux_fx.c is the build's prj_flux_update: every block of every stage
    gets its fluxes, face velocities and (MHD) face EMF inputs from
    fx_flux_update_block, while everything around the call --
    prj_timeint_step and its stages, prj_flux_div, the CT update and
    resistive EMFs, the ghost fill, the EOS, the CFL dt -- is the
    reference's own code;
  * prj_timeint_fx.c takes the reference's mesh onto the GPU (fx_mesh_create
    from the mesh's geometry, boundaries and gamma, the blocks' slot-0 state
    uploaded), steps it there with fx_step, and writes the state back into
    the blocks with the reference's own refill of eosvar and ghosts.

oracle/_ref/libprj_ref_{hydro,mhd}_fxface.so carry the third,
prj_riemann_fx.c: prj_riemann_hllc (hydro), or prj_riemann_hlld,
prj_mhd_emf_upwind and the prj_mhd_bf2bc forms (MHD), with the reference's
signatures, one GPU call per face, edge or block, inside the reference's own
prj_flux_update, CT edge loop and ghost fill.

Over several steps the state and the dt sequence must equal those of the
unmodified reference build bit for bit.  Wall times per step are printed.
"""
import time

import numpy as np
import pytest

import pyoracle as O
from fornaxamr_amd import problems as P

pytestmark = pytest.mark.gpu

CASES = {
    "hydro_periodic": (lambda: P.random_smooth(n=(32, 32, 32), seed=81, bc=(3,) * 6, gamma=1.4), 3),
    "hydro_mixed": (lambda: P.random_smooth(n=(32, 48, 32), seed=82, bc=(3, 3, 0, 1, 1, 0)), 3),
    "hydro_sedov": (lambda: P.sedov(n=(32, 32, 32)), 4),
    "mhd_outflow": (lambda: P.random_smooth(n=(32, 32, 32), seed=83, bc=(0,) * 6, mhd=True), 3),
    "mhd_resistive": (lambda: P.dense_resistive(n=(32, 32, 32), seed=84, eta=1.0e-3), 2),
}


def _run(pr, key, steps, t_end=1.0e300):
    sim = O.RefSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=pr.mhd, eta=pr.eta, key=key)
    sim.set_state(pr.W, pr.bf)
    dts, states = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        dts.append(sim.step(pr.cfl, t_end=t_end))
        st = [sim.get_prim(0)]
        if pr.mhd:
            st += sim.get_bf(0)
        states.append(st)
    wall = (time.perf_counter() - t0) / steps
    sim.close()
    return dts, states, wall


def _need(*keys):
    if not all(O.ref_available(k) for k in keys):
        pytest.skip("reference builds (oracle/_ref) not present")


@pytest.mark.parametrize("case", sorted(CASES))
def test_reference_stage_with_gpu_flux(gpu, case):
    make, steps = CASES[case]
    pr = make()
    base = "mhd" if pr.mhd else "hydro"
    _need(base, base + "_fxflux")
    dts_ref, st_ref, wall_ref = _run(pr, base, steps)
    dts_fx, st_fx, wall_fx = _run(pr, base + "_fxflux", steps)
    assert dts_fx == dts_ref
    for s in range(steps):
        for a, b in zip(st_fx[s], st_ref[s]):
            assert np.array_equal(a, b), f"{case}: state differs after step {s}"
    nblocks = int(np.prod(pr.n)) // 4096
    print(f"\n[{case}] {nblocks} blocks, {steps} steps: reference {wall_ref * 1e3:.1f} ms/step, "
          f"reference with GPU prj_flux_update {wall_fx * 1e3:.1f} ms/step")


ESSPRK_CASES = {
    "hydro_periodic": (lambda: P.random_smooth(n=(32, 32, 32), seed=85, bc=(3,) * 6, gamma=1.4), 2),
    "mhd_outflow": (lambda: P.random_smooth(n=(32, 32, 32), seed=86, bc=(0,) * 6, mhd=True), 2),
}


@pytest.mark.parametrize("case", sorted(ESSPRK_CASES))
def test_reference_essprk10_4_with_gpu_flux(gpu, case):
    """The per-block shim under the reference's default integrator: ten
    prj_flux_update calls per block per step, with the integrator's slot
    blends between them."""
    make, steps = ESSPRK_CASES[case]
    pr = make()
    base = ("mhd" if pr.mhd else "hydro") + "_essprk10_4"
    _need(base, base + "_fxflux")
    dts_ref, st_ref, _ = _run(pr, base, steps)
    dts_fx, st_fx, _ = _run(pr, base + "_fxflux", steps)
    assert dts_fx == dts_ref
    for s in range(steps):
        for a, b in zip(st_fx[s], st_ref[s]):
            assert np.array_equal(a, b), f"{case}: state differs after step {s}"


MESH_CASES = {
    "hydro_rk2_periodic": ("hydro", lambda: P.random_smooth(n=(32, 48, 32), seed=87, bc=(3,) * 6,
                                                            gamma=1.4), 4),
    "hydro_rk2_sedov": ("hydro", lambda: P.sedov(n=(32, 32, 32)), 4),
    "hydro_essprk_mixed": ("hydro_essprk10_4",
                           lambda: P.random_smooth(n=(32, 32, 48), seed=88, bc=(3, 3, 0, 1, 1, 0)), 2),
    "mhd_rk2_outflow": ("mhd", lambda: P.random_smooth(n=(32, 32, 32), seed=89, bc=(0,) * 6,
                                                      mhd=True), 3),
    "mhd_rk2_resistive": ("mhd", lambda: P.dense_resistive(n=(32, 32, 32), seed=90, eta=1.0e-3), 2),
    # the reference implements outflow magnetic-field boundaries only
    # (prj_boundary_apply_bf_axis), so every MHD case is outflow
    "mhd_essprk_outflow": ("mhd_essprk10_4",
                           lambda: P.random_smooth(n=(32, 48, 32), seed=91, bc=(0,) * 6, mhd=True), 2),
}


@pytest.mark.parametrize("case", sorted(MESH_CASES))
def test_reference_mesh_stepped_on_gpu(gpu, case):
    """prj_fx_attach / prj_fx_calc_dt / prj_fx_step / prj_fx_download on the
    reference's own mesh, in the reference's main loop (growth limit, dt_step
    clipped to t_end): the dt of every step and the state written back into
    the blocks equal the unmodified reference's, including after a download
    in the middle of the run (the device state carries on unchanged).  t_end
    falls inside the last step, so that step is a clipped one."""
    base, make, steps = MESH_CASES[case]
    pr = make()
    _need(base, base + "_fxflux")
    dts_free, _, _ = _run(pr, base, steps)
    t_end = sum(dts_free[:-1]) + 0.5 * dts_free[-1]
    dts_ref, st_ref, wall_ref = _run(pr, base, steps, t_end)
    assert dts_ref[:-1] == dts_free[:-1] and dts_ref[-1] < dts_free[-1]

    sim = O.RefSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=pr.mhd, eta=pr.eta, key=base + "_fxflux")
    sim.set_state(pr.W, pr.bf)
    sim.fx_attach()
    dts = []
    t_step = t_down = 0.0
    for s in range(steps):
        t0 = time.perf_counter()
        dts.append(sim.fx_step(pr.cfl, t_end=t_end))
        t1 = time.perf_counter()
        t_step += t1 - t0
        if s == steps // 2 - 1 or s == steps - 1:
            sim.fx_download()
            t_down += time.perf_counter() - t1
            got = [sim.get_prim(0)] + (sim.get_bf(0) if pr.mhd else [])
            for a, b in zip(got, st_ref[s]):
                assert np.array_equal(a, b), f"{case}: state differs after step {s}"
    sim.fx_detach()
    sim.close()
    assert dts == dts_ref
    nblocks = int(np.prod(pr.n)) // 4096
    print(f"\n[{case}] {nblocks} blocks, {steps} steps: reference {wall_ref * 1e3:.1f} ms/step; "
          f"mesh on the GPU {t_step / steps * 1e3:.2f} ms/step (first-launch costs included), "
          f"download + the reference's refill {t_down / 2 * 1e3:.1f} ms each")


FACE_CASES = {
    # (the reference's ghost fill needs two blocks along a periodic axis)
    "hydro_periodic": (lambda: P.random_smooth(n=(32, 32, 16), seed=92, bc=(3, 3, 3, 3, 0, 0),
                                               gamma=1.4), 1),
    "hydro_mixed": (lambda: P.random_smooth(n=(32, 16, 16), seed=93, bc=(3, 3, 0, 1, 1, 0)), 1),
    "mhd_outflow": (lambda: P.random_smooth(n=(16, 16, 32), seed=94, bc=(0,) * 6, mhd=True), 1),
    "mhd_resistive": (lambda: P.dense_resistive(n=(16, 16, 16), seed=95, eta=1.0e-3), 1),
}


@pytest.mark.parametrize("case", sorted(FACE_CASES))
def test_reference_with_gpu_face_edge_functions(gpu, case):
    """The same-signature per-face / per-edge / per-block shim: the
    reference's own loops call prj_riemann_hllc or prj_riemann_hlld once per
    face, prj_mhd_emf_upwind once per edge (with the resistive term on the
    resistive mesh) and prj_mhd_bf2bc_all_mesh once per block of each ghost
    fill, every call on the GPU; dts and states equal the unmodified
    reference's bit for bit."""
    make, steps = FACE_CASES[case]
    pr = make()
    base = "mhd" if pr.mhd else "hydro"
    _need(base, base + "_fxface")
    dts_ref, st_ref, wall_ref = _run(pr, base, steps)
    dts_fx, st_fx, wall_fx = _run(pr, base + "_fxface", steps)
    assert dts_fx == dts_ref
    for s in range(steps):
        for a, b in zip(st_fx[s], st_ref[s]):
            assert np.array_equal(a, b), f"{case}: state differs after step {s}"
    nblocks = int(np.prod(pr.n)) // 4096
    print(f"\n[{case}] {nblocks} blocks, {steps} steps: reference {wall_ref * 1e3:.1f} ms/step, "
          f"with one GPU call per face / edge {wall_fx * 1e3:.0f} ms/step")
