"""Throughput of the FP64 RK2 Godunov stage on one MI355X.

Workload (the GPU-at-scale configuration): hydro, PLM (MC) + HLLC(theta), RK2,
256^3 cells as 4096 blocks of 16^3, periodic box with a seeded smooth density
perturbation, uniform pressure and a smoothed Gaussian velocity field of rms
0.3 (fornaxamr_amd.problems.turbulent_box), CFL 0.4.  A step is one RK2 step
(CFL dt on the device, stage 1, stage 2) on the device-resident mesh; the
metric counts zone-updates (one active cell through one RK stage).

Prints one JSON line.  Also reports the hydro stage kernel against the HBM
roofline (event-timed live on the mesh's stream) and a CPU baseline: the
serial reference build (oracle/_ref) when present, else the C restatement,
timed on this host's cores over a bounded sample (the same scheme on a 64^3
periodic box, one process per core).

The same line carries the batched-ensemble figure: 48 independent periodic
turbulent boxes with edges drawn from {32,48,64,96,128} and gamma from
[1.3,1.7] (fornaxamr_amd.problems.ragged_ensemble), advanced by one batched
launch per stage captured in a hipGraph with the autotuned schedule, against
one set of launches per instance, and the cost of the discrete adjoint (reverse
sweep with recompute) relative to the forward on the 64^3 acoustic box, and
the per-block drop-in (fx_flux_update_blocks: prj_flux_update on 4096
device-resident blocks in the reference's own layouts, one launch).
The MHD stage (magnetized Sedov, PLM(MC)+HLLD+upwind CT) at 256^3 and 128^3
is reported in the same line under ``mhd``; ``--workload mhd`` prints a line
for the MHD stage alone instead.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per zone-update: stage 1 reads W (6) + P, gamma (2) and
# writes W (6); stage 2 also reads W^n (6)  ->  112 B and 160 B
BYTES_STAGE = {1: 112.0, 2: 160.0}
# MHD: + B (3 cell, 3 face) read and written, + W^n and Bf^n in stage 2
BYTES_STAGE_MHD = {1: 208.0, 2: 304.0}
METRIC = "FP64 zone-updates/s per RK stage (plain vs batched) + adjoint/forward cost ratio"


def _progress(msg):
    """A line on stderr per phase, so a long run shows that it is alive."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _timed_steps(step, sync, steps):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    sync()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def run_ensemble(steps, warmup, count=48):
    """Batched (autotuned, graph-captured) vs per-instance launches."""
    from fornaxamr_amd import problems as P
    from fornaxamr_amd.ensemble import GpuEnsemble
    probs = P.ragged_ensemble(count=count)
    e = GpuEnsemble(probs)
    cfl = 0.4

    def measure(graph):
        e.reset()
        e.set_graph(graph)
        for _ in range(warmup):
            e.step_async(cfl)
        e.sync()
        return _timed_steps(lambda: e.step_async(cfl), e.sync, steps) / steps

    e.set_schedule(0, 0, 1, 0)
    t_inst = measure(False)
    t_inst_graph = measure(True)
    e.set_graph(True)
    best = e.autotune(cfl, 1.2, reps=3)
    t_batch = measure(True)
    zu = e.ncells * 2
    e.close()
    sizes = sorted({p.n[0] for p in probs})
    return {
        "value": zu / t_batch, "unit": "zone-updates/s",
        "per_instance_value": zu / t_inst,
        "per_instance_graph_value": zu / t_inst_graph,
        "speedup_vs_per_instance": t_inst / t_batch,
        "ms_per_step": t_batch * 1000.0, "per_instance_ms_per_step": t_inst * 1000.0,
        "schedule": {k: best[k] for k in ("mode", "persistent_wgs", "xcd", "table")},
        "workload": (f"{count} periodic turbulent boxes, edges from {sizes}, gamma U[1.3,1.7], "
                     f"{e.nblocks} blocks, {e.ncells} cells, hipGraph per step"),
    }


def run_adjoint(reps=3):
    """Cost of the discrete adjoint (reverse sweep incl. recompute) against the
    forward, on the acoustic box: 64^3, 40 RK2 steps, checkpoint every 8."""
    import torch
    from fornaxamr_amd import problems as P
    from fornaxamr_amd.adjoint import AdjointSolver
    pr = P.acoustic_box(n=(64, 64, 64))
    dt = P.cfl_dt(pr.W, pr.gamma, [1.0 / 64] * 3, pr.cfl)
    solver = AdjointSolver(pr.n, pr.xlim, pr.gamma, checkpoint_every=8)
    W0 = torch.from_numpy(pr.W).cuda()
    mask = torch.from_numpy(P.subbox_mask(pr.n, pr.xlim)).cuda()
    nsteps = 40
    WT = solver.forward(W0, nsteps, dt)
    gWT = torch.zeros_like(WT)
    gWT[0] = 2.0 * WT[0] * mask
    solver.backward(gWT)
    torch.cuda.synchronize()
    tf = tb = 1e30
    for _ in range(reps):
        a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        a.record()
        solver.forward(W0, nsteps, dt)
        b.record()
        solver.backward(gWT)
        c.record()
        torch.cuda.synchronize()
        tf = min(tf, a.elapsed_time(b))
        tb = min(tb, b.elapsed_time(c))
    solver.close()
    zu = pr.ncells * 2 * nsteps
    return {"ratio": tb / tf, "forward_ms": tf, "backward_ms": tb,
            "forward_zone_updates_per_s": zu / (tf / 1000.0),
            "workload": "van Albada + HLLE RK2, 64^3 periodic acoustic box, 40 steps, "
                        "checkpoint every 8 (reverse sweep recomputes each segment)"}


# per-block drop-in: algorithmic bytes of one prj_flux_update call on a hydro
# block in the reference's layouts -- read the W slot (6 x 20^3) and eosvar P,
# gamma (2 x 20^3), write the 13,056 faces' fluxes (6) and face velocities (3)
BLOCK_CELLS = 20 ** 3
BLOCK_FACES = 3 * 17 * 16 * 16
BYTES_BLOCK = 8.0 * (8 * BLOCK_CELLS + 9 * BLOCK_FACES)


def _host_block_call_ms(lib, mhd, reps=50):
    """Wall time of one fx_flux_update_block call (prj_flux_update on one block
    in host memory: copies in, one launch, copies out, wait), median of reps."""
    import ctypes as C
    import numpy as np
    import pyoracle as O
    from fornaxamr_amd import _native
    slab, gam, bf = O.block_inputs(5, bool(mhd))
    nh = 9 if mhd else 6
    eos = O.block_eosvar(slab, 0, gam)
    W = np.ascontiguousarray(slab.reshape(nh, -1))
    keep = [W, eos]
    io = _native.BlockIO()
    io.W = W.ctypes.data
    io.var_stride = 2 * BLOCK_CELLS
    io.eosvar = eos.ctypes.data
    for d in range(3):
        for name, size in (("flux", nh * BLOCK_CELLS), ("v_riemann", 3 * BLOCK_CELLS)):
            a = np.zeros(size)
            keep.append(a)
            getattr(io, name)[d] = a.ctypes.data
        if mhd:
            b = np.ascontiguousarray(bf[d])
            keep.append(b)
            io.bf[d] = b.ctypes.data
            for name in ("Bv1", "Bv2"):
                a = np.zeros(BLOCK_CELLS)
                keep.append(a)
                getattr(io, name)[d] = a.ctypes.data
    times = []
    for r in range(reps + 3):
        t0 = time.perf_counter()
        _native.check(lib.fx_flux_update_block(C.byref(io), mhd))
        if r >= 3:
            times.append(time.perf_counter() - t0)
    return float(np.median(times)) * 1000.0


def run_per_block(nblocks=4096, reps=10):
    """fx_flux_update_blocks (prj_flux_update over device-resident blocks in the
    reference's own WIDX / EIDX / VIDX / VRIDX layouts), hydro, one launch."""
    import ctypes as C
    import torch
    from fornaxamr_amd import _native
    lib = _native.load()
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    gam = 5.0 / 3.0
    # W_mhd slab: [var][slot][cell]; slot 0 holds the stage input
    W = torch.empty((nblocks, 6, 2, BLOCK_CELLS), dtype=torch.float64, device=dev)
    W[:, 0] = 1.0 + 0.1 * torch.rand((nblocks, 2, BLOCK_CELLS), generator=g, device=dev,
                                       dtype=torch.float64)
    W[:, 1:4] = 0.3 * torch.randn((nblocks, 3, 2, BLOCK_CELLS), generator=g, device=dev,
                                  dtype=torch.float64)
    W[:, 4] = 1.5 + 0.1 * torch.rand((nblocks, 2, BLOCK_CELLS), generator=g, device=dev,
                                     dtype=torch.float64)
    W[:, 5] = 0.1
    # eosvar (AoSoA, 3 components: P, T, gamma); every entry is positive, which
    # is all the flux update needs of it here
    E = torch.empty((nblocks, 3 * BLOCK_CELLS), dtype=torch.float64, device=dev)
    E.uniform_(0.5, 1.0, generator=g)
    E.view(nblocks, -1, 3, 4)[:, :, 2, :] = gam
    F = torch.empty((3, nblocks, 6 * BLOCK_CELLS), dtype=torch.float64, device=dev)
    V = torch.empty((3, nblocks, 3 * BLOCK_CELLS), dtype=torch.float64, device=dev)
    descs = (_native.BlockIO * nblocks)()
    for b in range(nblocks):
        io = descs[b]
        io.W = W[b].data_ptr()
        io.var_stride = 2 * BLOCK_CELLS
        io.eosvar = E[b].data_ptr()
        for d in range(3):
            io.flux[d] = F[d, b].data_ptr()
            io.v_riemann[d] = V[d, b].data_ptr()
    raw = torch.frombuffer(bytearray(bytes(descs)), dtype=torch.uint8).to(dev)
    status = torch.zeros(8, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def launch():
        _native.check(lib.fx_flux_update_blocks(nblocks, C.c_void_p(raw.data_ptr()), 0,
                                                C.c_void_p(status.data_ptr()),
                                                C.c_void_p(stream.cuda_stream)))
    launch()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        launch()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    gbs = BYTES_BLOCK * nblocks / (ms * 1e-3) / 1e9
    host = {name: _host_block_call_ms(lib, mhd) for name, mhd in (("hydro", 0), ("mhd", 1))}
    return {"kernel": "block_pencil_kernel<0> (x, y faces) + block_face_kernel<0, 2> (z faces): "
                      "prj_flux_update on the reference's block layouts",
            "blocks": nblocks, "ms_per_launch": ms,
            "host_array_call_ms": host,
            "faces_per_s": BLOCK_FACES * nblocks / (ms * 1e-3),
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS}}


def _cpu_worker(args):
    kind, nsteps = args
    import numpy as np  # noqa: F401
    import pyoracle as O
    from fornaxamr_amd import problems as P
    pr = P.turbulent_box(n=(64, 64, 64), seed=3)
    if kind == "reference":
        sim = O.RefSim(pr.n, pr.xlim, pr.bc, pr.gamma)
    else:
        sim = O.OracleSim(pr.n, pr.xlim, pr.bc, pr.gamma)
    sim.set_state(pr.W)
    sim.step(pr.cfl)  # warm-up
    t0 = time.perf_counter()
    for _ in range(nsteps):
        sim.step(pr.cfl)
    dt = time.perf_counter() - t0
    return pr.ncells * 2 * nsteps, dt


# CPU share of one GPU on the GPU hosts: worker pools are sized to it
CPU_SHARE = 16


def cpu_baseline(nsteps=6):
    """Run before any GPU call: the workers are forked from this process.
    One process per core the process may run on (its affinity set), up to the
    16-core share of one GPU; each steps the C2 scheme and generator on a
    64^3 sample box (the reference's per-core rate does not depend on the mesh
    size: 1.9-2.2e6 zone-updates/s per core at 64^3 and 128^3 alike)."""
    import multiprocessing as mp
    import pyoracle as O
    kind = "reference" if O.ref_available() else "port"
    if kind == "port":
        O.ensure_built()
        os.environ["OMP_NUM_THREADS"] = "1"
    affinity = len(os.sched_getaffinity(0))
    cores = max(1, min(CPU_SHARE, affinity))
    ctx = mp.get_context("fork")
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(kind, nsteps)] * cores)
    zu = sum(r[0] for r in res)
    wall = max(r[1] for r in res)
    return {"value": zu / wall, "unit": "zone-updates/s", "cores": cores, "kind": kind,
            "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "per_core": zu / wall / cores,
            "sample": (f"{cores} processes x ({nsteps} RK2 steps of the C2 scheme, "
                       f"turbulent_box generator at 64^3 = 64 blocks of 16^3), one process per "
                       f"core (all {affinity} cores of the affinity set, at most the "
                       f"{CPU_SHARE}-core share of one GPU), "
                       f"{'serial reference build' if kind == 'reference' else 'C restatement'}")}


def run_mesh(pr, bytes_stage, steps, warmup, integrator="rk2", schedule=None):
    """Steps of one device-resident mesh; stage kernels event-timed on its stream.
    Returns zone-updates/s counting every stage (eSSPRK(10,4): 10 substeps).
    schedule: (ct_split, side_stream) of an MHD mesh, None for the automatic one."""
    from fornaxamr_amd.mesh import GpuMesh
    mesh = GpuMesh.from_problem(pr)
    mesh.set_integrator(integrator)
    if schedule is not None:
        mesh.set_mhd_schedule(*schedule)
    stages = 10 if integrator == "essprk10_4" else 2
    for _ in range(warmup):
        mesh.step_async(pr.cfl)
    mesh.sync()
    mesh.set_timing(True)
    elapsed = _timed_steps(lambda: mesh.step_async(pr.cfl), mesh.sync, steps)
    ms1, n1 = mesh.kernel_time(1)
    ms2, n2 = mesh.kernel_time(2)
    mesh.set_timing(False)
    mesh.close()
    ncells = pr.ncells
    bytes_launch_avg = ncells * (n1 * bytes_stage[1] + n2 * bytes_stage[2]) / max(n1 + n2, 1)
    avg_launch_s = (ms1 + ms2) / 1000.0 / (n1 + n2)
    achieved = bytes_launch_avg / avg_launch_s / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "avg_launch_ms": avg_launch_s * 1000.0,
            "stage_ms": [ms1 / max(n1, 1), ms2 / max(n2, 1)]}
    return ncells * stages * steps / elapsed, elapsed / steps * 1000.0, roof


def run_mhd_line(n, steps, warmup, schedule=None):
    """The MHD stage on the magnetized Sedov problem at n^3 (C3 at 128^3; the
    256^3 line is the headline MHD target): zone-updates/s and the roofline of
    the stage's kernels (208 / 304 B per zone, event-timed live).  schedule:
    (ct_split, side_stream) in place of the automatic launch schedule."""
    from fornaxamr_amd import problems as P
    pr = P.magnetized_sedov(n=(n, n, n))
    value, ms, roof = run_mesh(pr, BYTES_STAGE_MHD, steps, warmup, schedule=schedule)
    roof["kernel"] = "mhd_flux_kernel + mhd_ghostface_kernel + mhd_ct_kernel"
    roof["traffic"] = _traffic(f"mhd_{n}")
    line = {"value": value, "unit": "zone-updates/s", "ms_per_step": ms,
            "workload": (f"MHD PLM(MC)+HLLD+upwind CT RK2, {n}^3 magnetized Sedov (outflow), "
                         f"{pr.nblocks} blocks of 16^3, CFL 0.4"),
            "roofline": roof}
    if schedule is not None:
        line["schedule"] = {"ct_split": schedule[0], "side_stream": schedule[1]}
    return line


def _traffic(name):
    """HBM bytes per launch from the committed counter summary, if present."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["hydro", "mhd"], default="hydro")
    ap.add_argument("--integrator", choices=["rk2", "essprk10_4"], default="rk2",
                    help="time integrator of the hydro line (essprk10_4: the reference's "
                         "default eSSPRK(10,4); stage kernels are its forward-Euler substeps)")
    ap.add_argument("--n", type=int, default=None, help="cells per axis (multiple of 16); "
                    "default 256 (hydro), 128 (mhd)")
    ap.add_argument("--mhd-schedule", default=None, metavar="CT,SIDE",
                    help="MHD launch schedule instead of the automatic one: CT workgroups "
                         "per block (1, 2, 4) and the ghost-layer faces in order (0) or on "
                         "a second stream (1), e.g. 1,0 (fx_mesh_set_mhd_schedule)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble figure")
    ap.add_argument("--no-mhd", action="store_true",
                    help="skip the MHD lines (magnetized Sedov at 256^3 and 128^3)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost ratio")
    ap.add_argument("--no-per-block", action="store_true",
                    help="skip the per-block drop-in figure (fx_flux_update_blocks)")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only the batched-ensemble figure (for profiling)")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="first run a plain device copy of this many GiB (3 times): the "
                         "calibration kernel for HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the stage shards by blocks over GPUs; on one GPU every block is resident)",
              file=sys.stderr)
        sys.exit(2)
    mhd = args.workload == "mhd"
    n = args.n or (128 if mhd else 256)

    _progress("cpu baseline")
    cpu = None if (args.no_cpu or mhd or args.ensemble_only) else cpu_baseline()
    _progress("gpu lines")

    import torch
    from fornaxamr_amd import problems as P

    torch.cuda.set_device(0)
    if args.copy_gib > 0:
        nelem = int(args.copy_gib * (1 << 30)) // 8
        src = torch.ones(nelem, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.ensemble_only:
        ens = run_ensemble(args.steps, args.warmup)
        print(json.dumps({"ensemble": ens}), flush=True)
        return

    schedule = (tuple(int(v) for v in args.mhd_schedule.split(","))
                if args.mhd_schedule else None)
    if mhd:
        line = run_mhd_line(n, args.steps, args.warmup, schedule)
        value, ms, roof, workload = (line["value"], line["ms_per_step"], line["roofline"],
                                     line["workload"])
        pr = P.magnetized_sedov(n=(n, n, n))
    else:
        pr = P.turbulent_box(n=(n, n, n), seed=1)
        value, ms, roof = run_mesh(pr, BYTES_STAGE, args.steps, args.warmup, args.integrator)
        roof["kernel"] = "hydro_stage_kernel<1|2>"
        roof["traffic"] = _traffic(f"hydro_{n}")
        scheme = "RK2" if args.integrator == "rk2" else "eSSPRK(10,4)"
        workload = (f"hydro PLM(MC)+HLLC {scheme}, {n}^3 periodic turbulent box, "
                    f"{pr.nblocks} blocks of 16^3, CFL 0.4")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "cells": pr.ncells, "blocks": pr.nblocks,
                   "stages_per_step": 10 if args.integrator == "essprk10_4" else 2},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if not mhd and not args.no_mhd:
        out["mhd"] = {}
        for m in (256, 128):
            _progress(f"mhd {m}^3")
            out["mhd"][str(m)] = run_mhd_line(m, args.steps, args.warmup, schedule)
    if not mhd and not args.no_ensemble:
        _progress("ensemble")
        out["batched"] = run_ensemble(args.steps, args.warmup)
    if not mhd and not args.no_adjoint:
        _progress("adjoint")
        out["adjoint"] = run_adjoint()
    if not mhd and not args.no_per_block:
        _progress("per-block")
        out["per_block"] = run_per_block()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
