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


def cpu_baseline(nsteps=6):
    """Run before any GPU call: the workers are forked from this process."""
    import multiprocessing as mp
    import pyoracle as O
    kind = "reference" if O.ref_available() else "port"
    if kind == "port":
        O.ensure_built()
        os.environ["OMP_NUM_THREADS"] = "1"
    cores = min(16, os.cpu_count() or 1)
    ctx = mp.get_context("fork")
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(kind, nsteps)] * cores)
    zu = sum(r[0] for r in res)
    wall = max(r[1] for r in res)
    return {"value": zu / wall, "unit": "zone-updates/s", "cores": cores, "kind": kind,
            "sample": (f"{cores} processes x ({nsteps} RK2 steps of the same scheme on a 64^3 "
                       f"periodic box, 64 blocks of 16^3), one process per core, "
                       f"{'serial reference build' if kind == 'reference' else 'C restatement'}")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="cells per axis (multiple of 16)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="first run a plain device copy of this many GiB (3 times): the "
                         "calibration kernel for HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the stage shards by blocks; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    cpu = None if args.no_cpu else cpu_baseline()

    import numpy as np  # noqa: F401
    import torch
    from fornaxamr_amd import problems as P
    from fornaxamr_amd.mesh import GpuMesh

    n = args.n
    pr = P.turbulent_box(n=(n, n, n), seed=1)
    torch.cuda.set_device(0)
    if args.copy_gib > 0:
        nelem = int(args.copy_gib * (1 << 30)) // 8
        src = torch.ones(nelem, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    mesh = GpuMesh.from_problem(pr)
    for _ in range(args.warmup):
        mesh.step_async(pr.cfl)
    mesh.sync()

    mesh.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mesh.step_async(pr.cfl)
    mesh.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    ms1, n1 = mesh.kernel_time(1)
    ms2, n2 = mesh.kernel_time(2)
    mesh.set_timing(False)
    ncells = pr.ncells
    zu = ncells * 2 * args.steps
    value = zu / elapsed
    bytes_launch_avg = ncells * (BYTES_STAGE[1] + BYTES_STAGE[2]) / 2.0
    avg_launch_s = (ms1 + ms2) / 1000.0 / (n1 + n2)
    achieved = bytes_launch_avg / avg_launch_s / 1e9
    out = {
        "metric": "FP64 zone-updates/s per RK stage",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1000.0,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"hydro PLM(MC)+HLLC RK2, {n}^3 periodic turbulent box, "
                               f"{pr.nblocks} blocks of 16^3, CFL 0.4",
                   "cells": ncells, "blocks": pr.nblocks, "stages_per_step": 2},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "hydro_stage_kernel<1|2>",
                     "avg_launch_ms": avg_launch_s * 1000.0,
                     "stage_ms": [ms1 / max(n1, 1), ms2 / max(n2, 1)]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
