"""Write the golden fixtures of tests/golden/ from the serial reference build.

Runs in the development container only (it needs oracle/_ref, which
oracle/Makefile compiles from the reference sources).  Fixtures are data:
per-stencil / per-face / per-edge input/output vectors of the reference's own
prj_reconstruct_mc_face_slope, prj_riemann_hllc, prj_riemann_hlld and
prj_mhd_emf_upwind, and, for whole
RK2 trajectories, the SHA-256 of the input state and of the state after every
step plus the dt sequence and a few sampled values; for the per-block entry
points, the SHA-256 of the reference's prj_flux_update / prj_flux_div outputs
on seeded raw blocks; for the coarse-fine corrections of refined meshes,
the mesh topologies the reference's own AMR code builds (cf_mesh_*.npz:
block geometry, levels and neighbour slots) and the SHA-256 of the
reference's prj_riemann_flux_send / prj_mhd_emf_send outputs on seeded
block arrays (cf_corrections.json).

    python tests/golden/make_golden.py [--only-block | --only-cf | --only-emf-block |
                                        --only-trajectories | --only-bf2bc | --only-mc]
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as O  # noqa: E402
from fornaxamr_amd import problems as P  # noqa: E402

_dp = C.POINTER(C.c_double)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def rand_states(rng, n, nh, mhd):
    """Face states covering sub- and supersonic flow in both directions."""
    W = np.zeros((2, n, nh))
    for s in range(2):
        W[s, :, 0] = rng.uniform(0.1, 3.0, n)                 # rho
        W[s, :, 1] = rng.normal(0.0, 1.5, n)                  # vn
        W[s, :, 2:4] = rng.normal(0.0, 1.0, (n, 2))           # vt
        W[s, :, 4] = rng.uniform(0.2, 4.0, n)                 # eint
        W[s, :, 5] = rng.uniform(0.05, 0.5, n)                # ye
        if mhd:
            W[s, :, 6:9] = rng.normal(0.0, 1.0, (n, 3))
    # some strongly supersonic pairs
    W[:, : n // 8, 1] += 6.0
    W[:, n // 8: n // 4, 1] -= 6.0
    return W


def hllc_vectors(lib, rng, n=512):
    W = rand_states(rng, n, 6, False)
    g = rng.uniform(1.3, 1.7, (2, n))
    p = (g - 1.0) * W[:, :, 0] * W[:, :, 4] * rng.uniform(0.8, 1.2, (2, n))
    d = rng.normal(0.0, 0.5, (3, n))
    F = np.zeros((n, 6))
    V = np.zeros((n, 3))
    for i in range(n):
        wl = np.ascontiguousarray(W[0, i])
        wr = np.ascontiguousarray(W[1, i])
        f = np.zeros(6)
        v = np.zeros(3)
        lib.prj_riemann_hllc(wl.ctypes.data_as(_dp), wr.ctypes.data_as(_dp), p[0, i], p[1, i],
                             g[0, i], g[1, i], None, f.ctypes.data_as(_dp), v.ctypes.data_as(_dp),
                             d[0, i], d[1, i], d[2, i])
        F[i] = f
        V[i] = v
    return dict(WL=W[0], WR=W[1], pL=p[0], pR=p[1], gL=g[0], gR=g[1], du=d[0], dv=d[1], dw=d[2],
                flux=F, v_face=V)


def hlld_vectors(lib, rng, n=512):
    W = rand_states(rng, n, 9, True)
    g = np.full((2, n), 5.0 / 3.0)
    p = (g - 1.0) * W[:, :, 0] * W[:, :, 4]
    bn = rng.normal(0.0, 1.0, n)
    bn[:32] = 0.0                                   # bn_small switch
    bn[32:48] *= 1e-6
    # strong rarefactions: exercise the HLL fallback (p*_T <= 0)
    W[0, 48:80, 1] = -8.0
    W[1, 48:80, 1] = 8.0
    W[:, 48:80, 4] = 0.01
    p[:, 48:80] = (g[:, 48:80] - 1.0) * W[:, 48:80, 0] * W[:, 48:80, 4]
    W[:, :, 6] = bn
    F = np.zeros((n, 9))
    V = np.zeros((n, 3))
    E = np.zeros((n, 2))
    for i in range(n):
        wl = np.ascontiguousarray(W[0, i])
        wr = np.ascontiguousarray(W[1, i])
        f = np.zeros(9)
        v = np.zeros(3)
        b1 = C.c_double()
        b2 = C.c_double()
        lib.prj_riemann_hlld(wl.ctypes.data_as(_dp), wr.ctypes.data_as(_dp), p[0, i], p[1, i],
                             g[0, i], g[1, i], None, bn[i], f.ctypes.data_as(_dp),
                             v.ctypes.data_as(_dp), C.byref(b1), C.byref(b2), 0.0, 0.0, 0.0)
        F[i] = f
        V[i] = v
        E[i] = (b1.value, b2.value)
    return dict(WL=W[0], WR=W[1], pL=p[0], pR=p[1], gL=g[0], gR=g[1], bn=bn, flux=F, v_face=V,
                bv=E)


def emf_vectors(lib, rng, n=256):
    ef = rng.normal(0.0, 1.0, (n, 4))
    ec = rng.normal(0.0, 1.0, (n, 4))
    vn = rng.normal(0.0, 1.0, (n, 4))
    vn[rng.uniform(size=(n, 4)) < 0.2] = 0.0
    out = np.zeros(n)
    for i in range(n):
        a = np.ascontiguousarray(ef[i])
        b = np.ascontiguousarray(ec[i])
        c = np.ascontiguousarray(vn[i])
        out[i] = lib.refh_emf_upwind(i % 3, a.ctypes.data_as(_dp), b.ctypes.data_as(_dp),
                                     c.ctypes.data_as(_dp))
    return dict(emf_face=ef, emf_cell=ec, v_norm=vn, emf=out)


def emf_block_vectors(lib, rng, n=384):
    """prj_mhd_emf_upwind with the resistive term on one block's face arrays:
    random faces (3 FACE_IDX arrays), edges of every direction over the range
    the reference's edge loop and its curl reach (block indices -1..16),
    anisotropic cell widths (not powers of two), eta zero on a quarter of the
    edges.  `stored` is what the reference wrote to block->emf."""
    bf = [rng.normal(0.0, 1.0, O.BLOCK_NFACES) for _ in range(3)]
    dx = np.array([0.03125, 0.05, 0.07])
    edge = np.zeros((n, 4), dtype=np.int32)
    edge[:, 0] = rng.integers(0, 3, n)
    edge[:, 1:] = rng.integers(-1, 17, (n, 3))
    ef = rng.normal(0.0, 1.0, (n, 4))
    ec = rng.normal(0.0, 1.0, (n, 4))
    vn = rng.normal(0.0, 1.0, (n, 4))
    vn[rng.uniform(size=(n, 4)) < 0.2] = 0.0
    eta = rng.uniform(0.0, 2.0, n)
    eta[rng.uniform(size=n) < 0.25] = 0.0
    out = np.zeros(n)
    stored = np.zeros(n)
    bp = [np.ascontiguousarray(b) for b in bf]
    for m in range(n):
        a = [np.ascontiguousarray(x[m]) for x in (ef, ec, vn)]
        st = C.c_double()
        out[m] = lib.refh_emf_upwind_block(int(edge[m, 0]), int(edge[m, 1]), int(edge[m, 2]),
                                           int(edge[m, 3]), *[x.ctypes.data_as(_dp) for x in a],
                                           *[b.ctypes.data_as(_dp) for b in bp],
                                           dx.ctypes.data_as(_dp), float(eta[m]), C.byref(st))
        stored[m] = st.value
    return dict(edge=edge, emf_face=ef, emf_cell=ec, v_norm=vn, eta=eta, bfx=bf[0], bfy=bf[1],
                bfz=bf[2], dx=dx, emf=out, stored=stored)


TRAJECTORIES = {
    # name: (factory, kwargs, steps)
    "sedov_c1_64": ("sedov", dict(n=(64, 64, 64)), 20),
    "periodic_32": ("random_smooth", dict(n=(32, 32, 32), seed=11, bc=(3,) * 6, gamma=1.4), 6),
    "reflect_mixed_32x48x32": ("random_smooth", dict(n=(32, 48, 32), seed=12, bc=(1, 0, 3, 3, 0, 1)), 6),
    "mhd_random_32": ("random_smooth", dict(n=(32, 32, 32), seed=13, bc=(0,) * 6, mhd=True), 6),
    "mhd_sedov_32": ("magnetized_sedov", dict(n=(32, 32, 32)), 6),
    # resistive CT: uniform mhd_eta, and the anomalous mask on a dense patch
    "mhd_eta_32": ("random_smooth", dict(n=(32, 32, 32), seed=14, bc=(0,) * 6, mhd=True,
                                         eta=1.0e-3), 4),
    "mhd_eta_mask_32": ("dense_resistive", dict(n=(32, 32, 32), seed=3, eta=1.0e-3), 4),
}


def mc_vectors(lib, rng, n=4096):
    """prj_reconstruct_mc_face_slope on stencils (qm, q0, qp): random smooth
    and rough data, and the special cases of the limiter -- equal neighbours,
    sign changes, ties between the three candidates, signed zeros,
    subnormal differences, differences that overflow when doubled,
    infinities and NaN."""
    q = rng.normal(0.0, 1.0, (n, 3))
    q[: n // 8] = np.cumsum(rng.uniform(0.0, 1.0, (n // 8, 3)), axis=1)     # monotone up
    q[n // 8: n // 4] = -np.cumsum(rng.uniform(0.0, 1.0, (n // 8, 3)), axis=1)
    special = []
    tiny, huge, inf, nan = 5e-324, 1.7e308, np.inf, np.nan
    vals = [0.0, -0.0, 1.0, -1.0, 2.0, 3.0, tiny, -tiny, 1e-310, -1e-310, huge, -huge, inf, -inf,
            nan, 0.5, 1.5, 1e-300]
    for a in vals:
        for b in vals:
            for c in vals:
                special.append((a, b, c))
    special += [(0.0, 1.0, 3.0), (0.0, 1.0, 1.5), (0.0, 2.0, 3.0), (3.0, 1.0, 0.0),
                (1.0, 1.0, 2.0), (0.0, 0.0, 0.0), (-0.0, 0.0, -0.0)]
    q = np.concatenate([q, np.array(special)])
    qm, q0, qp = (np.ascontiguousarray(q[:, i]) for i in range(3))
    out = np.empty(len(q))
    lib.refh_mc_face_slope(len(q), qm.ctypes.data_as(_dp), q0.ctypes.data_as(_dp),
                           qp.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return dict(qm=qm, q0=q0, qp=qp, slope=out)


def make_problem(name):
    fac, kw, steps = TRAJECTORIES[name]
    return getattr(P, fac)(**kw), steps


def trajectory(name):
    pr, steps = make_problem(name)
    ref = O.RefSim(pr.n, pr.xlim, pr.bc, pr.gamma, mhd=pr.mhd, eta=pr.eta)
    ref.set_state(pr.W, pr.bf)
    rec = dict(problem=name, steps=steps, cfl=pr.cfl, input_sha=sha(pr.W),
               input_bf_sha=[sha(b) for b in pr.bf] if pr.mhd else None,
               dt=[], W_sha=[], bf_sha=[], samples=[])
    rng = np.random.default_rng(7)
    idx = [tuple(int(rng.integers(0, s)) for s in pr.n) for _ in range(16)]
    for _ in range(steps):
        rec["dt"].append(ref.step(pr.cfl))
        W = ref.get_prim(0)
        rec["W_sha"].append(sha(W))
        if pr.mhd:
            rec["bf_sha"].append([sha(b) for b in ref.get_bf(0)])
    W = ref.get_prim(0)
    rec["samples"] = [[list(c), [float(W[(v,) + c]) for v in range(W.shape[0])]] for c in idx]
    ref.close()
    return rec


BLOCK_CASES = {"hydro": (0, 5, 1), "mhd": (1, 6, 1), "mhd_slot0": (1, 7, 0)}


def block_flux():
    """prj_flux_update and prj_flux_div of the reference on one raw block per
    case (inputs from pyoracle.block_inputs(seed, mhd)): SHA-256 of every
    output array (untouched entries NaN) and a few sampled values."""
    recs = {}
    for name, (mhd, seed, slot) in BLOCK_CASES.items():
        slab, gam, bf = O.block_inputs(seed, mhd)
        eos = O.block_eosvar(slab, slot, gam)
        out = O.ref_flux_update_raw(mhd, slot, slab, eos, bf)
        fdiv, area, vol = O.ref_flux_div_raw(mhd, out["flux"])
        rec = dict(mhd=mhd, seed=seed, slot=slot, input_sha=sha(slab), eosvar_sha=sha(eos),
                   bf_sha=[sha(b) for b in bf] if mhd else None,
                   out_sha={k: [sha(a) for a in v] for k, v in out.items()},
                   flux_div_sha=sha(fdiv), area=list(map(float, area)), vol=vol,
                   samples=[[int(i), float(out["flux"][0][i])]
                            for i in np.flatnonzero(np.isfinite(out["flux"][0]))[::1500]])
        recs[name] = rec
    return recs


BF2BC_CASES = {"all_slot0": (31, 0, False), "all_slot1": (32, 1, False),
               "done_slot1": (33, 1, True)}


def bf2bc_fixtures():
    """prj_mhd_bf2bc(_all) of the reference on raw MHD blocks (inputs from
    pyoracle.block_inputs(seed, True), the done flags from the seed): SHA-256
    of the W slab it leaves."""
    recs = {}
    for name, (seed, slot, use_done) in BF2BC_CASES.items():
        slab, _, bf = O.block_inputs(seed, True)
        done = O.bf2bc_done_flags(seed) if use_done else None
        out = O.ref_bf2bc_raw(slot, slab, bf, done)
        recs[name] = dict(seed=seed, slot=slot, use_done=use_done, input_sha=sha(slab),
                          out_sha=sha(out))
    return recs


CF_MESHES = {
    # two levels, outflow
    "two_level": dict(root=(2, 2, 2), xlim=(-1, 1, -1, 1, -1, 1), bc=(0,) * 6, max_level=2,
                      ball=(0.3, 0.2, -0.4, 0.2)),
    # same refinement, periodic: slots across the seam see the far block
    "two_level_periodic": dict(root=(2, 2, 2), xlim=(-1, 1, -1, 1, -1, 1), bc=(3,) * 6,
                               max_level=2, ball=(0.3, 0.2, -0.4, 0.2)),
    # three active levels on a 2x1x1 root (blocks that both receive and send)
    "three_level": dict(root=(2, 1, 1), xlim=(-2, 2, -1, 1, -1, 1), bc=(0,) * 6, max_level=3,
                        ball=(0.55, 0.3, -0.45, 0.03)),
}
CF_SEEDS = {"flux6": 101, "flux9": 102, "emf": 103}


def cf_fixtures():
    recs = {}
    for name, cfg in CF_MESHES.items():
        rec = dict(cfg)
        mesh = O.AmrMesh.from_reference(mhd=False, **cfg)
        mesh.save(os.path.join(HERE, f"cf_mesh_{name}.npz"))
        flux = O.cf_inputs(mesh, 6, CF_SEEDS["flux6"])
        rec["flux6_in"] = O.cf_digest(flux)
        rec["flux6_out"] = O.cf_digest(O.ref_cf_flux_send(False, mesh, flux))
        mesh_m = O.AmrMesh.from_reference(mhd=True, **cfg)
        assert np.array_equal(mesh_m.slots, mesh.slots) and np.array_equal(mesh_m.geo, mesh.geo)
        flux = O.cf_inputs(mesh, 9, CF_SEEDS["flux9"])
        rec["flux9_in"] = O.cf_digest(flux)
        rec["flux9_out"] = O.cf_digest(O.ref_cf_flux_send(True, mesh_m, flux))
        emf = O.cf_emf_inputs(mesh, CF_SEEDS["emf"])
        rec["emf_in"] = O.cf_digest(emf)
        e, f = O.ref_cf_emf_send(mesh_m, emf)
        loc = mesh.local()
        rec["emf_out"] = O.cf_digest(e)
        rec["fidelity_local_out"] = O.cf_digest(f, loc)
        rec["seeds"] = CF_SEEDS
        recs[name] = rec
    with open(os.path.join(HERE, "cf_corrections.json"), "w") as fh:
        json.dump(recs, fh, indent=1)


def main():
    if "--only-cf" in sys.argv:
        cf_fixtures()
        return
    if "--only-emf-block" in sys.argv:
        O.RefSim((32, 32, 32), (0, 1, 0, 1, 0, 1), (0,) * 6, 5.0 / 3.0, mhd=True)
        np.savez_compressed(os.path.join(HERE, "emf_block_vectors.npz"),
                            **emf_block_vectors(O.ref_lib(True), np.random.default_rng(20261203)))
        return
    if "--only-mc" in sys.argv:
        np.savez_compressed(os.path.join(HERE, "mc_vectors.npz"),
                            **mc_vectors(O.ref_lib(False), np.random.default_rng(20261204)))
        return
    if "--only-trajectories" in sys.argv:
        recs = {name: trajectory(name) for name in TRAJECTORIES}
        with open(os.path.join(HERE, "trajectories.json"), "w") as f:
            json.dump(recs, f, indent=1)
        return
    if "--only-bf2bc" in sys.argv:
        with open(os.path.join(HERE, "bf2bc.json"), "w") as f:
            json.dump(bf2bc_fixtures(), f, indent=1)
        return
    if "--only-block" in sys.argv:
        with open(os.path.join(HERE, "block_flux.json"), "w") as f:
            json.dump(block_flux(), f, indent=1)
        return
    rng = np.random.default_rng(20261202)
    libh = O.ref_lib(False)
    libm = O.ref_lib(True)
    # per-edge entry needs an initialised MHD mesh for its scratch block
    O.RefSim((32, 32, 32), (0, 1, 0, 1, 0, 1), (0,) * 6, 5.0 / 3.0, mhd=True)
    np.savez(os.path.join(HERE, "hllc_vectors.npz"), **hllc_vectors(libh, rng))
    np.savez(os.path.join(HERE, "hlld_vectors.npz"), **hlld_vectors(libm, rng))
    np.savez(os.path.join(HERE, "emf_vectors.npz"), **emf_vectors(libm, rng))
    np.savez_compressed(os.path.join(HERE, "emf_block_vectors.npz"),
                        **emf_block_vectors(libm, np.random.default_rng(20261203)))
    np.savez_compressed(os.path.join(HERE, "mc_vectors.npz"),
                        **mc_vectors(libh, np.random.default_rng(20261204)))
    recs = {name: trajectory(name) for name in TRAJECTORIES}
    with open(os.path.join(HERE, "trajectories.json"), "w") as f:
        json.dump(recs, f, indent=1)
    with open(os.path.join(HERE, "block_flux.json"), "w") as f:
        json.dump(block_flux(), f, indent=1)
    with open(os.path.join(HERE, "bf2bc.json"), "w") as f:
        json.dump(bf2bc_fixtures(), f, indent=1)
    cf_fixtures()
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
