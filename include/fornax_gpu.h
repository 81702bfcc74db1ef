/*
 * fornax_gpu.h -- C ABI of the MI355X (gfx950) implementation of FornaxAMR's
 * per-block Godunov RK2 stage.
 *
 * Plain C types only.  Every entry point returns FX_OK (0) or a negative
 * FX_E* code; fx_last_error() then holds a message (thread-local).  Physical
 * failures the reference aborts on (HLLD states, non-finite EMFs) are
 * recorded on the device and reported by the next fx_sync()/fx_step() with
 * the reference's message text.
 *
 * Array conventions on the host side ("global" arrays over the active cells
 * of one uniform mesh; float64):
 *   cells  W[var][X][Y][Z]   (Z fastest), var = rho, v1, v2, v3, eint, ye
 *                            [, b1, b2, b3]   (src/prj_defs.h:466-494)
 *   faces  Bx[X+1][Y][Z], By[X][Y+1][Z], Bz[X][Y][Z+1]
 * Boundary codes are the reference's (src/prj_defs.h:594-599):
 *   0 outflow (mirror copy), 1 reflect, 3 periodic.
 */
#ifndef FORNAX_GPU_H
#define FORNAX_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FX_OK 0
#define FX_EINVAL (-1)      /* bad argument */
#define FX_EHIP (-2)        /* HIP runtime error */
#define FX_EPHYS (-3)       /* physical state the reference aborts on */
#define FX_ENOMEM (-4)

typedef struct fx_mesh fx_mesh;

typedef struct fx_mesh_desc {
    int n[3];           /* active cells per axis; multiples of 16 (16^3 blocks) */
    double xlim[6];     /* x1min, x1max, x2min, x2max, x3min, x3max */
    int bc[6];          /* x1 inner/outer, x2 inner/outer, x3 inner/outer */
    double gamma;       /* ideal-gas adiabatic index (eosvar pressure/gamma refill) */
    int mhd;            /* 0: hydro (6 prims), 1: MHD with CT (9 prims + faces) */
    double eta;         /* uniform CT resistivity, mesh->mhd_eta */
} fx_mesh_desc;

/* ---- mesh lifetime (replaces prj_mesh_init + block storage for a uniform
 *      single-level mesh: src/prj_mesh.c:1438, src/prj_mesh.c:420) ---------- */
int fx_mesh_create(const fx_mesh_desc *desc, fx_mesh **out);
int fx_mesh_destroy(fx_mesh *m);
/* Stream all launches of this mesh go to (a hipStream_t); NULL = default.
 * An MHD mesh with at most two blocks per CU also runs its ghost-layer face
 * kernel on a stream of its own, forked from and joined back into this one
 * with events within each stage (graph capture follows the fork). */
int fx_mesh_set_stream(fx_mesh *m, void *hip_stream);

/* ---- state transfer (host arrays in the global layout above) ------------ */
/* Sets both stage slots.  For MHD the cell-centred B of W is ignored and
 * recomputed from the faces given by fx_mesh_set_bf (prj_mhd_bf2bc_all,
 * src/prj_mhd.c:615-619), so call fx_mesh_set_bf after fx_mesh_set_prim. */
int fx_mesh_set_prim(fx_mesh *m, const double *W);
int fx_mesh_get_prim(fx_mesh *m, int slot, double *W);
int fx_mesh_set_bf(fx_mesh *m, const double *bx, const double *by, const double *bz);
int fx_mesh_get_bf(fx_mesh *m, int slot, double *bx, double *by, double *bz);

/* ---- the RK2 path -------------------------------------------------------- */
/* CFL step of slot 0 (prj_timeint_calc_dt, src/prj_timeint.c:2144-2256),
 * with the integrator's factor: x 6 on an eSSPRK(10,4) mesh. */
int fx_calc_dt(fx_mesh *m, double cfl, double *dt);
/* One RK2 stage over all blocks: the body of prj_timeint_stage1/2
 * (src/prj_timeint.c:2732-2973) -- domain ghost fill, reconstruction,
 * Riemann fluxes, flux divergence, CT EMF and face update (MHD), per-cell
 * update.  stage 1: slot 0 -> slot 1; stage 2: slots 0,1 -> slot 0. */
int fx_stage(fx_mesh *m, int stage, double dt);
/* One main-loop step with dt kept on the device: CFL dt, growth limit
 * dt <= dt_factor * dt_prev (src/main.c:510-522), stage 1, stage 2.
 * Does not synchronise.  Stage 2 records the CFL maxima of the state it
 * writes, and the next fx_step_async / fx_calc_dt forms dt from them (the
 * same value as a pass over the cells) unless the state was written by other
 * means since: fx_mesh_set_prim, fx_mesh_set_bf, fx_mesh_view_get and the
 * eSSPRK integrator drop the record. */
int fx_step_async(fx_mesh *m, double cfl, double dt_factor);
/* Same, then synchronises and returns the dt used. */
int fx_step(fx_mesh *m, double cfl, double dt_factor, double *dt_used);
/* One step of the mesh's integrator with the caller's dt: prj_timeint_step
 * (src/prj_timeint.c:3221-3283) called with dt_step, for a main loop that
 * forms dt on the host as src/main.c:510-545 does -- fx_calc_dt, the growth
 * limit, then dt clipped to the next output/event time and to t_end.  A later
 * fx_step takes this dt as its previous one.  Synchronises; dt must be finite
 * and > 0 (FX_EINVAL otherwise). */
int fx_step_dt(fx_mesh *m, double dt);
/* Time integrator of fx_step / fx_step_async (TIME_INTEGRATION of
 * src/prj_defs.h:136-176).  FX_RK2 (default): the two stages above.
 * FX_ESSPRK10_4: Ketcheson's eSSPRK(10,4), the reference's default, as in
 * prj_timeint_eSSPRK10_4_advance (src/prj_timeint.c:2661-2730) -- CFL dt x 6,
 * ten forward-Euler substeps of dt/6 and three convex blends of the conserved
 * state; on MHD meshes the face fields are saved and blended too, before the
 * cells, whose field is then the face average of the blended faces
 * (src/prj_timeint.c:2287-2417).  fx_stage keeps its RK2 meaning. */
#define FX_RK2 0
#define FX_ESSPRK10_4 1
int fx_mesh_set_integrator(fx_mesh *m, int integrator);
/* Riemann solver of a hydro mesh's stage.  FX_RIEMANN_HLLC (default): the
 * reference's prj_riemann_hllc with the theta limiter (src/prj_riemann.c:
 * 1494-1599).  FX_RIEMANN_HLLE: HLL over the same Davis wave-speed bounds and
 * face states, in the form of the reference's HLL fallback
 * (src/prj_riemann.c:393-407); the reference has no hydro HLLE, so this
 * option has no reference counterpart.  MHD meshes take HLLD only. */
/* Launch schedule of an MHD mesh's stage (no reference counterpart; every
 * schedule gives the same bits).  ct_split: CT workgroups per block, 1, 2 or
 * 4 (each marches 16/ct_split of the block's x-planes), 0 automatic (more
 * where the blocks would not fill the chip's CT slots).  side_stream: 1 runs
 * the ghost-layer face kernel on a second stream beside the flux kernel, 0 in
 * order on the mesh's stream, -1 automatic (beside it where the flux kernel
 * fits the chip in one round).  Synchronises the mesh's streams first. */
int fx_mesh_set_mhd_schedule(fx_mesh *m, int ct_split, int side_stream);
#define FX_RIEMANN_HLLC 0
#define FX_RIEMANN_HLLE 1
int fx_mesh_set_riemann(fx_mesh *m, int solver);
/* Waits for the mesh's stream and reports any recorded device error. */
int fx_sync(fx_mesh *m);
/* dt of the last device step (synchronises). */
int fx_last_dt(fx_mesh *m, double *dt);

/* Event timing of the stage kernels (the work of fx_stage after the domain
 * ghost fill), recorded on the mesh's stream.  fx_mesh_set_timing(m, 1)
 * clears earlier records and starts recording; fx_mesh_kernel_time sums the
 * recorded launches of `stage` (1, 2, or 0 for both), synchronising first.
 * Measurement support only; no counterpart in the reference. */
int fx_mesh_set_timing(fx_mesh *m, int on);
int fx_mesh_kernel_time(fx_mesh *m, int stage, double *total_ms, int *launches);

/* Device addresses of the state, for zero-copy interop: slot s, variable v
 * of the cell arrays starts at W[s] + v*vstride; cell (x,y,z) at
 * origin + x*sx + y*sy + z.  Face component d of slot s: bf[s] + d*vstride.
 * Taking a view drops the recorded CFL maxima (see fx_step_async): a caller
 * that writes the state through a view after a later step calls
 * fx_mesh_view_get again before the next step. */
typedef struct fx_mesh_view {
    double *W[2];
    double *bf[3];
    long long vstride, sx, sy, origin;
    int n[3];
    int nh;
} fx_mesh_view;
int fx_mesh_view_get(fx_mesh *m, fx_mesh_view *out);

/* ---- per-face / per-edge entry points on device arrays (SoA: component c
 *      of item i at p[c*n + i]); stream may be NULL -------------------------- */
/* prj_reconstruct_mc_face_slope (src/prj_reconstruct.h:355-375), the MC
 * limited slope of the 3-cell stencil (qm, q0, qp), over n stencils of device
 * arrays; the face values are q0 +/- 0.5 * slope. */
int fx_reconstruct_mc_slope_batch(int n, const double *qm, const double *q0, const double *qp,
    double *slope, void *stream);
/* prj_riemann_hllc (src/prj_riemann.h:7-10): WL/WR 6 rotated prims; flux 6,
 * v_face 3. */
int fx_riemann_hllc_batch(int n, const double *WL, const double *WR, const double *pL,
    const double *pR, const double *gL, const double *gR, const double *du,
    const double *dv, const double *dw, double *flux, double *v_face, void *stream);
/* prj_riemann_hlld (src/prj_riemann.h:12-15): WL/WR 9 rotated prims (slot 6
 * is overwritten by bn), flux 9, v_face 3, Bv1, Bv2; status[i] = 0 or the
 * reference abort site (>0). */
int fx_riemann_hlld_batch(int n, const double *WL, const double *WR, const double *pL,
    const double *pR, const double *gL, const double *gR, const double *bn, double *flux,
    double *v_face, double *Bv1, double *Bv2, int *status, void *stream);
/* prj_mhd_emf_upwind without the resistive term (src/prj_mhd.h:18-20, with
 * eta = 0): emf_face/emf_cell/v_norm 4 components each. */
int fx_emf_upwind_batch(int n, const double *emf_face, const double *emf_cell,
    const double *v_norm, double *emf, void *stream);
/* prj_mhd_emf_upwind (src/prj_mhd.h:18-20, src/prj_mhd.c:627-718) with the
 * resistive term, over n edges of one block in the reference's block layouts.
 * Item m is the edge (dir, i, j, k) = edge[4m .. 4m+3] (block indices, as the
 * reference's); emf_face/emf_cell/v_norm 4 components each (device, SoA);
 * eta[m] the edge's resistivity (mhd_eta plus the mask maximum, as
 * src/prj_timeint.c:404-405 passes it); bf[3] (a host array of device
 * pointers) the block's face fields of the stage (FACE_IDX), read only where
 * eta != 0; dx[3] (host) the block's cell widths.  emf[m] receives the value
 * and, where block_emf (host array of device pointers) is non-NULL,
 * block_emf[dir][EDGE_IDX(dir, i, j, k)] too, as the reference stores it.
 * status[m] = 0, or where the reference aborts: 37 edge outside the block's
 * storage, 31 non-finite input, 38 missing face field for eta != 0, 36
 * non-finite curl B, 39 non-finite result; 35 a curl read outside the face
 * storage (an out-of-bounds read in the reference).  emf[m] is NaN where
 * status != 0.
 * Items run concurrently: where one edge appears more than once, its
 * block_emf entry holds the value of one of them, unspecified which. */
int fx_emf_upwind_block_batch(int n, const int *edge, const double *emf_face,
    const double *emf_cell, const double *v_norm, const double *eta, const double *const *bf,
    const double *dx, double *const *block_emf, double *emf, int *status, void *stream);

/* ---- per-block entry points in the reference's own block layouts -------- *
 * A block is the reference's 20^3 box (16^3 active cells + 2 ghost layers,
 * src/prj_defs.h:383-459):
 *   W          prims of one stage slot: variable v at W[v*var_stride + LIDX]
 *              (W_mhd slot s: W_mhd + s*8000, var_stride = NSTAGES*8000 = 16000)
 *   eosvar     EIDX AoSoA (P, T, gamma)
 *   bf[3]      MHD face fields of the stage slot, FACE_IDX
 *   flux[3]    out, VIDX AoSoA, 6 (hydro) or 9 (MHD) conserved variables
 *   v_riemann  out, VRIDX AoSoA; Bv1/Bv2 out, IDX (MHD)
 * Faces written are the reference's (src/prj_flux.c:2221-2262): normal index
 * 0..16; transverse 0..15 (hydro) or -1..16 (MHD); nothing else is touched.
 * Hydro flux and v_riemann arrays on 16-byte boundaries (as hipMalloc and
 * malloc return them) are written with paired 16-byte stores; others with one
 * store per value, same results. */
typedef struct fx_block_io {
    const double *W;
    long long var_stride;
    const double *eosvar;
    const double *bf[3];
    double *flux[3];
    double *v_riemann[3];
    double *Bv1[3];
    double *Bv2[3];
} fx_block_io;
/* prj_flux_update (src/prj_flux.h:4-5) over n blocks whose arrays are device
 * memory; `blocks` is a device array of n descriptors.  One launch on
 * `stream`, no synchronisation.  MHD: status (device int[5], zeroed by the
 * caller) receives the first HLLD failure {code, dir, i, j, k}. */
int fx_flux_update_blocks(int n, const fx_block_io *blocks, int mhd, int *status, void *stream);
/* One prj_flux_update call on host arrays: copies the block into device memory
 * the library keeps (allocated on the first call on a device, reused after),
 * runs the same kernel with n = 1 on a stream of its own, copies the outputs
 * back and waits for that stream only; an HLLD failure returns FX_EPHYS with
 * the reference's message.  Calls from several host threads are serialised. */
int fx_flux_update_block(const fx_block_io *block, int mhd);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc): prj_flux_update on
 * n host-array blocks, as fx_flux_update_block, without allocating anything or
 * creating a stream.  The blocks are staged through a fixed device array of
 * 29,264,128 bytes that the library holds on each GPU it runs on until the
 * process ends; n outside 1..FX_NOALLOC_MAX_BLOCKS returns FX_EINVAL.  Runs on
 * `stream` (NULL: the null stream) and waits for it.  Every call on one GPU
 * shares that array, which keeps the last call's contents (nothing is read
 * from it that the call has not written); calls from several host threads are
 * serialised by a lock, as overlapping calls would race on the array. */
#define FX_NOALLOC_MAX_BLOCKS 8
int fx_flux_update_block_noalloc(int n, const fx_block_io *blocks, int mhd, void *stream);
/* prj_mhd_bf2bc / prj_mhd_bf2bc_all (src/prj_mhd.h:7-13, src/prj_mhd.c:555-625):
 * the cell-centred field W[B1..B3] of every cell of a block, ghosts included,
 * from the face average of the slot's face fields.  The _mesh variants differ
 * only in the GR metric factor, which is 1 on this path. */
typedef struct fx_bf2bc_io {
    double *W;              /* prims of the stage slot: variable v at W[v*var_stride + LIDX] */
    long long var_stride;
    const double *bf[3];    /* the slot's face fields (FACE_IDX) */
    const int *done;        /* block->cell_derived_done (IDX): flagged cells are skipped
                             * (prj_mhd_bf2bc); NULL: every cell (prj_mhd_bf2bc_all) */
} fx_bf2bc_io;
/* n blocks whose arrays are device memory; `blocks` is a device array of n
 * descriptors.  One launch on `stream`, no synchronisation.  status (device
 * int[5], zeroed by the caller, or NULL) receives the first non-finite cell
 * field {1, block, i, j, k}, where the reference aborts with
 * "prj_mhd_bf2bc: non-finite cell-centered magnetic field" (that cell is not
 * written). */
int fx_bf2bc_blocks(int n, const fx_bf2bc_io *blocks, int *status, void *stream);
/* One call on host arrays: stages the block through device memory the library
 * keeps, runs the same kernel with n = 1 on a stream of its own and waits for
 * it; a non-finite field returns FX_EPHYS with the reference's message. */
int fx_bf2bc_block(const fx_bf2bc_io *block);
/* prj_flux_div (src/prj_flux.h:10) at every active cell of one block, on host
 * VIDX flux arrays of nvar variables: fluxdiv[v][i][j][k] (16^3 per v). */
int fx_flux_div_block(double *const *flux, const double *area, double vol, int nvar,
    double *fluxdiv);

/* ---- coarse-fine corrections of a refined mesh (per-block layouts) ------ *
 * The local phase of prj_riemann_flux_send (src/prj_riemann.h:25,
 * src/prj_riemann.c:1701-1768) and prj_mhd_emf_send (src/prj_mhd.h,
 * src/prj_mhd.c:779-1029) on device-resident block arrays.
 *
 * The caller describes the mesh once per topology: a table of blocks (index
 * = the reference's block id; geometry of every block, arrays of the local
 * ones) and the neighbour slots of its local blocks whose neighbour is
 * coarser, in the reference's loop order (blocks by id, slots 0..55).
 * fx_cf_plan_create turns them into device correction lists; each stage
 * then calls fx_cf_flux_send after the flux update and fx_cf_emf_send after
 * the edge EMFs, both stream-ordered, neither synchronising.  The results,
 * the order of overlapping writes included, are the reference's. */
typedef struct fx_cf_block {
    double xmin[3], xmax[3], dx[3];  /* prj_block xmin, xmax, dx */
    int local;                       /* id >= 0, active, owned by this rank */
    double *flux[3];                 /* device flux[axis], nvar x 8000 (VIDX), or NULL */
    double *emf[3];                  /* device emf[dir], 8820 (EDGE_IDX), or NULL */
    int *edge_fidelity[3];           /* device edge_fidelity[dir], or NULL (not kept) */
} fx_cf_block;

typedef struct fx_cf_link {          /* slot n of a local block (src/prj_types.h:77-105) */
    int block;                       /* the block's table index */
    int neighbor;                    /* slot->id */
    int rel_level;                   /* slot->rel_level; >= 0 is skipped */
    int type;                        /* slot->type: 1 face, 2 edge, 3 corner */
    int neighbor_local;              /* slot->rank is this rank */
    int recv_start[3];               /* slot->recv_loc_start */
    int recv_end[3];                 /* slot->recv_loc_end */
    int send_start[3];               /* slot->send_loc_start */
} fx_cf_link;

typedef struct fx_cf_plan fx_cf_plan;
/* nvar: flux components (NVAR_CONS: 6 hydro, 9 MHD).  A part whose arrays
 * are missing stays unavailable: its send call then fails with FX_EINVAL. */
int fx_cf_plan_create(int nblocks, const fx_cf_block *blocks, int nlinks,
    const fx_cf_link *links, int nvar, fx_cf_plan **out);
/* info[4] = flux records, flux launch levels, EMF records, EMF launch levels */
int fx_cf_plan_info(const fx_cf_plan *p, int *info);
int fx_cf_plan_destroy(fx_cf_plan *p);
/* prj_riemann_flux_send, local phase: coarse face fluxes from the fine faces */
int fx_cf_flux_send(fx_cf_plan *p, void *stream);
/* prj_mhd_emf_send: edge fidelity reset, coarse edge EMFs from the fine
 * edges.  status (device int[4], zeroed by the caller, or NULL) receives the
 * first non-finite restricted value {1, block, dir, edge index}, where the
 * reference aborts with "prj_mhd_emf_send: invalid emf write". */
int fx_cf_emf_send(fx_cf_plan *p, int *status, void *stream);

/* ---- ragged ensembles of independent hydro meshes (batched stage) ------- *
 * No counterpart in the reference, which runs one problem per process: an
 * ensemble advances n independent periodic/outflow/reflect hydro meshes of
 * different sizes and gammas together, each with its own CFL dt.  Schedule
 * mode 0 launches every instance's kernels separately (the per-instance
 * baseline); mode 1 runs each stage as one batched launch over a table of
 * (instance, block) entries -- one workgroup per entry, or persistent_wgs
 * persistent workgroups draining the table through atomic queues (per XCD when
 * xcd != 0); table 0 is instance-major, 1 interleaves the instances.  With the
 * graph on, a whole step is captured once into a hipGraph and replayed. */
typedef struct fx_ensemble fx_ensemble;
int fx_ensemble_create(int n, const fx_mesh_desc *descs, fx_ensemble **out);
int fx_ensemble_destroy(fx_ensemble *e);
int fx_ensemble_set_prim(fx_ensemble *e, int inst, const double *W);
int fx_ensemble_get_prim(fx_ensemble *e, int inst, int slot, double *W);
int fx_ensemble_set_schedule(fx_ensemble *e, int mode, int persistent_wgs, int xcd, int table);
int fx_ensemble_set_graph(fx_ensemble *e, int on);
/* One RK2 step of every instance (CFL dt on the device, growth limit, two
 * stages); does not synchronise. */
int fx_ensemble_step_async(fx_ensemble *e, double cfl, double dt_factor);
int fx_ensemble_sync(fx_ensemble *e);
int fx_ensemble_last_dt(fx_ensemble *e, double *dt);
int fx_ensemble_info(fx_ensemble *e, int *nblocks, long long *ncells);
/* Times `reps` steps of each candidate schedule (advancing the state) and
 * keeps the fastest; out[4] = mode, persistent_wgs, xcd, table. */
int fx_ensemble_autotune(fx_ensemble *e, double cfl, double dt_factor, int reps, int *out,
    double *best_ms);

/* ---- differentiable scheme and its discrete adjoint ---------------------- *
 * No counterpart in the reference (it has neither the van Albada limiter nor
 * HLLE): PLM with the smooth van Albada slope (regularisation va_eps2) + HLLE
 * with Davis speeds, RK2 with a fixed dt, periodic box, ideal gas.  Device
 * arrays in the compact layout W[var][X][Y][Z] (6 prims, Z fastest).
 * fx_adj_forward keeps a checkpoint every checkpoint_every steps;
 * fx_adj_backward maps the cotangent of W_T to that of W_0 for the last
 * forward, recomputing one segment at a time.  Stream-ordered on `stream`. */
typedef struct fx_adj fx_adj;
int fx_adj_create(const int *n, const double *xlim, double gamma, double va_eps2,
    int checkpoint_every, fx_adj **out);
int fx_adj_destroy(fx_adj *a);
int fx_adj_forward(fx_adj *a, const double *W0, double *WT, int nsteps, double dt, void *stream);
int fx_adj_backward(fx_adj *a, const double *gWT, double *gW0, void *stream);

const char *fx_last_error(void);
/* The reference's abort message for a per-item status code of the batched
 * entry points (HLLD 11-20, EMF 31 and 35-39); "ok" for 0. */
const char *fx_status_message(int status);
/* Library build string (arch, flags). */
const char *fx_version(void);

#ifdef __cplusplus
}
#endif

#endif
