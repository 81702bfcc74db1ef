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
/* Stream all launches of this mesh go to (a hipStream_t); NULL = default. */
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
/* CFL step of slot 0 (prj_timeint_calc_dt, src/prj_timeint.c:2144-2256). */
int fx_calc_dt(fx_mesh *m, double cfl, double *dt);
/* One RK2 stage over all blocks: the body of prj_timeint_stage1/2
 * (src/prj_timeint.c:2732-2973) -- domain ghost fill, reconstruction,
 * Riemann fluxes, flux divergence, CT EMF and face update (MHD), per-cell
 * update.  stage 1: slot 0 -> slot 1; stage 2: slots 0,1 -> slot 0. */
int fx_stage(fx_mesh *m, int stage, double dt);
/* One main-loop step with dt kept on the device: CFL dt, growth limit
 * dt <= dt_factor * dt_prev (src/main.c:510-522), stage 1, stage 2.
 * Does not synchronise. */
int fx_step_async(fx_mesh *m, double cfl, double dt_factor);
/* Same, then synchronises and returns the dt used. */
int fx_step(fx_mesh *m, double cfl, double dt_factor, double *dt_used);
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
 * origin + x*sx + y*sy + z.  Face component d of slot s: bf[s] + d*vstride. */
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
/* prj_mhd_emf_upwind without the resistive term (src/prj_mhd.h:18-20):
 * emf_face/emf_cell/v_norm 4 components each. */
int fx_emf_upwind_batch(int n, const double *emf_face, const double *emf_cell,
    const double *v_norm, double *emf, void *stream);

const char *fx_last_error(void);
/* Library build string (arch, flags). */
const char *fx_version(void);

#ifdef __cplusplus
}
#endif

#endif
