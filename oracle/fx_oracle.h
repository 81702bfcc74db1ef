/*
 * CPU restatement of FornaxAMR's per-block Godunov RK2 stage (hydro and MHD).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the "port" CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product library never links it.
 *
 * The restatement works on a uniform single-level mesh held as "global" arrays
 * over active cells (cell arrays [var][X][Y][Z], Z fastest; face arrays
 * Bx[X+1][Y][Z], By[X][Y+1][Z], Bz[X][Y][Z+1]).  Ghost values that the
 * reference fills before each stage (src/prj_boundary.c:1150-1241,
 * src/prj_eos.c:1109-1162, src/prj_mhd.c:555-603) are produced by index
 * mapping; on a single-level mesh the two are the same values.
 */
#ifndef FX_ORACLE_H
#define FX_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FXO_BC_OUTFLOW = 0, FXO_BC_REFLECT = 1, FXO_BC_PERIODIC = 3 };

typedef struct fxo_grid {
    int n[3];        /* active cells per axis */
    int bc[6];       /* x1 inner, x1 outer, x2 inner, x2 outer, x3 inner, x3 outer */
    double dx[3];
    double area[3];
    double vol;
    double gamma;    /* ideal-gas adiabatic index */
    int mhd;         /* 0: NHYDRO = 6, 1: NHYDRO = 9 with CT */
    double eta;      /* uniform CT resistivity (mesh->mhd_eta) */
} fxo_grid;

int fxo_grid_init(fxo_grid *g, int nx, int ny, int nz, const double *xlim,
    const int *bc, double gamma, int mhd, double eta);

double fxo_calc_dt(const fxo_grid *g, const double *W, double cfl);

/* One RK2 stage.  stage 1 reads W0/bf0 and writes Wout/bfout (slot 1);
 * stage 2 reads W1/bf1 for the fluxes, W0/bf0 for the blend, and writes
 * Wout/bfout (slot 0).  bf* and emf_out may be NULL for hydro; emf_out may be
 * NULL always.  Wout must not alias W0 or W1; bfout must not alias bf1.
 * Returns 0, or a nonzero error code with a message in fxo_last_error(). */
int fxo_stage(const fxo_grid *g, int stage, double dt,
    const double *W0, const double *W1, double *Wout,
    const double *const *bf0, const double *const *bf1, double *const *bfout,
    double *const *emf_out);

/* Cell-centred B of active cells from face fields (bf2bc). */
void fxo_bf2bc(const fxo_grid *g, const double *const *bf, double *W);
/* eSSPRK convex blend of the conserved state (src/prj_timeint.c:2382-2417),
 * hydro: W <- prim(w_s u(W_saved) + (1 - w_s) u(W)) */
int fxo_blend(const fxo_grid *g, double *W, const double *Wsaved, double saved_weight);
/* the same with face fields (src/prj_timeint.c:2347-2417, MHD): faces
 * bf <- w_s bf_saved + (1 - w_s) bf, then cells as above with the cell field
 * from the blended faces */
int fxo_blend_mhd(const fxo_grid *g, double *W, const double *Wsaved, double *const *bf,
    const double *const *bfsaved, double saved_weight);

/* Per-face and per-edge entry points with the reference's argument meaning. */
double fxo_mc_slope(double qm, double q0, double qp);
void fxo_hllc(const double *WL, const double *WR, double pL, double pR,
    double gL, double gR, double *flux, double *v_face,
    double du, double dv, double dw);
/* HLLE over the reference's HLLC wave-speed bounds and face states (not a
 * reference solver); v_face from the HLL average state */
void fxo_hlle(const double *WL, const double *WR, double pL, double pR,
    double gL, double gR, double *flux, double *v_face);
/* Riemann solver of hydro stages: 0 HLLC (default), 1 HLLE */
void fxo_set_riemann(int solver);
int fxo_hlld(const double *WL, const double *WR, double pL, double pR,
    double gL, double gR, double bn, double *flux, double *v_face,
    double *Bv1, double *Bv2);
int fxo_emf_upwind(const double *emf_face, const double *emf_cell,
    const double *v_norm, double *emf);
/* prj_mhd_emf_upwind with the resistive term on one block's FACE_IDX arrays */
int fxo_emf_upwind_block(int dir, int i, int j, int k, const double *emf_face,
    const double *emf_cell, const double *v_norm, const double *const *bf, const double *dx,
    double eta, double *result);

/* log of the resistivity mask: 0 = the host libm's (as the reference build),
 * 1 = correctly rounded (as the GPU path) */
void fxo_set_log_mode(int correctly_rounded);
double fxo_log_cr(double x);

const char *fxo_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
