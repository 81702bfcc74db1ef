/*
 * Driver harness for a serial build of the reference FornaxAMR sources.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is compiled together with the
 * reference's own C99 translation units (everything under src/ except main.c
 * and prj_io.c, plus rk_tableau/) by oracle/Makefile into
 * oracle/_ref/libprj_ref_{hydro,mhd}.so.  Nothing in the product links it;
 * tests and the CPU-baseline leg of bench.py load it through ctypes to pin
 * the C restatement in oracle/fx_oracle.c and to time the reference CPU path.
 *
 * It plays the role of src/main.c for a uniform single-level mesh:
 *   - mesh set-up: prj_mesh_init + prj_mpi_decompose/prepare
 *     (src/main.c:354-370, problems/prj_problem_sedov.c:329-344);
 *   - initial fill: eos_fill_active_cells, boundary fill, eos_fill_mesh
 *     (src/main.c:437-447);
 *   - the step: prj_timeint_calc_dt + dt_factor limit, prj_timeint_step
 *     (src/main.c:510-545);
 * and supplies the problem-side ideal-gas EOS plug-in that a USER-EOS build
 * requires (src/prj_eos.h:78-85), with gamma chosen at run time instead of at
 * compile time (problems/prj_problem_ideal_eos.h:6-32 fixes it per problem).
 *
 * Data exchanged with the caller uses "global" arrays over the active cells of
 * the whole mesh: cell arrays are [var][X][Y][Z] with Z fastest; face arrays
 * are Bx[X+1][Y][Z], By[X][Y+1][Z], Bz[X][Y][Z+1].
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "prj.h"

#define REFH_ENERGY_SCALE 0.95655684e18

static prj_sim g_sim;
static prj_mpi g_mpi;
static double g_gamma = 5.0 / 3.0;
static int g_nb[3];
static int g_ready = 0;

/* Ideal-gas EOS plug-in (USER provider).  Same closed forms as the reference's
 * problem-side ideal EOS: p = (g-1) rho e, T = (g-1) e / scale. */
void eos_rty(double rho, double temperature, double ye,
    double *eint, double *pressure, double *gamma, double *eta,
    double deint[3], double dpressure[3])
{
    const double g = g_gamma;
    (void)ye;
    *eint = REFH_ENERGY_SCALE * temperature / (g - 1.0);
    *pressure = rho * REFH_ENERGY_SCALE * temperature;
    *gamma = g;
    *eta = 0.0;
    deint[0] = 0.0;
    deint[1] = REFH_ENERGY_SCALE / (g - 1.0);
    deint[2] = 0.0;
    dpressure[0] = REFH_ENERGY_SCALE * temperature;
    dpressure[1] = rho * REFH_ENERGY_SCALE;
    dpressure[2] = 0.0;
}

void eos_rey(double rho, double eint, double ye,
    double *temperature, double *pressure, double *gamma)
{
    const double g = g_gamma;
    (void)ye;
    *temperature = (g - 1.0) * eint / REFH_ENERGY_SCALE;
    *pressure = (g - 1.0) * rho * eint;
    *gamma = g;
}

int refh_nhydro(void) { return PRJ_NHYDRO; }
int refh_mhd(void) { return PRJ_MHD; }
int refh_block_size(void) { return PRJ_BLOCK_SIZE; }

static prj_block *refh_block(int bx, int by, int bz)
{
    int id = (bx * g_nb[1] + by) * g_nb[2] + bz;
    return &g_sim.mesh.blocks[id];
}

static void refh_free(void)
{
    if (g_ready) {
        prj_mesh_destroy(&g_sim.mesh);
        g_ready = 0;
    }
}

/* bc[6] = {x1 inner, x1 outer, x2 inner, x2 outer, x3 inner, x3 outer} with the
 * reference codes 0 outflow, 1 reflect, 3 periodic. */
int refh_init(int nbx, int nby, int nbz, const double *xlim, const int *bc,
    double gamma, double mhd_eta)
{
    int axis;
    int bidx;

    refh_free();
    memset(&g_sim, 0, sizeof(g_sim));
    memset(&g_mpi, 0, sizeof(g_mpi));
    g_mpi.rank = 0;
    g_mpi.totrank = 1;
    g_gamma = gamma;
    g_nb[0] = nbx;
    g_nb[1] = nby;
    g_nb[2] = nbz;

    g_sim.coord.x1min = xlim[0];
    g_sim.coord.x1max = xlim[1];
    g_sim.coord.x2min = xlim[2];
    g_sim.coord.x2max = xlim[3];
    g_sim.coord.x3min = xlim[4];
    g_sim.coord.x3max = xlim[5];
    g_sim.bc.bc_x1_inner = bc[0];
    g_sim.bc.bc_x1_outer = bc[1];
    g_sim.bc.bc_x2_inner = bc[2];
    g_sim.bc.bc_x2_outer = bc[3];
    g_sim.bc.bc_x3_inner = bc[4];
    g_sim.bc.bc_x3_outer = bc[5];
    for (axis = 0; axis < 3; ++axis) {
        int inner_p = bc[2 * axis] == PRJ_BC_PERIODIC;
        int outer_p = bc[2 * axis + 1] == PRJ_BC_PERIODIC;
        if (inner_p != outer_p) {
            return 1;
        }
        g_sim.mesh.periodic[axis] = inner_p;
    }
    g_sim.mesh.max_blocks = nbx * nby * nbz;
    g_sim.mesh.mhd_eta = mhd_eta;
    g_sim.dt_factor = 1.2;
    if (prj_mesh_init(&g_sim.mesh, nbx, nby, nbz, 0, &g_sim.coord, 0) != 0) {
        return 2;
    }
    prj_mpi_decompose(&g_sim.mesh, &g_mpi);
    prj_mpi_prepare(&g_sim.mesh, &g_mpi);
#if PRJ_MHD
    g_sim.mhd_init_type = PRJ_MHD_INIT_UNIFORM;
    g_sim.mhd_B_norm = 0.0;
    g_sim.mhd_B_scale = 1.0;
    prj_mhd_init(&g_sim, &g_mpi);
#endif
    for (bidx = 0; bidx < g_sim.mesh.nblocks; ++bidx) {
        prj_block *b = &g_sim.mesh.blocks[bidx];
        int bxi = (int)floor((b->xmin[0] - xlim[0]) / (b->dx[0] * PRJ_BLOCK_SIZE) + 0.5);
        int byi = (int)floor((b->xmin[1] - xlim[2]) / (b->dx[1] * PRJ_BLOCK_SIZE) + 0.5);
        int bzi = (int)floor((b->xmin[2] - xlim[4]) / (b->dx[2] * PRJ_BLOCK_SIZE) + 0.5);
        if ((bxi * nby + byi) * nbz + bzi != b->id) {
            return 3;
        }
    }
    g_ready = 1;
    return 0;
}

void refh_destroy(void)
{
    refh_free();
}

void refh_geometry(double *dx, double *area, double *vol)
{
    const prj_block *b = &g_sim.mesh.blocks[0];
    int d;
    for (d = 0; d < 3; ++d) {
        dx[d] = b->dx[d];
        area[d] = b->area[d];
    }
    *vol = b->vol;
}

#define NXG (g_nb[0] * PRJ_BLOCK_SIZE)
#define NYG (g_nb[1] * PRJ_BLOCK_SIZE)
#define NZG (g_nb[2] * PRJ_BLOCK_SIZE)

static size_t refh_gidx(int v, int x, int y, int z)
{
    return (((size_t)v * NXG + (size_t)x) * NYG + (size_t)y) * NZG + (size_t)z;
}

/* Active-cell primitives for stage slot `slot` (0 or 1). */
int refh_set_prim(int slot, const double *W)
{
    int bx, by, bz, v, i, j, k;
    if (!g_ready) return 1;
    for (bx = 0; bx < g_nb[0]; ++bx)
    for (by = 0; by < g_nb[1]; ++by)
    for (bz = 0; bz < g_nb[2]; ++bz) {
        prj_block *b = refh_block(bx, by, bz);
        double *Ws = prj_block_prim_stage(b, slot);
        for (v = 0; v < PRJ_NHYDRO; ++v)
        for (i = 0; i < PRJ_BLOCK_SIZE; ++i)
        for (j = 0; j < PRJ_BLOCK_SIZE; ++j)
        for (k = 0; k < PRJ_BLOCK_SIZE; ++k) {
            Ws[WIDX(v, i, j, k)] = W[refh_gidx(v, bx * PRJ_BLOCK_SIZE + i,
                by * PRJ_BLOCK_SIZE + j, bz * PRJ_BLOCK_SIZE + k)];
        }
    }
    return 0;
}

int refh_get_prim(int slot, double *W)
{
    int bx, by, bz, v, i, j, k;
    if (!g_ready) return 1;
    for (bx = 0; bx < g_nb[0]; ++bx)
    for (by = 0; by < g_nb[1]; ++by)
    for (bz = 0; bz < g_nb[2]; ++bz) {
        prj_block *b = refh_block(bx, by, bz);
        const double *Ws = prj_block_prim_stage(b, slot);
        for (v = 0; v < PRJ_NHYDRO; ++v)
        for (i = 0; i < PRJ_BLOCK_SIZE; ++i)
        for (j = 0; j < PRJ_BLOCK_SIZE; ++j)
        for (k = 0; k < PRJ_BLOCK_SIZE; ++k) {
            W[refh_gidx(v, bx * PRJ_BLOCK_SIZE + i, by * PRJ_BLOCK_SIZE + j,
                bz * PRJ_BLOCK_SIZE + k)] = Ws[WIDX(v, i, j, k)];
        }
    }
    return 0;
}

/* Full ghost-padded prim array of one block, [var][20^3] in LIDX order. */
int refh_get_block_prim_padded(int id, int slot, double *out)
{
    int v, c;
    const prj_block *b;
    if (!g_ready || id < 0 || id >= g_sim.mesh.nblocks) return 1;
    b = &g_sim.mesh.blocks[id];
    for (v = 0; v < PRJ_NHYDRO; ++v) {
        const double *Ws = prj_block_prim_stage((prj_block *)b, slot);
        for (c = 0; c < PRJ_BLOCK_NCELLS; ++c) {
            out[(size_t)v * PRJ_BLOCK_NCELLS + c] = Ws[(size_t)v * PRJ_BLOCK_NSTAGES * PRJ_BLOCK_NCELLS + c];
        }
    }
    return 0;
}

int refh_get_eosvar(double *P, double *T, double *G)
{
    int bx, by, bz, i, j, k;
    if (!g_ready) return 1;
    for (bx = 0; bx < g_nb[0]; ++bx)
    for (by = 0; by < g_nb[1]; ++by)
    for (bz = 0; bz < g_nb[2]; ++bz) {
        const prj_block *b = refh_block(bx, by, bz);
        for (i = 0; i < PRJ_BLOCK_SIZE; ++i)
        for (j = 0; j < PRJ_BLOCK_SIZE; ++j)
        for (k = 0; k < PRJ_BLOCK_SIZE; ++k) {
            size_t g = refh_gidx(0, bx * PRJ_BLOCK_SIZE + i, by * PRJ_BLOCK_SIZE + j,
                bz * PRJ_BLOCK_SIZE + k);
            P[g] = b->eosvar[EIDX(PRJ_EOSVAR_PRESSURE, i, j, k)];
            T[g] = b->eosvar[EIDX(PRJ_EOSVAR_TEMPERATURE, i, j, k)];
            G[g] = b->eosvar[EIDX(PRJ_EOSVAR_GAMMA, i, j, k)];
        }
    }
    return 0;
}

#if PRJ_MHD
static int refh_face_max(int dir, int axis)
{
    return dir == axis ? PRJ_BLOCK_SIZE : PRJ_BLOCK_SIZE - 1;
}

static size_t refh_fidx(int dir, int x, int y, int z)
{
    size_t nx = (size_t)NXG + (dir == 0);
    size_t ny = (size_t)NYG + (dir == 1);
    size_t nz = (size_t)NZG + (dir == 2);
    (void)nx;
    return ((size_t)x * ny + (size_t)y) * nz + (size_t)z;
}

static int refh_copy_bf(int slot, double *f[3], int to_block)
{
    int bx, by, bz, dir, i, j, k;
    if (!g_ready) return 1;
    for (bx = 0; bx < g_nb[0]; ++bx)
    for (by = 0; by < g_nb[1]; ++by)
    for (bz = 0; bz < g_nb[2]; ++bz) {
        prj_block *b = refh_block(bx, by, bz);
        for (dir = 0; dir < 3; ++dir) {
            double *bf = prj_block_bf_stage(b, dir, slot);
            for (i = 0; i <= refh_face_max(dir, 0); ++i)
            for (j = 0; j <= refh_face_max(dir, 1); ++j)
            for (k = 0; k <= refh_face_max(dir, 2); ++k) {
                size_t g = refh_fidx(dir, bx * PRJ_BLOCK_SIZE + i,
                    by * PRJ_BLOCK_SIZE + j, bz * PRJ_BLOCK_SIZE + k);
                if (to_block) {
                    bf[FACE_IDX(dir, i, j, k)] = f[dir][g];
                } else {
                    f[dir][g] = bf[FACE_IDX(dir, i, j, k)];
                }
            }
        }
    }
    return 0;
}

int refh_set_bf(int slot, double *bx, double *by, double *bz)
{
    double *f[3];
    f[0] = bx; f[1] = by; f[2] = bz;
    return refh_copy_bf(slot, f, 1);
}

int refh_get_bf(int slot, double *bx, double *by, double *bz)
{
    double *f[3];
    f[0] = bx; f[1] = by; f[2] = bz;
    return refh_copy_bf(slot, f, 0);
}

/* Edge EMFs of the last stage, global layout: E_d has +1 extent across d. */
int refh_get_emf(double *ex, double *ey, double *ez)
{
    int bx, by, bz, dir, i, j, k;
    double *f[3];
    f[0] = ex; f[1] = ey; f[2] = ez;
    if (!g_ready) return 1;
    for (bx = 0; bx < g_nb[0]; ++bx)
    for (by = 0; by < g_nb[1]; ++by)
    for (bz = 0; bz < g_nb[2]; ++bz) {
        prj_block *b = refh_block(bx, by, bz);
        for (dir = 0; dir < 3; ++dir) {
            int imax = dir == 0 ? PRJ_BLOCK_SIZE - 1 : PRJ_BLOCK_SIZE;
            int jmax = dir == 1 ? PRJ_BLOCK_SIZE - 1 : PRJ_BLOCK_SIZE;
            int kmax = dir == 2 ? PRJ_BLOCK_SIZE - 1 : PRJ_BLOCK_SIZE;
            size_t ny = (size_t)NYG + (dir != 1);
            size_t nz = (size_t)NZG + (dir != 2);
            for (i = 0; i <= imax; ++i)
            for (j = 0; j <= jmax; ++j)
            for (k = 0; k <= kmax; ++k) {
                size_t g = ((size_t)(bx * PRJ_BLOCK_SIZE + i) * ny +
                    (size_t)(by * PRJ_BLOCK_SIZE + j)) * nz + (size_t)(bz * PRJ_BLOCK_SIZE + k);
                f[dir][g] = b->emf[dir][EDGE_IDX(dir, i, j, k)];
            }
        }
    }
    return 0;
}
#endif

/* Initial fill, as src/main.c:437-447 does after the problem set-up. */
int refh_prepare(void)
{
    if (!g_ready) return 1;
#if PRJ_MHD
    {
        int bidx;
        for (bidx = 0; bidx < g_sim.mesh.nblocks; ++bidx) {
            prj_mhd_bf2bc_all_mesh(&g_sim.eos, &g_sim.mesh, &g_sim.mesh.blocks[bidx], 0);
        }
    }
#endif
    prj_eos_fill_active_cells(&g_sim.mesh, &g_sim.eos, &g_mpi, 1, PRJ_EOS_CTX_MAIN);
    prj_boundary_fill_ghosts_and_bf(&g_sim.mesh, &g_mpi, &g_sim.bc, 1, 0, &g_sim.eos, 0,
        &g_sim.rad, PRJ_BOUNDARY_TIMER_SCOPE_NONE);
    prj_eos_fill_mesh(&g_sim.mesh, &g_sim.eos, &g_mpi, 1, PRJ_EOS_CTX_MAIN);
    g_sim.dt = 0.0;
    g_sim.time = 0.0;
    return 0;
}

/* CFL time step of the current state (src/prj_timeint.c:2144-2256). */
double refh_calc_dt(double cfl)
{
    return prj_timeint_calc_dt(&g_sim.mesh, &g_sim.eos, &g_sim.rad, &g_mpi, cfl,
        &PRJ_TIMEINT_TABLEAU_NAME);
}

void refh_stage(int stage, double dt)
{
    double dt_src = 1.0e100;
    if (stage == 1) {
        prj_timeint_stage1(&g_sim.mesh, &g_sim.coord, &g_sim.bc, &g_sim.eos, &g_sim.rad,
            &g_sim.grav, &g_mpi, dt, &dt_src, 0);
    } else {
        prj_timeint_stage2(&g_sim.mesh, &g_sim.coord, &g_sim.bc, &g_sim.eos, &g_sim.rad,
            &g_sim.grav, &g_mpi, dt, &dt_src, 0);
    }
}

/* One main-loop step: dt from the CFL condition with the dt_factor growth
 * limit (src/main.c:510-522), then the RK2 step.  Returns the dt used. */
/* main.c's growth limit of sim.dt, then dt_step clipped to t_end
 * (src/main.c:517-535); returns dt_step */
static double refh_limit_and_clip(double dt_new, double dt_factor, double t_end)
{
    double dt_step;
    if (g_sim.dt > 0.0) {
        double dt_limit = dt_factor * g_sim.dt;
        g_sim.dt = dt_new <= dt_limit ? dt_new : dt_limit;
    } else {
        g_sim.dt = dt_new;
    }
    dt_step = g_sim.dt;
    if (g_sim.time + dt_step > t_end) {
        dt_step = t_end - g_sim.time;
    }
    return dt_step;
}

/* one main-loop step with the run ending at t_end (src/main.c:510-549) */
double refh_step_until(double cfl, double dt_factor, double t_end)
{
    double dt_src = 1.0e100;
    double dt_step = refh_limit_and_clip(refh_calc_dt(cfl), dt_factor, t_end);
    g_sim.mesh.time_seconds = g_sim.time;
    prj_timeint_step(&g_sim.mesh, &g_sim.coord, &g_sim.bc, &g_sim.eos, &g_sim.rad,
        &g_sim.grav, &g_mpi, &PRJ_TIMEINT_TABLEAU_NAME, dt_step, &dt_src, 0);
    g_sim.time += dt_step;
    return dt_step;
}

double refh_step(double cfl, double dt_factor)
{
    return refh_step_until(cfl, dt_factor, 1.0e300);
}

#ifdef REFH_FX_MESH
/* The mesh-level drop-in (fornaxamr_amd/shim/prj_timeint_fx.c) on this mesh:
 * attach after refh_prepare, step on the GPU, download into the blocks. */
void prj_fx_attach(prj_sim *sim);
double prj_fx_calc_dt(prj_sim *sim);
void prj_fx_step(prj_sim *sim, double dt_step);
void prj_fx_download(prj_sim *sim, prj_mpi *mpi);
void prj_fx_detach(void);

void refh_fx_attach(void) { prj_fx_attach(&g_sim); }

/* refh_step_until with the two reference calls replaced by the shim's */
double refh_fx_step_until(double cfl, double dt_factor, double t_end)
{
    double dt_step;
    g_sim.cfl = cfl;
    dt_step = refh_limit_and_clip(prj_fx_calc_dt(&g_sim), dt_factor, t_end);
    prj_fx_step(&g_sim, dt_step);
    g_sim.time += dt_step;
    return dt_step;
}

void refh_fx_download(void) { prj_fx_download(&g_sim, &g_mpi); }

void refh_fx_detach(void) { prj_fx_detach(); }
#endif

/* Per-edge EMF entry (src/prj_mhd.c:627-718) on a scratch block, so tests can
 * call the reference upwind average with arbitrary inputs. */
#if PRJ_MHD
double refh_emf_upwind(int dir, const double *emf_face, const double *emf_cell,
    const double *v_norm)
{
    prj_block *b = &g_sim.mesh.blocks[0];
    return prj_mhd_emf_upwind(b, dir, 0, 0, 0, emf_face, emf_cell, v_norm, 0, 0.0);
}

#endif

/* The reference's MC face slope (a static inline function of
 * src/prj_reconstruct.h) over n stencils. */
void refh_mc_face_slope(int n, const double *qm, const double *q0, const double *qp,
    double *slope)
{
    int i;
    for (i = 0; i < n; ++i) {
        slope[i] = prj_reconstruct_mc_face_slope(qm[i], q0[i], qp[i]);
    }
}

#if PRJ_MHD
/* The same with the resistive term: edge (dir, i, j, k) of the scratch block
 * with its cell widths set to dx, face fields bf[3] (FACE_IDX arrays of
 * NFACES), resistivity eta.  Returns the value; *stored receives what the
 * reference wrote to block->emf[dir][EDGE_IDX(dir, i, j, k)]. */
double refh_emf_upwind_block(int dir, int i, int j, int k, const double *emf_face,
    const double *emf_cell, const double *v_norm, const double *bfx, const double *bfy,
    const double *bfz, const double *dx, double eta, double *stored)
{
    prj_block *b = &g_sim.mesh.blocks[0];
    const double *bf[3];
    double keep[3];
    double r;
    int d;
    bf[0] = bfx;
    bf[1] = bfy;
    bf[2] = bfz;
    for (d = 0; d < 3; ++d) {
        keep[d] = b->dx[d];
        b->dx[d] = dx[d];
    }
    r = prj_mhd_emf_upwind(b, dir, i, j, k, emf_face, emf_cell, v_norm, bf, eta);
    *stored = b->emf[dir][EDGE_IDX(dir, i, j, k)];
    for (d = 0; d < 3; ++d) b->dx[d] = keep[d];
    return r;
}
#endif

/* prj_mhd_bf2bc (use_done != 0, with block->cell_derived_done = done) or
 * prj_mhd_bf2bc_all (src/prj_mhd.c:555-625) on raw block arrays: the W_mhd
 * slab and the face fields of `slot` are copied into block 0, the call runs,
 * and the slab is copied out whole. */
#if PRJ_MHD
int refh_bf2bc_raw(int slot, double *W_slab, const double *bfx, const double *bfy,
    const double *bfz, const int *done, int use_done)
{
    prj_block *b;
    const double *bf[3];
    int d;
    if (!g_ready || slot < 0 || slot >= PRJ_BLOCK_NSTAGES) return 1;
    b = &g_sim.mesh.blocks[0];
    bf[0] = bfx;
    bf[1] = bfy;
    bf[2] = bfz;
    memcpy(b->W_mhd, W_slab, sizeof(double) * (size_t)PRJ_NHYDRO * PRJ_BLOCK_NSTAGES *
        PRJ_BLOCK_NCELLS);
    for (d = 0; d < 3; ++d) {
        memcpy(prj_block_bf_stage(b, d, slot), bf[d], sizeof(double) * PRJ_BLOCK_NFACES);
    }
    if (use_done) {
        memcpy(b->cell_derived_done, done, sizeof(int) * (size_t)PRJ_BLOCK_NCELLS);
        prj_mhd_bf2bc(&g_sim.eos, b, slot);
    } else {
        prj_mhd_bf2bc_all(&g_sim.eos, b, slot);
    }
    memcpy(W_slab, b->W_mhd, sizeof(double) * (size_t)PRJ_NHYDRO * PRJ_BLOCK_NSTAGES *
        PRJ_BLOCK_NCELLS);
    return 0;
}
#endif

/* prj_flux_update (src/prj_flux.c:2149-2521) on raw block arrays.  Inputs in
 * the reference's own block layouts: W_slab the whole W_mhd slab
 * (NHYDRO x NSTAGES x NCELLS, WIDX), eosvar (EIDX), bf[d] the face fields of
 * `slot` (FACE_IDX, MHD).  They are copied into block 0 of the mesh, the
 * output arrays of that block are filled with NaN, prj_flux_update runs for
 * `slot`, and flux[d] (VIDX), v_riemann[d] (VRIDX) and, for MHD, Bv1[d]/Bv2[d]
 * (IDX) are copied out whole, so untouched entries stay NaN. */
int refh_flux_update_raw(int slot, const double *W_slab, const double *eosvar,
    const double *bfx, const double *bfy, const double *bfz,
    double *flux0, double *flux1, double *flux2,
    double *vr0, double *vr1, double *vr2,
    double *bv1_0, double *bv1_1, double *bv1_2,
    double *bv2_0, double *bv2_1, double *bv2_2)
{
    prj_block *b;
    double *fl[3] = {flux0, flux1, flux2};
    double *vr[3] = {vr0, vr1, vr2};
    size_t n;
    int d;
    (void)bfx; (void)bfy; (void)bfz;
    (void)bv1_0; (void)bv1_1; (void)bv1_2;
    (void)bv2_0; (void)bv2_1; (void)bv2_2;
    if (!g_ready || slot < 0 || slot >= PRJ_BLOCK_NSTAGES) return 1;
    b = &g_sim.mesh.blocks[0];
    memcpy(b->W_mhd, W_slab, sizeof(double) * (size_t)PRJ_NHYDRO * PRJ_BLOCK_NSTAGES *
        PRJ_BLOCK_NCELLS);
    memcpy(b->eosvar, eosvar, sizeof(double) * (size_t)PRJ_NVAR_EOSVAR * PRJ_BLOCK_NCELLS);
#if PRJ_MHD
    {
        const double *bf[3] = {bfx, bfy, bfz};
        double *bvo1[3] = {bv1_0, bv1_1, bv1_2};
        double *bvo2[3] = {bv2_0, bv2_1, bv2_2};
        for (d = 0; d < 3; ++d) {
            memcpy(prj_block_bf_stage(b, d, slot), bf[d], sizeof(double) * PRJ_BLOCK_NFACES);
            for (n = 0; n < (size_t)PRJ_BLOCK_NCELLS; ++n) {
                b->Bv1[d][n] = NAN;
                b->Bv2[d][n] = NAN;
            }
        }
#endif
        for (d = 0; d < 3; ++d) {
            for (n = 0; n < (size_t)PRJ_NVAR_CONS * PRJ_BLOCK_NCELLS; ++n) b->flux[d][n] = NAN;
            for (n = 0; n < (size_t)3 * PRJ_BLOCK_NCELLS; ++n) b->v_riemann[d][n] = NAN;
        }
        prj_flux_update(&g_sim.eos, &g_sim.rad, &g_sim.mesh, b, prj_block_prim_stage(b, slot),
            b->eosvar, b->flux, slot);
        for (d = 0; d < 3; ++d) {
            memcpy(fl[d], b->flux[d], sizeof(double) * (size_t)PRJ_NVAR_CONS * PRJ_BLOCK_NCELLS);
            memcpy(vr[d], b->v_riemann[d], sizeof(double) * 3 * (size_t)PRJ_BLOCK_NCELLS);
#if PRJ_MHD
            memcpy(bvo1[d], b->Bv1[d], sizeof(double) * (size_t)PRJ_BLOCK_NCELLS);
            memcpy(bvo2[d], b->Bv2[d], sizeof(double) * (size_t)PRJ_BLOCK_NCELLS);
#endif
        }
#if PRJ_MHD
    }
#endif
    return 0;
}

/* prj_flux_div (src/prj_flux.c:2523-2533) at every active cell of block 0,
 * on flux arrays given in VIDX layout: out[v][i][j][k]. */
int refh_flux_div_raw(const double *flux0, const double *flux1, const double *flux2,
    double *out)
{
    prj_block *b;
    double *fl[3];
    int i, j, k, v;
    if (!g_ready) return 1;
    b = &g_sim.mesh.blocks[0];
    fl[0] = (double *)flux0;
    fl[1] = (double *)flux1;
    fl[2] = (double *)flux2;
    for (i = 0; i < PRJ_BLOCK_SIZE; ++i)
    for (j = 0; j < PRJ_BLOCK_SIZE; ++j)
    for (k = 0; k < PRJ_BLOCK_SIZE; ++k) {
        double fd[PRJ_NVAR_CONS];
        prj_flux_div(fl, b->area, b->vol, i, j, k, fd);
        for (v = 0; v < PRJ_NVAR_CONS; ++v)
            out[((size_t)v * PRJ_BLOCK_SIZE + i) * PRJ_BLOCK_SIZE * PRJ_BLOCK_SIZE +
                (size_t)j * PRJ_BLOCK_SIZE + k] = fd[v];
    }
    return 0;
}

void refh_block_geometry(double *area, double *vol)
{
    const prj_block *b = &g_sim.mesh.blocks[0];
    area[0] = b->area[0];
    area[1] = b->area[1];
    area[2] = b->area[2];
    *vol = b->vol;
}

/* ---- refined meshes: coarse-fine flux correction and EMF restriction ----
 *
 * refh_amr_init builds a two-or-more-level mesh with the reference's own AMR
 * code: root blocks from prj_mesh_init, then, level by level, blocks whose box
 * meets the ball (cx, cy, cz, r) are marked, the 2:1 balance is enforced and
 * the marked blocks refined, and the neighbour slots rebuilt -- the loop of
 * problems/prj_problem_sedov.c:134-176.  Block data of every new block is
 * allocated by prj_amr_refine_block itself (single rank). */
static int refh_box_meets_ball(const prj_block *b, const double *ball)
{
    double d2 = 0.0;
    int a;
    for (a = 0; a < 3; ++a) {
        double c = ball[a];
        double q = c < b->xmin[a] ? b->xmin[a] - c : (c > b->xmax[a] ? c - b->xmax[a] : 0.0);
        d2 += q * q;
    }
    return d2 < ball[3] * ball[3];
}

int refh_amr_init(int nbx, int nby, int nbz, const double *xlim, const int *bc,
    int max_level, const double *ball, int max_blocks)
{
    int axis;
    int level;

    refh_free();
    memset(&g_sim, 0, sizeof(g_sim));
    memset(&g_mpi, 0, sizeof(g_mpi));
    g_mpi.rank = 0;
    g_mpi.totrank = 1;
    g_nb[0] = nbx;
    g_nb[1] = nby;
    g_nb[2] = nbz;
    g_sim.coord.x1min = xlim[0];
    g_sim.coord.x1max = xlim[1];
    g_sim.coord.x2min = xlim[2];
    g_sim.coord.x2max = xlim[3];
    g_sim.coord.x3min = xlim[4];
    g_sim.coord.x3max = xlim[5];
    g_sim.bc.bc_x1_inner = bc[0];
    g_sim.bc.bc_x1_outer = bc[1];
    g_sim.bc.bc_x2_inner = bc[2];
    g_sim.bc.bc_x2_outer = bc[3];
    g_sim.bc.bc_x3_inner = bc[4];
    g_sim.bc.bc_x3_outer = bc[5];
    for (axis = 0; axis < 3; ++axis) {
        int inner_p = bc[2 * axis] == PRJ_BC_PERIODIC;
        int outer_p = bc[2 * axis + 1] == PRJ_BC_PERIODIC;
        if (inner_p != outer_p) return 1;
        g_sim.mesh.periodic[axis] = inner_p;
    }
    g_sim.mesh.max_blocks = max_blocks;
    if (prj_mesh_init(&g_sim.mesh, nbx, nby, nbz, max_level, &g_sim.coord, 0) != 0) {
        return 2;
    }
    prj_mpi_decompose(&g_sim.mesh, &g_mpi);
    prj_amr_init_neighbors(&g_sim.mesh);
    for (level = 0; level < max_level; ++level) {
        int bidx;
        int any = 0;
        for (bidx = 0; bidx < g_sim.mesh.nblocks; ++bidx) {
            prj_block *b = &g_sim.mesh.blocks[bidx];
            if (b->id < 0 || b->active != 1 || b->level != level) continue;
            b->refine_flag = refh_box_meets_ball(b, ball) ? 1 : 0;
            any |= b->refine_flag;
        }
        if (!any) break;
        prj_amr_enforce_two_to_one(&g_sim.mesh, &g_mpi);
        if (prj_amr_refine_marked_blocks(&g_sim.mesh, &g_mpi) == 0) break;
        prj_amr_init_neighbors(&g_sim.mesh);
    }
    prj_mesh_update_max_active_level(&g_sim.mesh);
    g_ready = 1;
    return 0;
}

int refh_amr_nblocks(void)
{
    return g_ready ? g_sim.mesh.nblocks : -1;
}

/* geo[9] = xmin[3], xmax[3], dx[3]; info[4] = id, active, level, rank */
int refh_amr_block(int id, double *geo, int *info)
{
    const prj_block *b;
    int d;
    if (!g_ready || id < 0 || id >= g_sim.mesh.nblocks) return 1;
    b = &g_sim.mesh.blocks[id];
    for (d = 0; d < 3; ++d) {
        geo[d] = b->xmin[d];
        geo[3 + d] = b->xmax[d];
        geo[6 + d] = b->dx[d];
    }
    info[0] = b->id;
    info[1] = b->active;
    info[2] = b->level;
    info[3] = b->rank;
    return 0;
}

/* out[56][16]: id, rank, rel_level, type, send_loc_start[3], send_loc_end[3],
 * recv_loc_start[3], recv_loc_end[3] of every neighbour slot (src/prj_types.h:77-105). */
int refh_amr_slots(int id, int *out)
{
    const prj_block *b;
    int n, d;
    if (!g_ready || id < 0 || id >= g_sim.mesh.nblocks) return 1;
    b = &g_sim.mesh.blocks[id];
    for (n = 0; n < 56; ++n) {
        const prj_neighbor *s = &b->slot[n];
        int *o = out + 16 * n;
        o[0] = s->id;
        o[1] = s->rank;
        o[2] = s->rel_level;
        o[3] = s->type;
        for (d = 0; d < 3; ++d) {
            o[4 + d] = s->send_loc_start[d];
            o[7 + d] = s->send_loc_end[d];
            o[10 + d] = s->recv_loc_start[d];
            o[13 + d] = s->recv_loc_end[d];
        }
    }
    return 0;
}

/* Copy block->flux[axis] (NVAR_CONS x NCELLS, VIDX) in (set = 1) or out. */
int refh_amr_flux(int id, int axis, double *buf, int set)
{
    prj_block *b;
    size_t n = (size_t)PRJ_NVAR_CONS * PRJ_BLOCK_NCELLS;
    if (!g_ready || id < 0 || id >= g_sim.mesh.nblocks || axis < 0 || axis > 2) return 1;
    b = &g_sim.mesh.blocks[id];
    if (b->flux[axis] == 0) return 2;
    if (set) memcpy(b->flux[axis], buf, n * sizeof(double));
    else memcpy(buf, b->flux[axis], n * sizeof(double));
    return 0;
}

void refh_amr_flux_send(void)
{
    prj_riemann_flux_send(&g_sim.mesh, &g_mpi);
}

#if PRJ_MHD
/* Copy block->emf[dir] (NEDGES, EDGE_IDX) in (set = 1) or out. */
int refh_amr_emf(int id, int dir, double *buf, int set)
{
    prj_block *b;
    if (!g_ready || id < 0 || id >= g_sim.mesh.nblocks || dir < 0 || dir > 2) return 1;
    b = &g_sim.mesh.blocks[id];
    if (b->emf[dir] == 0) return 2;
    if (set) memcpy(b->emf[dir], buf, (size_t)PRJ_BLOCK_NEDGES * sizeof(double));
    else memcpy(buf, b->emf[dir], (size_t)PRJ_BLOCK_NEDGES * sizeof(double));
    return 0;
}

int refh_amr_fidelity(int id, int dir, int *out)
{
    const prj_block *b;
    if (!g_ready || id < 0 || id >= g_sim.mesh.nblocks || dir < 0 || dir > 2) return 1;
    b = &g_sim.mesh.blocks[id];
    if (b->edge_fidelity[dir] == 0) return 2;
    memcpy(out, b->edge_fidelity[dir], (size_t)PRJ_BLOCK_NEDGES * sizeof(int));
    return 0;
}

void refh_amr_emf_send(void)
{
    prj_mhd_emf_send(&g_sim.mesh, &g_mpi);
}
#endif
