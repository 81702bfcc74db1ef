/*
 * Reference-side drop-in for the main-loop step of a uniform single-level
 * mesh: prj_timeint_calc_dt and prj_timeint_step (src/main.c:510-545,
 * src/prj_timeint.c:2144-2256, 3221-3283) on a mesh whose state stays in GPU
 * memory.
 *
 * A FornaxAMR maintainer compiles this file into the reference and links
 * libfornax_gpu.so.  main.c calls prj_fx_attach() once the initial state is
 * set and filled (after the initial ghost fill, src/main.c:437-447); in the
 * main loop it calls prj_fx_calc_dt() where it called prj_timeint_calc_dt()
 * and prj_fx_step(sim, dt_step) where it called prj_timeint_step(), keeping
 * its own growth limit and its clipping of dt_step to the next output/event
 * time and to t_end; and it calls prj_fx_download() before anything reads the
 * blocks again (a dump, an analysis, AMR).  The integrator follows
 * TIME_INTEGRATION: RK2 or eSSPRK(10,4).
 *
 * The blocks keep the reference's layouts; the state crosses the boundary as
 * the C ABI's global arrays of the active cells: W[var][X][Y][Z] with Z
 * fastest and, for MHD, Bx[X+1][Y][Z], By[X][Y+1][Z], Bz[X][Y][Z+1].  A
 * physical failure prints the reference's message and exits, as the
 * reference does.
 */
#include <stdio.h>
#include <stdlib.h>

#include "prj.h"
#include "fornax_gpu.h"

static fx_mesh *g_fx;
static double *g_W;              /* global [var][X][Y][Z] staging */
static int g_n[3];
#if PRJ_MHD
static double *g_bf[3];          /* global face fields */
#endif

static void fx_die(int rc)
{
    if (rc != FX_OK) {
        fprintf(stderr, "%s\n", fx_last_error());
        exit(1);
    }
}

/* the block's first active cell in global cell coordinates */
static void block_origin(const prj_mesh *m, const prj_block *b, int o[3])
{
    o[0] = (int)((b->xmin[0] - m->coord.x1min) / b->dx[0] + 0.5);
    o[1] = (int)((b->xmin[1] - m->coord.x2min) / b->dx[1] + 0.5);
    o[2] = (int)((b->xmin[2] - m->coord.x3min) / b->dx[2] + 0.5);
}

/* slot-0 primitives of every block <-> the global staging array */
static void cells_exchange(prj_mesh *m, int to_global)
{
    int b;
    for (b = 0; b < m->nblocks; ++b) {
        prj_block *bl = &m->blocks[b];
        double *W = prj_block_prim_stage(bl, 0);
        int o[3], v, i, j, k;
        block_origin(m, bl, o);
        for (v = 0; v < PRJ_NHYDRO; ++v)
            for (i = 0; i < PRJ_BLOCK_SIZE; ++i)
                for (j = 0; j < PRJ_BLOCK_SIZE; ++j)
                    for (k = 0; k < PRJ_BLOCK_SIZE; ++k) {
                        size_t g = (((size_t)v * g_n[0] + (size_t)(o[0] + i)) * g_n[1] +
                                    (size_t)(o[1] + j)) * g_n[2] + (size_t)(o[2] + k);
                        if (to_global) {
                            g_W[g] = W[WIDX(v, i, j, k)];
                        } else {
                            W[WIDX(v, i, j, k)] = g_W[g];
                        }
                    }
    }
}

#if PRJ_MHD
/* slot-0 face fields of every block (the block's active faces: 0..16 along
 * the face normal, 0..15 across) <-> the global face arrays */
static void faces_exchange(prj_mesh *m, int to_global)
{
    int b, d;
    for (b = 0; b < m->nblocks; ++b) {
        prj_block *bl = &m->blocks[b];
        int o[3];
        block_origin(m, bl, o);
        for (d = 0; d < 3; ++d) {
            double *bf = prj_block_bf_stage(bl, d, 0);
            int ny = g_n[1] + (d == 1), nz = g_n[2] + (d == 2);
            int imax = PRJ_BLOCK_SIZE - 1 + (d == 0);
            int jmax = PRJ_BLOCK_SIZE - 1 + (d == 1);
            int kmax = PRJ_BLOCK_SIZE - 1 + (d == 2);
            int i, j, k;
            for (i = 0; i <= imax; ++i)
                for (j = 0; j <= jmax; ++j)
                    for (k = 0; k <= kmax; ++k) {
                        size_t g = ((size_t)(o[0] + i) * ny + (size_t)(o[1] + j)) * nz +
                                   (size_t)(o[2] + k);
                        if (to_global) {
                            g_bf[d][g] = bf[FACE_IDX(d, i, j, k)];
                        } else {
                            bf[FACE_IDX(d, i, j, k)] = g_bf[d][g];
                        }
                    }
        }
    }
}
#endif

void prj_fx_detach(void);

/* one mesh per process: attaching again releases the previous one first */
void prj_fx_attach(prj_sim *sim)
{
    prj_mesh *m = &sim->mesh;
    fx_mesh_desc desc = {{0}};
    if (g_fx != 0) {
        prj_fx_detach();
    }
    const double xl[6] = {m->coord.x1min, m->coord.x1max, m->coord.x2min,
                          m->coord.x2max, m->coord.x3min, m->coord.x3max};
    const int bc[6] = {sim->bc.bc_x1_inner, sim->bc.bc_x1_outer, sim->bc.bc_x2_inner,
                       sim->bc.bc_x2_outer, sim->bc.bc_x3_inner, sim->bc.bc_x3_outer};
    int d;
    for (d = 0; d < 3; ++d) {
        g_n[d] = desc.n[d] = m->root_nx[d] * PRJ_BLOCK_SIZE;
    }
    for (d = 0; d < 6; ++d) {
        desc.xlim[d] = xl[d];
        desc.bc[d] = bc[d];
    }
    /* the ideal-gas gamma the problem's USER EOS plug-in returns, as the
     * initial eos fill stored it (PRJ_DEFINE_IDEAL_EOS,
     * problems/prj_problem_ideal_eos.h) */
    desc.gamma = m->blocks[0].eosvar[EIDX(PRJ_EOSVAR_GAMMA, 0, 0, 0)];
    desc.mhd = PRJ_MHD;
    desc.eta = m->mhd_eta;
    fx_die(fx_mesh_create(&desc, &g_fx));
#if TIME_INTEGRATION == PRJ_TIMEINT_ESSPRK10_4
    fx_die(fx_mesh_set_integrator(g_fx, FX_ESSPRK10_4));
#endif
    g_W = malloc(sizeof(double) * PRJ_NHYDRO * (size_t)g_n[0] * g_n[1] * g_n[2]);
    if (g_W == 0) {
        fprintf(stderr, "prj_fx_attach: out of host memory\n");
        exit(1);
    }
    cells_exchange(m, 1);
    fx_die(fx_mesh_set_prim(g_fx, g_W));
#if PRJ_MHD
    for (d = 0; d < 3; ++d) {
        g_bf[d] = malloc(sizeof(double) * (size_t)(g_n[0] + (d == 0)) * (g_n[1] + (d == 1)) *
                         (g_n[2] + (d == 2)));
        if (g_bf[d] == 0) {
            fprintf(stderr, "prj_fx_attach: out of host memory\n");
            exit(1);
        }
    }
    faces_exchange(m, 1);
    fx_die(fx_mesh_set_bf(g_fx, g_bf[0], g_bf[1], g_bf[2]));
#endif
}

/* prj_timeint_calc_dt on the device state (x 6 under eSSPRK(10,4)) */
double prj_fx_calc_dt(prj_sim *sim)
{
    double dt = 0.0;
    fx_die(fx_calc_dt(g_fx, sim->cfl, &dt));
    return dt;
}

/* prj_timeint_step(..., dt_step, ...) on the device state */
void prj_fx_step(prj_sim *sim, double dt_step)
{
    (void)sim;
    fx_die(fx_step_dt(g_fx, dt_step));
}

/* the device state back into the blocks' slot 0, then the reference's own
 * refill of eosvar and ghosts (src/main.c:437-447) */
void prj_fx_download(prj_sim *sim, prj_mpi *mpi)
{
    prj_mesh *m = &sim->mesh;
    int b;
    fx_die(fx_mesh_get_prim(g_fx, 0, g_W));
    cells_exchange(m, 0);
#if PRJ_MHD
    fx_die(fx_mesh_get_bf(g_fx, 0, g_bf[0], g_bf[1], g_bf[2]));
    faces_exchange(m, 0);
    for (b = 0; b < m->nblocks; ++b) {
        prj_mhd_bf2bc_all_mesh(&sim->eos, m, &m->blocks[b], 0);
    }
#else
    (void)b;
#endif
    prj_eos_fill_active_cells(m, &sim->eos, mpi, 1, PRJ_EOS_CTX_MAIN);
    prj_boundary_fill_ghosts_and_bf(m, mpi, &sim->bc, 1, 0, &sim->eos, 0, &sim->rad,
                                    PRJ_BOUNDARY_TIMER_SCOPE_NONE);
    prj_eos_fill_mesh(m, &sim->eos, mpi, 1, PRJ_EOS_CTX_MAIN);
}

void prj_fx_detach(void)
{
    int d;
    fx_mesh_destroy(g_fx);
    g_fx = 0;
    free(g_W);
    g_W = 0;
#if PRJ_MHD
    for (d = 0; d < 3; ++d) {
        free(g_bf[d]);
        g_bf[d] = 0;
    }
#else
    (void)d;
#endif
}
