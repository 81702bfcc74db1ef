/*
 * Reference-side drop-ins, with the reference's own signatures, for the
 * per-face, per-edge and per-block helpers of the stage:
 *
 *   prj_riemann_hllc   (src/prj_riemann.h:7-10)     hydro builds
 *   prj_riemann_hlld   (src/prj_riemann.h:12-15)    MHD builds
 *   prj_mhd_emf_upwind (src/prj_mhd.h:18-20)        MHD builds
 *   prj_mhd_bf2bc, prj_mhd_bf2bc_all, prj_mhd_bf2bc_mesh,
 *   prj_mhd_bf2bc_all_mesh (src/prj_mhd.h:7-13)     MHD builds
 *
 * Each call runs the library's batched entry point with one item on the
 * null stream: the call's inputs go to device memory in one copy, the
 * outputs come back in one copy, which also waits for the kernel.  These
 * forms exist so that the reference's own block loops (prj_flux_update,
 * prj_timeint_mhd_update_emf, the ghost fill) can run unchanged on the GPU
 * functions, one item at a time; they cost a host-device round trip per call.
 * The batched and mesh-level entry points are the fast paths.
 *
 * A maintainer compiles this file into the reference with -I<repo>/include
 * and the HIP headers, links libfornax_gpu.so and the HIP runtime, and
 * renames the reference's definitions (prj_riemann.c, prj_mhd.c compiled
 * with -D<name>=<name>_cpu).  Where the reference aborts, the reference's
 * message is printed and the process exits, as the reference does.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "prj.h"
#include "fornax_gpu.h"

/* device memory of one call: scalars first, then (MHD) the block's three
 * face arrays for the resistive term; allocated on the first call, held
 * until the process ends (the calls are serial, as the reference's are) */
enum { CALL_DOUBLES = 64 };
static double *g_dev;
#if PRJ_MHD
static double *g_dev_bf;
#endif

static void die(const char *msg)
{
    fprintf(stderr, "%s\n", msg);
    exit(EXIT_FAILURE);
}

static void hip_ok(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        exit(EXIT_FAILURE);
    }
}

static void fx_ok(int rc)
{
    if (rc != FX_OK) die(fx_last_error());
}

static double *call_buffer(void)
{
    if (g_dev == 0) {
        hip_ok(hipMalloc((void **)&g_dev, CALL_DOUBLES * sizeof(double)), "hipMalloc");
#if PRJ_MHD
        hip_ok(hipMalloc((void **)&g_dev_bf, 3 * (size_t)PRJ_BLOCK_NFACES * sizeof(double)),
               "hipMalloc");
#endif
    }
    return g_dev;
}

#if !PRJ_MHD
void prj_riemann_hllc(const double *WL, const double *WR,
    double pL, double pR, double gL, double gR,
    const prj_eos *eos, double *flux, double v_face[3],
    double deltau, double deltav, double deltaw)
{
    /* in: WL[6] WR[6] pL pR gL gR du dv dw (0..18); out: flux[6] v_face[3] (19..27) */
    double h[28];
    double *d = call_buffer();
    (void)eos;
    memcpy(h, WL, 6 * sizeof(double));
    memcpy(h + 6, WR, 6 * sizeof(double));
    h[12] = pL;
    h[13] = pR;
    h[14] = gL;
    h[15] = gR;
    h[16] = deltau;
    h[17] = deltav;
    h[18] = deltaw;
    hip_ok(hipMemcpy(d, h, 19 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    fx_ok(fx_riemann_hllc_batch(1, d, d + 6, d + 12, d + 13, d + 14, d + 15, d + 16, d + 17,
                                d + 18, d + 19, d + 25, NULL));
    hip_ok(hipMemcpy(h + 19, d + 19, 9 * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    memcpy(flux, h + 19, 6 * sizeof(double));
    memcpy(v_face, h + 25, 3 * sizeof(double));
}
#endif

#if PRJ_MHD
void prj_riemann_hlld(const double *WL, const double *WR,
    double pL, double pR, double gL, double gR,
    const prj_eos *eos, double bn, double *flux, double v_face[3],
    double *Bv1, double *Bv2, double deltau, double deltav, double deltaw)
{
    /* in: WL[9] WR[9] pL pR gL gR bn (0..22); out: flux[9] v_face[3] Bv1 Bv2
     * (23..36), status (int at double 37) */
    double h[38];
    double *d = call_buffer();
    int st = 0;
    (void)eos;
    (void)deltau;
    (void)deltav;
    (void)deltaw;
    memcpy(h, WL, 9 * sizeof(double));
    memcpy(h + 9, WR, 9 * sizeof(double));
    h[18] = pL;
    h[19] = pR;
    h[20] = gL;
    h[21] = gR;
    h[22] = bn;
    hip_ok(hipMemcpy(d, h, 23 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    fx_ok(fx_riemann_hlld_batch(1, d, d + 9, d + 18, d + 19, d + 20, d + 21, d + 22, d + 23,
                                d + 32, d + 35, d + 36, (int *)(d + 37), NULL));
    hip_ok(hipMemcpy(h + 23, d + 23, 15 * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    memcpy(&st, h + 37, sizeof(int));
    if (st != 0) die(fx_status_message(st));
    memcpy(flux, h + 23, 9 * sizeof(double));
    memcpy(v_face, h + 32, 3 * sizeof(double));
    *Bv1 = h[35];
    *Bv2 = h[36];
}

double prj_mhd_emf_upwind(prj_block *block, int dir, int i, int j, int k,
    const double emf_face[4], const double emf_cell[4], const double v_norm[4],
    const double *bf[3], double eta)
{
    /* in: edge (4 ints in double 0..1), emf_face[4] emf_cell[4] v_norm[4] eta
     * (2..14); out: emf (15), status (int at double 16) */
    double h[17];
    int edge[4];
    double *d = call_buffer();
    const double *dbf[3] = {0, 0, 0};
    int st = 0;
    if (block == 0 || dir < 0 || dir >= 3 || block->emf[dir] == 0) {
        die("prj_mhd_emf_upwind: invalid emf destination");
    }
    if (emf_face == 0 || emf_cell == 0 || v_norm == 0) {
        die("prj_mhd_emf_upwind: null input array");
    }
    edge[0] = dir;
    edge[1] = i;
    edge[2] = j;
    edge[3] = k;
    memcpy(h, edge, sizeof(edge));
    memcpy(h + 2, emf_face, 4 * sizeof(double));
    memcpy(h + 6, emf_cell, 4 * sizeof(double));
    memcpy(h + 10, v_norm, 4 * sizeof(double));
    h[14] = eta;
    if (eta != 0.0 && bf != 0 && bf[0] != 0 && bf[1] != 0 && bf[2] != 0) {
        int c;
        for (c = 0; c < 3; ++c) {
            dbf[c] = g_dev_bf + (size_t)c * PRJ_BLOCK_NFACES;
            hip_ok(hipMemcpy((void *)dbf[c], bf[c], PRJ_BLOCK_NFACES * sizeof(double),
                             hipMemcpyHostToDevice), "hipMemcpy");
        }
    }
    hip_ok(hipMemcpy(d, h, 15 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    fx_ok(fx_emf_upwind_block_batch(1, (const int *)d, d + 2, d + 6, d + 10, d + 14, dbf,
                                    block->dx, NULL, d + 15, (int *)(d + 16), NULL));
    hip_ok(hipMemcpy(h + 15, d + 15, 2 * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    memcpy(&st, h + 16, sizeof(int));
    if (st != 0) die(fx_status_message(st));
    block->emf[dir][EDGE_IDX(dir, i, j, k)] = h[15];
    return h[15];
}

/* the cell-centred field of every cell of the block (ghosts included) from
 * the slot's face fields; cells flagged in done (or none) are skipped */
static void bf2bc(prj_block *block, int use_bf1, const int *done)
{
    fx_bf2bc_io io;
    const int slot = prj_stage_slot_from_bf_arg(use_bf1);
    int d;
    memset(&io, 0, sizeof(io));
    io.W = prj_block_prim_stage(block, slot);
    io.var_stride = (long long)PRJ_BLOCK_NSTAGES * PRJ_BLOCK_NCELLS;
    for (d = 0; d < 3; ++d) {
        io.bf[d] = prj_block_bf_stage(block, d, slot);
    }
    io.done = done;
    fx_ok(fx_bf2bc_block(&io));
}

void prj_mhd_bf2bc(prj_eos *eos, prj_block *block, int use_bf1)
{
    (void)eos;
    bf2bc(block, use_bf1, block->cell_derived_done);
}

void prj_mhd_bf2bc_all(prj_eos *eos, prj_block *block, int use_bf1)
{
    (void)eos;
    bf2bc(block, use_bf1, 0);
}

/* the GR metric factor of the _mesh forms is 1 on this path */
void prj_mhd_bf2bc_mesh(prj_eos *eos, const prj_mesh *mesh, prj_block *block, int use_bf1)
{
    (void)eos;
    (void)mesh;
    bf2bc(block, use_bf1, block->cell_derived_done);
}

void prj_mhd_bf2bc_all_mesh(prj_eos *eos, const prj_mesh *mesh, prj_block *block,
                            int use_bf1)
{
    (void)eos;
    (void)mesh;
    bf2bc(block, use_bf1, 0);
}
#endif
