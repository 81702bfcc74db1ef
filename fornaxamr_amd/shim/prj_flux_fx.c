/*
 * Reference-side drop-in for prj_flux_update (src/prj_flux.h:4-5).
 *
 * A FornaxAMR maintainer compiles this file into the reference in place of
 * the prj_flux_update of src/prj_flux.c (there: renamed, or left out of the
 * link) and links libfornax_gpu.so.  The reference's own stage drivers
 * (prj_timeint_stage1/2, src/prj_timeint.c:2732-2973, and eSSPRK,
 * :2470-2730) then get every block's fluxes, face velocities and, for MHD,
 * face EMF inputs Bv1/Bv2 from the gfx950 kernel; everything else -- flux
 * divergence, the CT update, the ghost fill, the EOS -- stays the reference's.
 *
 * The block arrays keep the reference's layouts (W in WIDX form at the stage
 * slot passed in, eosvar/flux/v_riemann AoSoA, Bf in FACE_IDX form, Bv1/Bv2 in
 * IDX form); fx_flux_update_block copies the block to device memory that the
 * library keeps, runs the kernel and copies the outputs back before it
 * returns, so the call is synchronous like the reference's.  A physical
 * failure (the HLLD abort sites of src/prj_riemann.c) prints the reference's
 * message and exits, as the reference does.
 *
 * Build (from the reference's src/): the reference sources with
 * -Dprj_flux_update=prj_flux_update_cpu on prj_flux.c only, this file with
 * -I<repo>/include, and -L<repo>/fornaxamr_amd -lfornax_gpu.
 */
#include <stdio.h>
#include <stdlib.h>

#include "prj.h"
#include "fornax_gpu.h"

void prj_flux_update(prj_eos *eos, prj_rad *rad, const prj_mesh *mesh,
    prj_block *block, double *W, double *eosvar, double *flux[3], int use_bf1)
{
    fx_block_io io = {0};
    int d;

    (void)eos;
    (void)rad;
    (void)mesh;
#if !PRJ_MHD
    (void)use_bf1;
#endif
    io.W = W;
    io.var_stride = (long long)PRJ_BLOCK_NSTAGES * PRJ_BLOCK_NCELLS;
    io.eosvar = eosvar;
    for (d = 0; d < 3; ++d) {
        io.flux[d] = flux[d];
        io.v_riemann[d] = block->v_riemann[d];
#if PRJ_MHD
        io.bf[d] = prj_block_bf_stage(block, d, prj_stage_slot_from_bf_arg(use_bf1));
        io.Bv1[d] = block->Bv1[d];
        io.Bv2[d] = block->Bv2[d];
#endif
    }
    if (fx_flux_update_block(&io, PRJ_MHD) != FX_OK) {
        fprintf(stderr, "%s\n", fx_last_error());
        exit(1);
    }
}
