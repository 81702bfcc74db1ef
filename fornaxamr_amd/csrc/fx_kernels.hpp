// Kernel argument blocks and launchers shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/fornax_gpu.h"
#include "fx_layout.hpp"

namespace fx {

struct HydroStageArgs {
    Geom g;
    const double* Win;      // stage input W (padded layout, 6 variables)
    const double* W0;       // W^n for the stage-2 blend (== Win for stage 1)
    double* Wout;           // stage output (may alias W0 for stage 2)
    const double* dt_ptr;   // device dt, or null to use dt
    double dt;
    // stage 2: cfl_part[b] = largest CFL denominator of block b's output
    // (b = (bx * nby + by) * nbz + bz); null: not recorded
    double* cfl_part = nullptr;
    int riemann = 0;        // FX_RIEMANN_HLLC or FX_RIEMANN_HLLE
};

struct MhdStageArgs {
    Geom g;
    const double* Win;      // 9 variables
    const double* W0;
    double* Wout;
    const double* bf_in;    // face fields of the stage input (3 components)
    const double* bf0;      // face fields at t^n (stage-2 blend)
    double* bf_out;
    double* fdat;           // scratch: face EMFs and normal velocities (3 dirs x 3)
    const double* zring;    // compact copies of the ghost planes z = -1, nz (zring_doubles)
    int* err;               // device error word: [code, dir, x, y, z]
    const double* dt_ptr;
    double dt;
    int ct_split = 1;       // CT workgroups per block (1, 2 or 4), each BS / ct_split x-planes
    // stage 2: cfl_part[w] = largest CFL denominator of the cells CT workgroup
    // w wrote (nblocks * ct_split entries); null: not recorded
    double* cfl_part = nullptr;
};

// ---- ragged ensembles of independent hydro meshes --------------------------
struct EnsInst {
    Geom g;
    double* W[2];           // stage slots of this instance (padded layout)
    double* dev;            // [0] dt, [1] dt_prev, [2] CFL reduction (bits of the largest denominator)
};

struct EnsBlock {
    int inst, bx, by, bz;
};

// Batched CFL step of an ensemble.  launch_ens_dt_reduce: dev[2] of every
// instance = the bits of its largest denominator, read from its state.
// launch_ens_dt_clear: dev[2] = 0 (before a batched stage 2, which raises it
// to the largest denominator of the state it writes).  launch_ens_dt_finalize:
// dt of every instance from dev[2] (dev[0]) and the factor limit (dev[1]).
hipError_t launch_ens_dt_reduce(const EnsInst* inst, int ninst, const EnsBlock* blocks,
                                int nblocks, hipStream_t s);
hipError_t launch_ens_dt_clear(const EnsInst* inst, int ninst, hipStream_t s);
hipError_t launch_ens_dt_finalize(const EnsInst* inst, int ninst, double cfl, double factor,
                                  hipStream_t s);
// persistent_wgs == 0: one workgroup per table entry (XCD remap if xcd_remap);
// otherwise that many persistent workgroups pull table entries from `queues`
// (8 counters, one per XCD chunk when xcd_queues, else only the first).
hipError_t launch_ens_stage(const EnsInst* inst, const EnsBlock* blocks, int nblocks, int stage,
                            int persistent_wgs, int xcd_remap, int xcd_queues, int* queues,
                            hipStream_t s);

// Compact copies of the MHD ghost planes z = -1 and z = nz (the 9 cell
// variables over x, y in [-2, n + 2)), which the ghost-layer face kernel reads
// for its x- and y-faces on those planes instead of one double per line of the
// padded arrays: copy 0 with y fastest, [v][zc][x + 2][y + 2]; copy 1 with x
// fastest, [v][zc][y + 2][x + 2]; zc = 0 for z = -1, 1 for z = nz.
__host__ __device__ inline int64_t zring_plane(const Geom& g)
{
    return (int64_t)(g.nx + 4) * (g.ny + 4);
}
__host__ __device__ inline int64_t zring_doubles(const Geom& g)
{
    return 2 * 9 * 2 * zring_plane(g);
}

// fill the two-cell ghost shell of the domain (periodic / outflow / reflect;
// MHD: cell B from the clamped outflow face fields), and (zring non-null) the
// compact copies of its planes z = -1 and z = nz.  The MHD stage needs it;
// the hydro stage reads out-of-domain cells at their sources.
hipError_t launch_boundary_fill(const Geom& g, double* W, const double* bf, double* zring,
                                hipStream_t s);

hipError_t launch_hydro_stage(const HydroStageArgs& a, int stage, hipStream_t s);

// side (optional, with the events fork/join): a second stream on which the
// ghost-layer face kernel runs beside the flux kernel; the CT kernel waits
// for both.  side == nullptr: all three kernels in order on s.
hipError_t launch_mhd_stage(const MhdStageArgs& a, int stage, hipStream_t s,
                            hipStream_t side = nullptr, hipEvent_t fork = nullptr,
                            hipEvent_t join = nullptr);
// CT workgroups per block: more on meshes with fewer blocks than CT slots
int mhd_ct_split(const Geom& g, int ncu);

// cell-centred B of the active cells from the faces (prj_mhd_bf2bc)
hipError_t launch_bf2bc(const Geom& g, double* W, const double* bf, hipStream_t s);

// CFL time step: dt_out[0] = min(1e99, min over active cells of cfl / denom)
// times `scale` (the integrator's SSP coefficient); if dt_prev is non-null
// and *dt_prev > 0, also applies min(dt, factor * dt_prev) and stores the
// result to *dt_prev.  partial: device scratch of CALC_DT_WGS doubles (one per
// workgroup of the reduction).
constexpr int CALC_DT_WGS = 4096;
hipError_t launch_calc_dt(const Geom& g, const double* W, double cfl, double* partial,
                          double* dt_out, double* dt_prev, double factor, hipStream_t s,
                          double scale = 1.0);
// The two halves of launch_calc_dt: the reduction writes the largest
// denominator of each of `nwg` stretches of the active cells to partial[0..nwg);
// the final launch forms dt from n partials -- those of the reduction or the
// per-block maxima a hydro stage 2 recorded (HydroStageArgs::cfl_part).
int calc_dt_wgs(const Geom& g);
hipError_t launch_dt_reduce(const Geom& g, const double* W, double* partial, int nwg,
                            hipStream_t s);
hipError_t launch_dt_final(const double* partial, int n, double cfl, double* dt_out,
                           double* dt_prev, double factor, double scale, hipStream_t s);
// eSSPRK convex blend of the active cells: W <- prim(w u(Ws) + (1 - w) u(W))
hipError_t launch_blend(const Geom& g, double* W, const double* Ws, double saved_weight,
                        hipStream_t s);
// the same on an MHD mesh: the face fields first (bf <- w bfs + (1 - w) bf),
// then the cells with their field from the blended faces
hipError_t launch_blend_mhd(const Geom& g, double* W, const double* Ws, double* bf,
                            const double* bfs, double saved_weight, hipStream_t s);
// dst[0] = src[0] / div on the device
hipError_t launch_div_scalar(const double* src, double* dst, double div, hipStream_t s);

// per-face / per-edge batched entry points (global-memory SoA arrays)
hipError_t launch_mc_slope_batch(int n, const double* qm, const double* q0, const double* qp,
                                double* slope, hipStream_t s);
hipError_t launch_hllc_batch(int n, const double* WL, const double* WR, const double* pL,
                             const double* pR, const double* gL, const double* gR,
                             const double* du, const double* dv, const double* dw,
                             double* flux, double* vface, hipStream_t s);
hipError_t launch_hlld_batch(int n, const double* WL, const double* WR, const double* pL,
                             const double* pR, const double* gL, const double* gR,
                             const double* bn, double* flux, double* vface, double* bv1,
                             double* bv2, int* status, hipStream_t s);
hipError_t launch_emf_batch(int n, const double* ef, const double* ec, const double* vn,
                            double* emf, hipStream_t s);
// prj_mhd_emf_upwind with the resistive term over n edges of one block in the
// reference's block layouts (edge[4m..4m+3] = dir, i, j, k)
struct EmfBlockArgs {
    int n;
    const int* edge;
    const double* ef;       // emf_face, emf_cell, v_norm: 4 x n (SoA)
    const double* ec;
    const double* vn;
    const double* eta;      // n
    const double* bf[3];    // the block's face fields (FACE_IDX), or null
    double dx[3];
    double* block_emf[3];   // the block's edge EMFs (EDGE_IDX), or null
    double* emf;            // n: the value (NaN where status != 0)
    int* status;            // n
};
hipError_t launch_emf_block_batch(const EmfBlockArgs& a, hipStream_t s);

// ---- differentiable scheme (van Albada + HLLE) and its adjoint --------------
// compact periodic layout [var][X][Y][Z], N = X*Y*Z
struct AdjGeom {
    int n[3];
    int64_t N;
    double dx[3], area[3], vol;
    double gamma;
    double eps2;            // van Albada regularisation
};

// one forward RK2 stage: F scratch 18*N; stage 2 may write Wout == W0
hipError_t launch_adj_stage(const AdjGeom& g, int stage, double dt, const double* Win,
                            const double* W0, double* F, double* Wout, hipStream_t s);
hipError_t launch_adj_w2u(const AdjGeom& g, const double* W, const double* Wb, double* Ub,
                          hipStream_t s);
hipError_t launch_adj_u2w(const AdjGeom& g, const double* W, const double* Ub, double* Wb,
                          hipStream_t s);
// reverse of one RK2 step (W0, W1 its stage inputs); U2b in: cotangent of the
// step's output in conserved form, out: that of W0; U1b 6*N and G 36*N scratch
hipError_t launch_adj_step_reverse(const AdjGeom& g, double dt, const double* W0, const double* W1,
                                   double* U2b, double* U1b, double* G, hipStream_t s);

// ---- reference block layouts (per-block drop-in) ----------------------------
hipError_t launch_block_flux(const fx_block_io* blocks_dev, int n, int mhd, int* err,
                             hipStream_t s);
hipError_t launch_block_bf2bc(const fx_bf2bc_io* blocks_dev, int n, int* err, hipStream_t s);
hipError_t launch_block_flux_div(const double* f0, const double* f1, const double* f2,
                                 const double* area, double vol, int nvar, double* out,
                                 hipStream_t s);

// ---- coarse-fine corrections of refined meshes (fx_amr.hip) ---------------
struct CfPtrs {             // device arrays of one block of the plan's table
    double* flux[3];
    double* emf[3];
    int* fid[3];
};
struct CfFlux {             // one coarse face cell: dst = 0.25 (s0 + s1 + s2 + s3)
    int dst, src, axis, dst_off;
    int s[4];               // BIDX(nvar, 0, ...) of the four fine faces
};
struct CfEmf {              // one coarse edge: dst = ((0 + a) + b) / 2
    int dst, src, dir, dst_idx;
    int a, b, store, pad;   // store 0: overwritten later, finiteness check only
};
struct CfHostPlan {
    std::vector<CfFlux> flux;
    std::vector<int> flux_off;      // level l: records [off[l], off[l+1])
    std::vector<CfEmf> emf;
    std::vector<int> emf_off;
    std::vector<int> fid_blocks;    // local blocks whose edge fidelity is reset
    std::string flux_err, emf_err;  // why a part is unavailable ("" if usable)
};
// fills p; a part whose inputs are unusable gets an empty list and a reason
void cf_build_plan(int nblocks, const fx_cf_block* blocks, int nlinks, const fx_cf_link* links,
                   int nvar, CfHostPlan& p);
hipError_t launch_cf_flux(const CfPtrs* tab, const CfFlux* recs, int n, int nvar, hipStream_t s);
hipError_t launch_cf_fid_init(const CfPtrs* tab, const int* blocks, int n, hipStream_t s);
hipError_t launch_cf_emf(const CfPtrs* tab, const CfEmf* recs, int n, int* status, hipStream_t s);

}  // namespace fx
