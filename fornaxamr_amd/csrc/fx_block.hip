// Per-block flux assembly in the reference's own block layouts: the drop-in
// form of prj_flux_update (src/prj_flux.c:2149-2521) and prj_flux_div
// (src/prj_flux.c:2523-2533) for blocks kept in the reference's 20^3 storage.
//
// Block layouts (src/prj_defs.h:383-459, 16^3 active cells, 2 ghost layers,
// AoSoA lane width 4):
//   LIDX(i,j,k)        = ((i+2)*20 + (j+2))*20 + (k+2)
//   BIDX(nc,v,i,j,k)   = ((((i+2)*20 + (j+2))*5 + (k+2)/4)*nc + v)*4 + (k+2)%4
//   FACE_IDX(dir,...)  = LIDX with 21 slots along dir
// Faces written (src/prj_flux.c:2221-2262): normal index 0..16 in every
// direction; transverse 0..15 for hydro, -1..16 for MHD (CT needs one more
// layer).  A face (i,j,k) of direction dir lies between cell (i,j,k) - e_dir
// (left) and (i,j,k) (right).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/fornax_gpu.h"
#include "fx_kernels.hpp"
#include "fx_physics.hpp"

namespace fx {
namespace {

constexpr int NT = 256;
constexpr int RB = 20;          // reference block edge incl. ghosts (PRJ_BS)
constexpr int RNG = 2;          // ghost layers
constexpr int AW = 4;           // AoSoA lane width (PRJ_AOSOA_W)
constexpr int SPLIT = 8;        // workgroups per block of the z-face kernel

__device__ __forceinline__ int lidx(int i, int j, int k)
{
    return ((i + RNG) * RB + (j + RNG)) * RB + (k + RNG);
}

__device__ __forceinline__ int bidx(int nc, int v, int i, int j, int k)
{
    return ((((i + RNG) * RB + (j + RNG)) * (RB / AW) + (k + RNG) / AW) * nc + v) * AW +
           (k + RNG) % AW;
}

__device__ __forceinline__ int face_idx(int dir, int i, int j, int k)
{
    if (dir == 0) return ((i + RNG) * RB + (j + RNG)) * RB + (k + RNG);
    if (dir == 1) return ((i + RNG) * (RB + 1) + (j + RNG)) * RB + (k + RNG);
    return ((i + RNG) * RB + (j + RNG)) * (RB + 1) + (k + RNG);
}

template <int MHD>
__device__ void block_face(const fx_block_io& b, int dir, int i, int j, int k, int* err)
{
    constexpr int NHV = MHD ? 9 : 6;
    const int64_t vs = b.var_stride;
    const double* W = b.W;
    const int ex = dir == 0, ey = dir == 1, ez = dir == 2;
    double WL[NHV], WR[NHV];
#pragma unroll
    for (int lv = 0; lv < NHV; ++lv) {
        const double* Wv = W + (int64_t)gmap(dir, lv) * vs;
        double q[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) q[s] = Wv[lidx(i + (s - 2) * ex, j + (s - 2) * ey, k + (s - 2) * ez)];
        WL[lv] = q[1] + mc_half(q[0], q[1], q[2]);
        WR[lv] = q[2] - mc_half(q[1], q[2], q[3]);
    }
    // pressure and gamma reconstructed from eosvar (src/prj_flux.c:555-562)
    double pq[4], gq[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int ci = i + (s - 2) * ex, cj = j + (s - 2) * ey, ck = k + (s - 2) * ez;
        pq[s] = b.eosvar[bidx(3, 0, ci, cj, ck)];
        gq[s] = b.eosvar[bidx(3, 2, ci, cj, ck)];
    }
    const double pL = pq[1] + mc_half(pq[0], pq[1], pq[2]);
    const double pR = pq[2] - mc_half(pq[1], pq[2], pq[3]);
    const double gL = gq[1] + mc_half(gq[0], gq[1], gq[2]);
    const double gR = gq[2] - mc_half(gq[1], gq[2], gq[3]);
    double Fl[NHV], vloc[3];
    const int c = lidx(i, j, k);
    if (MHD) {
        const double bn = b.bf[dir][face_idx(dir, i, j, k)];
        WL[B1] = bn;
        WR[B1] = bn;
        double bv1 = 0.0, bv2 = 0.0;
        const int rc = hlld(WL, WR, pL, pR, gL, gR, bn, Fl, vloc, &bv1, &bv2);
        if (rc != 0) {
            if (atomicCAS(err, 0, rc) == 0) {
                err[1] = dir;
                err[2] = i;
                err[3] = j;
                err[4] = k;
            }
            return;
        }
        b.Bv1[dir][c] = bv1;
        b.Bv2[dir][c] = bv2;
    } else {
        // velocity deltas (src/prj_flux.c:1967-2036)
        const int a1 = (dir + 1) % 3, a2 = (dir + 2) % 3;
        const int li = i - ex, lj = j - ey, lk = k - ez;
        const double* vn = W + (int64_t)(V1 + dir) * vs;
        const double du = vn[c] - vn[lidx(li, lj, lk)];
        double dtr[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int a = t == 0 ? a1 : a2;
            const int tx = a == 0, ty = a == 1, tz = a == 2;
            const double* va = W + (int64_t)(V1 + a) * vs;
            const double l0 = va[lidx(li, lj, lk)];
            const double r0 = va[c];
            const double dl = dmin_delta(l0 - va[lidx(li - tx, lj - ty, lk - tz)],
                                         va[lidx(li + tx, lj + ty, lk + tz)] - l0);
            const double dr = dmin_delta(r0 - va[lidx(i - tx, j - ty, k - tz)],
                                         va[lidx(i + tx, j + ty, k + tz)] - r0);
            dtr[t] = dmin_delta(dl, dr);
        }
        hllc(WL, WR, pL, pR, gL, gR, du, dtr[0], dtr[1], Fl, vloc);
    }
    double* F = b.flux[dir];
#pragma unroll
    for (int lv = 0; lv < NHV; ++lv) F[bidx(NHV, gmap(dir, lv), i, j, k)] = Fl[lv];
    double* vr = b.v_riemann[dir];
#pragma unroll
    for (int r = 0; r < 3; ++r) vr[bidx(3, gmap(dir, 1 + r) - 1, i, j, k)] = vloc[r];
}

// One face per thread, faces enumerated k fastest (coalesced for z-faces,
// whose stencil runs along the arrays' fastest axis); grid (block, SPLIT).
// The direction is a template constant so that the descriptor's per-direction
// pointers are selected at compile time.
template <int MHD, int DIR>
__global__ void __launch_bounds__(NT) block_face_kernel(const fx_block_io* __restrict__ blocks,
                                                        int* err)
{
    constexpr int dir = DIR;
    const fx_block_io b = blocks[blockIdx.x];
    const int t0 = MHD ? -1 : 0;
    const int tn = MHD ? 18 : 16;          // transverse count
    int lo[3], cnt[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = a == dir ? 0 : t0;
        cnt[a] = a == dir ? 17 : tn;
    }
    const int nf = cnt[0] * cnt[1] * cnt[2];
    for (int f = blockIdx.y * NT + threadIdx.x; f < nf; f += SPLIT * NT) {
        const int k = lo[2] + f % cnt[2];
        const int j = lo[1] + (f / cnt[2]) % cnt[1];
        const int i = lo[0] + f / (cnt[2] * cnt[1]);
        block_face<MHD>(b, dir, i, j, k, err);
    }
}

// One pencil of faces per thread: the faces n = 0..16 along `dir` through the
// transverse position (p1, p2), marched with a sliding window of cells so that
// every cell of the pencil is loaded once and its MC half-slopes (and, for
// hydro, its transverse velocity differences) are formed once and used by both
// of its faces -- the values prj_flux_update forms per face
// (src/prj_flux.c:373-445, 525-579, 1967-2036, 2149-2521): the same function
// of the same inputs, hence the same bits.
template <int MHD>
struct PencilCell {
    static constexpr int NHV = MHD ? 9 : 6;
    double q[NHV + 2];   // rotated primitives (local order), then P and gamma
    double d[2];         // hydro: min one-sided differences of v_t1, v_t2
};

// The loads of one pencil cell, issued a face ahead of their use: q as in
// PencilCell, and (hydro) the transverse neighbours of v_t1 and v_t2 that
// the velocity differences take.
template <int MHD>
struct PencilLoads {
    static constexpr int NHV = MHD ? 9 : 6;
    double q[NHV + 2];
    double nb[4];             // v_t1 at -/+ a1, v_t2 at -/+ a2
};

// The hydro face outputs of one wave (4 rows of 16 pencils) leave through a
// wave-private LDS tile: written [row][var][k] by the computing lanes, read
// back as 16-byte pairs (k, k+1) of one variable in the arrays' own order
// (AoSoA group, variable, slot pair), so that a wave stores each row's
// 16 x 6 fluxes as 48 16-byte pieces in increasing address order (3 store
// instructions per face for the fluxes, 2 for the face velocities) where each
// variable's per-lane store scattered over ~20 cache lines.  A row's first and
// last AoSoA group hold two active k (slots 2-3 and 0-1); the ghost slots are
// not written.  Row pitches 112 / 48 doubles put the two rows of a half-wave
// in complementary LDS banks.
constexpr int SFP = 112;   // flux tile row pitch (6 x 16 + 16)
constexpr int SVP = 48;    // face-velocity tile row pitch (3 x 16)

// Piece w (0 .. 8 nc - 1, in address order) of one row of an AoSoA array of
// nc variables over the 16 active k: its variable and the first k of its pair.
__device__ __forceinline__ void row_piece(int nc, int w, int& v, int& k0)
{
    if (w < nc) {                    // group 0: slots 2-3 = k 0, 1
        v = w;
        k0 = 0;
    } else if (w < 7 * nc) {         // groups 1..3: slot pairs 0-1 and 2-3
        const int t = w - nc;
        const int g = 1 + t / (2 * nc);
        const int rem = t % (2 * nc);
        v = rem >> 1;
        k0 = 4 * g - 2 + 2 * (rem & 1);
    } else {                         // group 4: slots 0-1 = k 14, 15
        v = w - 7 * nc;
        k0 = 14;
    }
}

// grid: (block, direction 0 or 1); hydro 256 pencils (transverse 0..15), MHD
// 324 (transverse -1..16, for the CT layer) in a 384-thread workgroup.  The
// pencil threads of a wave read rows along k, coalesced for x- and y-faces;
// z-faces, whose pencils would run along k, use block_face_kernel.
template <int MHD>
__global__ void __launch_bounds__(MHD ? 384 : 256) block_pencil_kernel(
    const fx_block_io* __restrict__ blocks, int* err)
{
    constexpr int NHV = MHD ? 9 : 6;
    constexpr int NQ = NHV + 2;
    constexpr int T = MHD ? 18 : 16;      // pencils per transverse axis
    constexpr int LO = MHD ? -1 : 0;
    // the hydro output tiles take a wave as 4 whole rows of 16 pencils
    static_assert(MHD || (T * T == NT && T == 16), "hydro pencil layout");
    const int pid = threadIdx.x;
    if (pid >= T * T) return;
    const fx_block_io* bp = blocks + blockIdx.x;
    const int dir = blockIdx.y;
    const int a1 = (dir + 1) % 3, a2 = (dir + 2) % 3;
    // this block's arrays; the per-direction pointers are read with the
    // (uniform) direction as index straight from the descriptor
    const double* __restrict__ W = bp->W;
    const int64_t vs = bp->var_stride;
    const double* __restrict__ eos = bp->eosvar;
    double* F = bp->flux[dir];
    double* vr = bp->v_riemann[dir];
    const double* bfn = MHD ? bp->bf[dir] : nullptr;
    // hydro: the wave's staging tiles; the paired stores need 16-byte
    // aligned output arrays (else each lane stores its own face)
    __shared__ __align__(16) double stg_f[MHD ? 1 : NT / 64][MHD ? 1 : 4 * SFP];
    __shared__ __align__(16) double stg_v[MHD ? 1 : NT / 64][MHD ? 1 : 4 * SVP];
    const bool paired = !MHD && ((reinterpret_cast<uintptr_t>(F) | reinterpret_cast<uintptr_t>(vr)) & 15) == 0;
    double* bv1o = MHD ? bp->Bv1[dir] : nullptr;
    double* bv2o = MHD ? bp->Bv2[dir] : nullptr;
    // transverse coordinates: the fastest-varying one is the array's fastest
    // axis where possible (k, except for z-faces, where it is j)
    const int pf = pid / T + LO, ps = pid % T + LO;
    auto ijk = [&](int n, int& i, int& j, int& k) {
        i = dir == 0 ? n : pf;
        j = dir == 0 ? pf : (dir == 1 ? n : ps);
        k = dir == 2 ? n : ps;
    };
    const int sa1 = a1 == 0 ? RB * RB : (a1 == 1 ? RB : 1);
    const int sa2 = a2 == 0 ? RB * RB : (a2 == 1 ? RB : 1);
    int gvar[NHV];
#pragma unroll
    for (int lv = 0; lv < NHV; ++lv) gvar[lv] = gmap(dir, lv);

    auto issue = [&](int n, PencilLoads<MHD>& r) {
        int i, j, k;
        ijk(n, i, j, k);
        const int o = lidx(i, j, k);
#pragma unroll
        for (int lv = 0; lv < NHV; ++lv) r.q[lv] = W[(int64_t)gvar[lv] * vs + o];
        r.q[NHV] = eos[bidx(3, 0, i, j, k)];
        r.q[NHV + 1] = eos[bidx(3, 2, i, j, k)];
        if (!MHD) {
            const double* v1 = W + (int64_t)(V1 + a1) * vs;
            const double* v2 = W + (int64_t)(V1 + a2) * vs;
            r.nb[0] = v1[o - sa1];
            r.nb[1] = v1[o + sa1];
            r.nb[2] = v2[o - sa2];
            r.nb[3] = v2[o + sa2];
        }
    };
    auto finish = [&](const PencilLoads<MHD>& r, PencilCell<MHD>& c) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) c.q[q] = r.q[q];
        if (!MHD) {
            // velocity deltas of this cell along the two transverse axes
            // (src/prj_flux.c:1967-2036): v_t1 = local 2, v_t2 = local 3
            c.d[0] = dmin_delta(c.q[2] - r.nb[0], r.nb[1] - c.q[2]);
            c.d[1] = dmin_delta(c.q[3] - r.nb[2], r.nb[3] - c.q[3]);
        }
    };
    auto load = [&](int n, PencilCell<MHD>& c) {
        PencilLoads<MHD> r;
        issue(n, r);
        finish(r, c);
    };
    // MC half-slopes of the middle cell (the B1 slot of MHD is replaced by bn)
    auto half = [&](const PencilCell<MHD>& m, const PencilCell<MHD>& c, const PencilCell<MHD>& p,
                    double* h) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) h[q] = (MHD && q == B1) ? 0.0 : mc_half(m.q[q], c.q[q], p.q[q]);
    };

    PencilCell<MHD> cm, c0, cp;      // cells n-1, n, n+1 at face n
    double hm[NQ], h0[NQ];           // half-slopes of cells n-1 and n
    PencilLoads<MHD> ahead;          // cell n+1's loads, issued at face n-1
    {
        PencilCell<MHD> cmm;
        load(-2, cmm);
        load(-1, cm);
        load(0, c0);
        issue(1, ahead);
        half(cmm, cm, c0, hm);
    }
    for (int n = 0; n <= BS; ++n) {
        finish(ahead, cp);
        // cell n+2 (the last pencil cell is 17, the block's outer ghost layer)
        if (n + 2 <= BS + 1) issue(n + 2, ahead);
        half(cm, c0, cp, h0);
        double WL[NHV], WR[NHV], Fl[NHV], vloc[3];
#pragma unroll
        for (int lv = 0; lv < NHV; ++lv) {
            WL[lv] = cm.q[lv] + hm[lv];
            WR[lv] = c0.q[lv] - h0[lv];
        }
        const double pL = cm.q[NHV] + hm[NHV], pR = c0.q[NHV] - h0[NHV];
        const double gL = cm.q[NHV + 1] + hm[NHV + 1], gR = c0.q[NHV + 1] - h0[NHV + 1];
        int i, j, k;
        ijk(n, i, j, k);
        bool ok = true;
        if (MHD) {
            const double bn = bfn[face_idx(dir, i, j, k)];
            WL[B1] = bn;
            WR[B1] = bn;
            double bv1 = 0.0, bv2 = 0.0;
            const int rc = hlld(WL, WR, pL, pR, gL, gR, bn, Fl, vloc, &bv1, &bv2);
            if (rc != 0) {
                if (atomicCAS(err, 0, rc) == 0) {
                    err[1] = dir;
                    err[2] = i;
                    err[3] = j;
                    err[4] = k;
                }
                ok = false;
            } else {
                const int c = lidx(i, j, k);
                bv1o[c] = bv1;
                bv2o[c] = bv2;
            }
        } else {
            const double du = c0.q[1] - cm.q[1];
            hllc(WL, WR, pL, pR, gL, gR, du, dmin_delta(cm.d[0], c0.d[0]),
                 dmin_delta(cm.d[1], c0.d[1]), Fl, vloc);
        }
        if (!MHD && paired) {
            const int wv = pid >> 6, lane = pid & 63, row = (pid >> 4) & 3, kk = pid & 15;
            double* sf = stg_f[MHD ? 0 : wv];
            double* sv = stg_v[MHD ? 0 : wv];
#pragma unroll
            for (int lv = 0; lv < NHV; ++lv) sf[row * SFP + gvar[lv] * 16 + kk] = Fl[lv];
#pragma unroll
            for (int r = 0; r < 3; ++r) sv[row * SVP + (gvar[1 + r] - 1) * 16 + kk] = vloc[r];
            // the tile is read by other lanes of this wave only: LDS operations
            // of a wave complete in order, the fences keep the compiler's order
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // 4 rows x 48 flux pieces, 4 rows x 24 velocity pieces
#pragma unroll 1
            for (int sidx = 0; sidx < 3; ++sidx) {
                const int c = sidx * 64 + lane;
                const int rr = c / (8 * NHV), w = c % (8 * NHV);
                int v, k0;
                row_piece(NHV, w, v, k0);
                const int prow = (pid >> 6) * 4 + rr;        // pf of row rr
                const int pi = dir == 0 ? n : prow, pj = dir == 0 ? prow : n;
                const double2 val = *reinterpret_cast<const double2*>(&sf[rr * SFP + v * 16 + k0]);
                *reinterpret_cast<double2*>(&F[bidx(NHV, v, pi, pj, k0)]) = val;
            }
#pragma unroll 1
            for (int sidx = 0; sidx < 2; ++sidx) {
                const int c = sidx * 64 + lane;
                if (c < 4 * 24) {
                    const int rr = c / 24, w = c % 24;
                    int v, k0;
                    row_piece(3, w, v, k0);
                    const int prow = (pid >> 6) * 4 + rr;
                    const int pi = dir == 0 ? n : prow, pj = dir == 0 ? prow : n;
                    const double2 val =
                        *reinterpret_cast<const double2*>(&sv[rr * SVP + v * 16 + k0]);
                    *reinterpret_cast<double2*>(&vr[bidx(3, v, pi, pj, k0)]) = val;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else if (ok) {
#pragma unroll
            for (int lv = 0; lv < NHV; ++lv) F[bidx(NHV, gvar[lv], i, j, k)] = Fl[lv];
#pragma unroll
            for (int r = 0; r < 3; ++r) vr[bidx(3, gvar[1 + r] - 1, i, j, k)] = vloc[r];
        }
        cm = c0;
        c0 = cp;
#pragma unroll
        for (int q = 0; q < NQ; ++q) hm[q] = h0[q];
    }
}

// prj_flux_div over the active cells: out[v][i][j][k]
__global__ void __launch_bounds__(NT) block_flux_div_kernel(const double* __restrict__ f0,
                                                           const double* __restrict__ f1,
                                                           const double* __restrict__ f2,
                                                           double a0, double a1, double a2,
                                                           double vol, int nvar, double* out)
{
    const int c = blockIdx.x * NT + threadIdx.x;
    if (c >= 4096) return;
    const int k = c & 15, j = (c >> 4) & 15, i = c >> 8;
    for (int v = 0; v < nvar; ++v) {
        out[v * 4096 + c] =
            -(a0 * (f0[bidx(nvar, v, i + 1, j, k)] - f0[bidx(nvar, v, i, j, k)]) +
              a1 * (f1[bidx(nvar, v, i, j + 1, k)] - f1[bidx(nvar, v, i, j, k)]) +
              a2 * (f2[bidx(nvar, v, i, j, k + 1)] - f2[bidx(nvar, v, i, j, k)])) / vol;
    }
}

// prj_mhd_bf2bc_impl (src/prj_mhd.c:555-603) over every cell of the 20^3
// block, ghosts included: cell B = face average (the metric factor is 1 off
// GR; dividing by 1.0 is the identity).  One workgroup per block.  Cells with
// a nonzero done flag are skipped (prj_mhd_bf2bc); done == null is
// prj_mhd_bf2bc_all.  A non-finite average is not stored and is reported in
// err {1, block, i, j, k} where the reference aborts.
__global__ void __launch_bounds__(NT) block_bf2bc_kernel(const fx_bf2bc_io* __restrict__ blocks,
                                                         int* err)
{
    const fx_bf2bc_io b = blocks[blockIdx.x];
    for (int c = threadIdx.x; c < RB * RB * RB; c += NT) {
        const int k = c % RB - RNG, j = (c / RB) % RB - RNG, i = c / (RB * RB) - RNG;
        if (b.done != nullptr && b.done[c] != 0) continue;
        const double b1 = 0.5 * (b.bf[0][face_idx(0, i, j, k)] + b.bf[0][face_idx(0, i + 1, j, k)]);
        const double b2 = 0.5 * (b.bf[1][face_idx(1, i, j, k)] + b.bf[1][face_idx(1, i, j + 1, k)]);
        const double b3 = 0.5 * (b.bf[2][face_idx(2, i, j, k)] + b.bf[2][face_idx(2, i, j, k + 1)]);
        if (!isfinite(b1) || !isfinite(b2) || !isfinite(b3)) {
            if (err != nullptr && atomicCAS(err, 0, 1) == 0) {
                err[1] = blockIdx.x;
                err[2] = i;
                err[3] = j;
                err[4] = k;
            }
            continue;
        }
        b.W[(int64_t)B1 * b.var_stride + c] = b1;
        b.W[(int64_t)B2 * b.var_stride + c] = b2;
        b.W[(int64_t)B3 * b.var_stride + c] = b3;
    }
}

}  // namespace

hipError_t launch_block_bf2bc(const fx_bf2bc_io* blocks_dev, int n, int* err, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(block_bf2bc_kernel, dim3(n), dim3(NT), 0, s, blocks_dev, err);
    return hipGetLastError();
}

hipError_t launch_block_flux(const fx_block_io* blocks_dev, int n, int mhd, int* err,
                             hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    // x- and y-faces as pencils, z-faces one per thread
    if (mhd) {
        hipLaunchKernelGGL(block_pencil_kernel<1>, dim3(n, 2), dim3(384), 0, s, blocks_dev, err);
        hipLaunchKernelGGL((block_face_kernel<1, 2>), dim3(n, SPLIT), dim3(NT), 0, s, blocks_dev, err);
    } else {
        hipLaunchKernelGGL(block_pencil_kernel<0>, dim3(n, 2), dim3(256), 0, s, blocks_dev, err);
        hipLaunchKernelGGL((block_face_kernel<0, 2>), dim3(n, SPLIT), dim3(NT), 0, s, blocks_dev, err);
    }
    return hipGetLastError();
}

hipError_t launch_block_flux_div(const double* f0, const double* f1, const double* f2,
                                 const double* area, double vol, int nvar, double* out,
                                 hipStream_t s)
{
    hipLaunchKernelGGL(block_flux_div_kernel, dim3(4096 / NT), dim3(NT), 0, s, f0, f1, f2,
                       area[0], area[1], area[2], vol, nvar, out);
    return hipGetLastError();
}

}  // namespace fx
