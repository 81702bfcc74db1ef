// Hydro RK2 stage on gfx950: MC reconstruction + HLLC(theta) (or HLLE) +
// flux divergence + per-cell RK2 update, fused in one kernel, one workgroup
// per 16^3 block.
//
// Reference path: prj_flux_update (src/prj_flux.c:2149-2521) for every block,
// then the per-cell stage update (src/prj_timeint.c:1255-1350 stage 1,
// :1562-1655 stage 2) with prj_flux_div (src/prj_flux.c:2523-2533) and the
// ideal prim<->cons of src/prj_eos.c:2112-2220.  The eosvar pressure and gamma
// the reference reconstructs are the ideal-EOS refill of the current W
// (src/prj_eos.c:240-261), so they are formed on the fly from W here.
//
// Work decomposition.  256 threads own the 16x16 (y,z) columns of the block
// and march along x.  Plane x=i of the block (with its 2-cell halo in y and z,
// 20x20 cells, 7 quantities rho,v1,v2,v3,e,ye,p) lives in LDS; planes i and
// i+1 are double-buffered, plus v1 of plane i-1 for the transverse velocity
// deltas.  Per plane each thread solves the y-face and z-face on the low side
// of its cell and the x-face on its high side; the 32 high-side y/z faces of
// the block (shared with the neighbour blocks) are solved by half of one wave,
// rotating through the waves plane by plane.  x-face states and fluxes are
// carried in registers along the column; y/z fluxes are exchanged through LDS
// (aliasing the dead part of plane i) to form the divergence in the
// reference's x, y, z summation order.  The next plane is prefetched into
// registers one step ahead.
#include <hip/hip_runtime.h>

#include "fx_kernels.hpp"
#include "fx_physics.hpp"

namespace fx {

namespace {

constexpr int PW = BS + 2 * NG;    // 20
constexpr int PC = PW * PW;        // 400 cells per plane incl. halo
constexpr int NQ = 7;              // rho v1 v2 v3 e ye p (x-column states)
constexpr int NV = 6;              // rho v1 v2 v3 e ye (LDS planes)
constexpr int NHALO = PC - BS * BS; // 144
constexpr int NT = 256;

struct PlaneRegs {
    double own[6];
    double halo[6];
    int fo, fh;            // sign flips still to apply (see load_plane)
};

// Loads one x-plane of the block: the thread's own cell and (has_halo) one
// halo cell, at offsets already mapped onto active cells.  fo / fh: bit a
// set where the cell stands for a ghost across a reflecting face normal to
// axis a, whose normal velocity is negated (the values the ghost fill of the
// reference would hold, src/prj_boundary.c:413-531).  The flips are applied
// where the values are consumed (flip_plane / store_plane), so that a
// prefetch does not wait for its loads.
__device__ __forceinline__ void load_plane(const double* __restrict__ W, int64_t vs,
                                           int64_t off_own, int64_t off_halo, bool has_halo,
                                           PlaneRegs& r, int fo, int fh)
{
#pragma unroll
    for (int q = 0; q < 6; ++q) r.own[q] = W[q * vs + off_own];
    if (has_halo) {
#pragma unroll
        for (int q = 0; q < 6; ++q) r.halo[q] = W[q * vs + off_halo];
    }
    r.fo = fo;
    r.fh = fh;
}

// -v where bit is set (a sign-bit xor: no branch, bit for bit -v)
__device__ __forceinline__ double negate_if(double v, int bit)
{
    return __longlong_as_double(__double_as_longlong(v) ^ ((long long)(bit & 1) << 63));
}

__device__ __forceinline__ void flip_plane(PlaneRegs& r, bool has_halo)
{
    r.own[1] = negate_if(r.own[1], r.fo);
    if (has_halo) {
#pragma unroll
        for (int q = 1; q <= 3; ++q) r.halo[q] = negate_if(r.halo[q], r.fh >> (q - 1));
    }
    r.fo = r.fh = 0;
}

__device__ __forceinline__ void store_plane(double (*S)[PC], PlaneRegs& r, int oidx,
                                            int hidx, bool has_halo)
{
    flip_plane(r, has_halo);
#pragma unroll
    for (int q = 0; q < NV; ++q) S[q][oidx] = r.own[q];
    if (has_halo) {
#pragma unroll
        for (int q = 0; q < NV; ++q) S[q][hidx] = r.halo[q];
    }
}

// prim -> cons, hydro (src/prj_eos.c:2112-2158)
__device__ __forceinline__ void prim2cons6(const double* w, double* u)
{
    double rho = w[RHO], v1 = w[V1], v2 = w[V2], v3 = w[V3];
    u[0] = rho;
    u[1] = rho * v1;
    u[2] = rho * v2;
    u[3] = rho * v3;
    u[4] = rho * w[EINT] + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3);
    u[5] = rho * w[YE];
}

// cons -> prim, hydro (src/prj_eos.c:2160-2220)
__device__ __forceinline__ void cons2prim6(const double* u, double* w)
{
    double rho = u[0];
    if (rho == 0.0) {
#pragma unroll
        for (int v = 0; v < 6; ++v) w[v] = 0.0;
        return;
    }
    double v1 = u[1] / rho, v2 = u[2] / rho, v3 = u[3] / rho;
    double kinetic = 0.5 * (v1 * v1 + v2 * v2 + v3 * v3);
    w[0] = rho;
    w[1] = v1;
    w[2] = v2;
    w[3] = v3;
    w[4] = u[4] / rho - kinetic;
    w[5] = u[5] / rho;
}

// Block index of this workgroup.  Workgroups are dealt to the 8 XCDs round
// robin; give each XCD a contiguous run of blocks so that neighbours that
// share halo cells run on the same L2.
__device__ __forceinline__ int block_of_workgroup(int wg, int nwg)
{
    if ((nwg & 7) == 0) {
        int per = nwg >> 3;
        return (wg & 7) * per + (wg >> 3);
    }
    return wg;
}

// LDS of one block's stage: planes i and i+1 (rho v1 v2 v3 e ye; pressure is
// formed from rho*e on use), v1 of planes i-1 and i (v1 of plane p in
// v1p[(p + 1) & 1]), the face fluxes that cross a wave boundary, and the
// per-cell MC half-slopes of plane i along y and z.
//
// Two barriers per plane.  B1 (after the slopes): plane i+1 stored, slopes
// written.  B2 (after the faces): fluxes exchanged, every read of plane i's
// slot done.  After B2 a wave may store plane i+2 into plane i's slot and run
// on into the next plane's slopes: the slopes read only plane i+1 and write
// hy/hz, whose readers (this plane's faces) are all behind B2; the next
// plane's faces, which read plane i+2 and write the exchange buffers that
// this plane's divergence reads, come after the next B1.  v1 of plane i is
// saved before B2 into the v1p buffer that held plane i-2.
struct HydroLds {
    double pl[2][NV][PC];
    double v1p[2][PC];
    double exup[6][4][BS];   // Fy(4(w+1), k) for the last row of wave w
    double ez16[6][BS];      // Fz(j, 16): the block's high-side z faces
    double hy[NQ][288];      // MC half-slopes along y, cells j=-1..16, k=0..15
    double hz[NQ][288];      // MC half-slopes along z, cells j=0..15, k=-1..16
};

// Riemann solver of the hydro stage: the reference's HLLC with the theta
// limiter, or HLLE (same face states and wave-speed bounds, no contact)
enum { RS_HLLC = 0, RS_HLLE = 1 };

// One RK2 stage of block (bx, by, bz) of the mesh described by g.  Called by
// the single-mesh kernel and by the batched ensemble kernels.  Stage 2 returns
// the largest CFL denominator of the cells this thread wrote (the next step's
// time-step reduction, taken while the new state is in registers); stage 1
// returns 0.  RS selects the Riemann solver (the velocity deltas feed HLLC's
// theta limiter only).
template <int STAGE, int RS = RS_HLLC>
__device__ __forceinline__ double hydro_block(HydroLds& L, const Geom& g,
                                            const double* __restrict__ Win,
                                            const double* __restrict__ W0,
                                            double* __restrict__ Wout, double dt, int bx, int by,
                                            int bz)
{
    auto& pl = L.pl;
    auto& v1p = L.v1p;
    auto& exup = L.exup;
    auto& ez16 = L.ez16;
    auto& hy = L.hy;
    auto& hz = L.hz;

    const int t = threadIdx.x;
    const int j = t >> 4;
    const int k = t & 15;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int wv = __builtin_amdgcn_readfirstlane(wave);   // the same, as a scalar
    const int x0 = bx * BS, y0 = by * BS, z0 = bz * BS;
    const int64_t vs = g.vstride;
    const double gm1 = g.gamma - 1.0;
    const double gam = g.gamma;
    double cfl_best = 0.0;

    // own cell and (for t < 144) one halo cell of the 20x20 plane
    const int oidx = (j + NG) * PW + (k + NG);
    const bool has_halo = t < NHALO;
    int hj, hk;
    if (t < 80) {
        int r = t / PW;
        hj = r < 2 ? r : r + BS;
        hk = t - r * PW;
    } else {
        int h = t - 80;
        int c = h & 3;
        hj = NG + (h >> 2);
        hk = c < 2 ? c : c + BS;
    }
    const int hidx = hj * PW + hk;
    // Cells outside the domain are read where they come from (periodic wrap
    // or mirror, map_axis), so the stage needs no ghost fill: y/z of the halo
    // cell once per thread, x once per plane (uniform over the workgroup).
    int hflip = 0, hym = 0, hzm = 0;
    if (has_halo) {
        hym = map_axis(g, 1, g.ny, y0 + hj - NG, hflip);
        hzm = map_axis(g, 2, g.nz, z0 + hk - NG, hflip);
    }
    const int64_t col_own = cell_off(g, 0, y0 + j, z0 + k);
    const int64_t col_halo = cell_off(g, 0, hym, hzm);

    // active planes (p in [0, BS)): the update and W^n
    auto own_at = [&](int p) { return col_own + (int64_t)(x0 + p) * g.sx; };
    // plane p of the stencil, mapped; fx receives its x flip bit
    auto plane_x = [&](int p, int& fx) {
        fx = 0;
        return (int64_t)map_axis(g, 0, g.nx, x0 + p, fx) * g.sx;
    };
    auto load_at = [&](int p, PlaneRegs& r) {
        int fx;
        const int64_t xo = plane_x(p, fx);
        load_plane(Win, vs, col_own + xo, col_halo + xo, has_halo, r, fx, fx | hflip);
    };

    double Ls[NQ];             // left state of x-face i+1 (rho..ye, p)
    double Fx[6];              // flux through x-face i
    PlaneRegs pf;              // prefetched plane i+2

    // ---- prologue: planes -2..2 -----------------------------------------
    {
        double cmm[NQ], cm[NQ], c0[NQ], c1[NQ];
        PlaneRegs r;
        load_at(-1, r);
        store_plane(pl[1], r, oidx, hidx, has_halo);        // plane -1 (temporary)
        v1p[0][oidx] = r.own[1];                            // plane -1
        if (has_halo) v1p[0][hidx] = r.halo[1];
#pragma unroll
        for (int q = 0; q < NV; ++q) cm[q] = r.own[q];
        load_at(0, r);
        store_plane(pl[0], r, oidx, hidx, has_halo);        // plane 0
        int fmm, f1;
        const int64_t omm = col_own + plane_x(-2, fmm), o1 = col_own + plane_x(1, f1);
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            c0[q] = r.own[q];
            cmm[q] = Win[q * vs + omm];
            c1[q] = Win[q * vs + o1];
        }
        cmm[1] = negate_if(cmm[1], fmm);
        c1[1] = negate_if(c1[1], f1);
        cmm[6] = gm1 * cmm[0] * cmm[4];
        cm[6] = gm1 * cm[0] * cm[4];
        c0[6] = gm1 * c0[0] * c0[4];
        c1[6] = gm1 * c1[0] * c1[4];
        __syncthreads();

        // x-face 0: L from cell -1, R from cell 0
        double WL[6], WR[6], pL, pR;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double hl = mc_half(cmm[q], cm[q], c0[q]);
            double h0 = mc_half(cm[q], c0[q], c1[q]);
            if (q < 6) { WL[q] = cm[q] + hl; WR[q] = c0[q] - h0; }
            else { pL = cm[q] + hl; pR = c0[q] - h0; }
            Ls[q] = c0[q] + h0;
        }
        const double* A2 = pl[1][2];
        const double* A3 = pl[1][3];
        const double* B2 = pl[0][2];
        const double* B3 = pl[0][3];
        const int c = oidx;
        double du = c0[1] - cm[1];
        double dv = dmin_delta(dmin_delta(A2[c] - A2[c - PW], A2[c + PW] - A2[c]),
                         dmin_delta(B2[c] - B2[c - PW], B2[c + PW] - B2[c]));
        double dw = dmin_delta(dmin_delta(A3[c] - A3[c - 1], A3[c + 1] - A3[c]),
                         dmin_delta(B3[c] - B3[c - 1], B3[c + 1] - B3[c]));
        double vf[3];
        if (RS == RS_HLLE) hlle_hydro(WL, WR, pL, pR, gam, gam, Fx, vf);
        else hllc(WL, WR, pL, pR, gam, gam, du, dv, dw, Fx, vf);
        __syncthreads();
        load_at(1, r);
        store_plane(pl[1], r, oidx, hidx, has_halo);        // plane 1
        load_at(2, pf);
        __syncthreads();
    }

    const double a1 = g.area[0], a2 = g.area[1], a3 = g.area[2], vol = g.vol;
    const double inv_vol = g.inv_vol;
    const bool vol_pow2 = g.vol_pow2 != 0;
    // stage 2: W^n of the own cell of the next plane, loaded an iteration
    // ahead of the update that blends with it
    double w0n[6];
    if (STAGE == 2) {
#pragma unroll
        for (int q = 0; q < 6; ++q) w0n[q] = W0[q * vs + own_at(0)];
    }

    for (int i = 0; i < BS; ++i) {
        const int A = i & 1;
        const int B = A ^ 1;
        // the wave with one pass more than the others before the next barrier
        // (halo slopes here, the extra faces below) issues first on its SIMD
        if (wv == ((i + 1) & 3)) __builtin_amdgcn_s_setprio(2);

        // ---- MC half-slopes of plane i along y and z, once per cell ---------
        // (the reference sweeps each variable once per direction and emits both
        // faces of a cell from one slope, src/prj_flux.c:373-445)
        {
            const double* R = pl[A][0];
            const double* E = pl[A][4];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                double qm_y, qp_y, qm_z, qp_z, q0;
                if (q < NV) {
                    const double* P = pl[A][q];
                    q0 = P[oidx];
                    qm_y = P[oidx - PW]; qp_y = P[oidx + PW];
                    qm_z = P[oidx - 1]; qp_z = P[oidx + 1];
                } else {
                    q0 = gm1 * R[oidx] * E[oidx];
                    qm_y = gm1 * R[oidx - PW] * E[oidx - PW];
                    qp_y = gm1 * R[oidx + PW] * E[oidx + PW];
                    qm_z = gm1 * R[oidx - 1] * E[oidx - 1];
                    qp_z = gm1 * R[oidx + 1] * E[oidx + 1];
                }
                hy[q][(j + 1) * BS + k] = mc_half(qm_y, q0, qp_y);
                hz[q][j * (BS + 2) + k + 1] = mc_half(qm_z, q0, qp_z);
            }
            // the 64 halo cells whose slopes the block's boundary faces need:
            // y: rows -1 and 16; z: columns -1 and 16 (one wave, one cell per lane)
            if (wave == ((i + 1) & 3)) {
                const int side = lane >> 4;          // 0,1: y rows -1,16; 2,3: z cols -1,16
                const int m = lane & 15;
                int cj, ck, S;
                if (side < 2) { cj = side == 0 ? -1 : BS; ck = m; S = PW; }
                else { cj = m; ck = side == 2 ? -1 : BS; S = 1; }
                const int c = (cj + NG) * PW + (ck + NG);
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    double a0, a1v, a2v;
                    if (q < NV) {
                        const double* P = pl[A][q];
                        a0 = P[c - S]; a1v = P[c]; a2v = P[c + S];
                    } else {
                        a0 = gm1 * R[c - S] * E[c - S];
                        a1v = gm1 * R[c] * E[c];
                        a2v = gm1 * R[c + S] * E[c + S];
                    }
                    double h = mc_half(a0, a1v, a2v);
                    if (side < 2) hy[q][(cj + 1) * BS + ck] = h;
                    else hz[q][cj * (BS + 2) + ck + 1] = h;
                }
            }
        }
        __syncthreads();   // B1: slopes written, plane i+1 stored

        // ---- faces: 0 = own y-face (low side), 1 = own z-face (low side),
        //      2 = extra high-side face (half of one wave), 3 = x-face i+1 ----
        double Fy[6], Fz[6], Fxn[6];
        const int xw = i & 3;
        const bool extra = wave == xw && lane < 32;
        __builtin_amdgcn_s_setprio(0);
        if (wv == xw) __builtin_amdgcn_s_setprio(2);
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            if (f == 2 && !extra) continue;
            double WL[6], WR[6], pL, pR, du, dv, dw;
            int dir, fj = j, fk = k;
            if (f == 3) {
                // x-face between own cells i (slot A) and i+1 (slot B); cell i+2 is pf
                double cm[NQ], c0[NQ], cp[NQ];
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    cm[q] = pl[A][q][oidx];
                    c0[q] = pl[B][q][oidx];
                    cp[q] = pf.own[q];
                }
                cp[1] = negate_if(cp[1], pf.fo);   // pf is flipped when stored
                cm[6] = gm1 * cm[0] * cm[4];
                c0[6] = gm1 * c0[0] * c0[4];
                cp[6] = gm1 * cp[0] * cp[4];
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    double h = mc_half(cm[q], c0[q], cp[q]);
                    double r = c0[q] - h;
                    if (q < 6) { WL[q] = Ls[q]; WR[q] = r; }
                    else { pL = Ls[6]; pR = r; }
                    Ls[q] = c0[q] + h;
                }
                const double* A2 = pl[A][2];
                const double* A3 = pl[A][3];
                const double* B2 = pl[B][2];
                const double* B3 = pl[B][3];
                const int c = oidx;
                du = c0[1] - cm[1];
                dv = dmin_delta(dmin_delta(A2[c] - A2[c - PW], A2[c + PW] - A2[c]),
                                dmin_delta(B2[c] - B2[c - PW], B2[c + PW] - B2[c]));
                dw = dmin_delta(dmin_delta(A3[c] - A3[c - 1], A3[c + 1] - A3[c]),
                                dmin_delta(B3[c] - B3[c - 1], B3[c + 1] - B3[c]));
                dir = 0;
            } else {
                if (f == 0) dir = 1;
                else if (f == 1) dir = 2;
                else if (lane < 16) { dir = 1; fj = BS; fk = lane; }
                else { dir = 2; fj = lane - 16; fk = BS; }
                const int c = (fj + NG) * PW + (fk + NG);   // right cell
                const int S = dir == 1 ? PW : 1;             // stride along the face normal
                const int T = dir == 1 ? 1 : PW;             // in-plane transverse stride
                const int cl = c - S;
                // slope-plane indices of the left and right cells
                const int hr = dir == 1 ? (fj + 1) * BS + fk : fj * (BS + 2) + fk + 1;
                const int hl = dir == 1 ? hr - BS : hr - 1;
                const double (*H)[288] = dir == 1 ? hy : hz;
                const int qn = dir == 1 ? 2 : 3;
                const int qt1 = dir == 1 ? 3 : 1;
                const int qt2 = dir == 1 ? 1 : 2;
                const int qmap[6] = {0, qn, qt1, qt2, 4, 5};
#pragma unroll
                for (int lv = 0; lv < 6; ++lv) {
                    const int q = qmap[lv];
                    const double* P = pl[A][q];
                    WL[lv] = P[cl] + H[q][hl];
                    WR[lv] = P[c] - H[q][hr];
                }
                {
                    const double* R = pl[A][0];
                    const double* E = pl[A][4];
                    pL = gm1 * R[cl] * E[cl] + H[6][hl];
                    pR = gm1 * R[c] * E[c] - H[6][hr];
                }
                // velocity deltas (src/prj_flux.c:1967-2036)
                const double* Vn = pl[A][qn];
                du = Vn[c] - Vn[cl];
                const double* Xm = v1p[i & 1];
                const double* Xc = pl[A][1];
                const double* Xp = pl[B][1];
                const double* Vt = pl[A][dir == 1 ? 3 : 2];
                double din = dmin_delta(dmin_delta(Vt[cl] - Vt[cl - T], Vt[cl + T] - Vt[cl]),
                                        dmin_delta(Vt[c] - Vt[c - T], Vt[c + T] - Vt[c]));
                double dx1 = dmin_delta(dmin_delta(Xc[cl] - Xm[cl], Xp[cl] - Xc[cl]),
                                        dmin_delta(Xc[c] - Xm[c], Xp[c] - Xc[c]));
                // y-face: dv = v3 along z (in plane), dw = v1 along x;
                // z-face: dv = v1 along x, dw = v2 along y (in plane)
                dv = dir == 1 ? din : dx1;
                dw = dir == 1 ? dx1 : din;
            }
            double Fl[6], vf[3];
            if (RS == RS_HLLE) hlle_hydro(WL, WR, pL, pR, gam, gam, Fl, vf);
            else hllc(WL, WR, pL, pR, gam, gam, du, dv, dw, Fl, vf);
            // back to the global variable order (src/prj_flux.c:609-690)
            double G[6];
            G[0] = Fl[0];
            G[1] = dir == 0 ? Fl[1] : (dir == 1 ? Fl[3] : Fl[2]);
            G[2] = dir == 0 ? Fl[2] : (dir == 1 ? Fl[1] : Fl[3]);
            G[3] = dir == 0 ? Fl[3] : (dir == 1 ? Fl[2] : Fl[1]);
            G[4] = Fl[4];
            G[5] = Fl[5];
            if (f == 0) {
#pragma unroll
                for (int v = 0; v < 6; ++v) Fy[v] = G[v];
                // first row of waves 1..3 is the high face of the previous wave's last row
                if (wave > 0 && lane < BS) {
#pragma unroll
                    for (int v = 0; v < 6; ++v) exup[v][wave - 1][k] = G[v];
                }
            } else if (f == 1) {
#pragma unroll
                for (int v = 0; v < 6; ++v) Fz[v] = G[v];
            } else if (f == 2) {
                if (dir == 1) {
#pragma unroll
                    for (int v = 0; v < 6; ++v) exup[v][3][fk] = G[v];
                } else {
#pragma unroll
                    for (int v = 0; v < 6; ++v) ez16[v][fj] = G[v];
                }
            } else {
#pragma unroll
                for (int v = 0; v < 6; ++v) Fxn[v] = G[v];
            }
        }
        // own cell of plane i for the update, and v1 of plane i for the next
        // plane's faces (slot A is overwritten after B2)
        double wi[6];
#pragma unroll
        for (int q = 0; q < NV; ++q) wi[q] = pl[A][q][oidx];
        v1p[(i + 1) & 1][oidx] = wi[1];
        if (has_halo) v1p[(i + 1) & 1][hidx] = pl[A][1][hidx];
        __syncthreads();   // B2: cross-wave fluxes written; reads of slot A, v1p, slopes done
        __builtin_amdgcn_s_setprio(0);

        // ---- divergence in the reference's x, y, z order ---------------------
        double fs[6];
#pragma unroll
        for (int v = 0; v < 6; ++v) {
            double fyu = __shfl_down(Fy[v], BS);
            double fzu = __shfl_down(Fz[v], 1);
            if (lane >= 48) fyu = exup[v][wave][k];
            if (k == BS - 1) fzu = ez16[v][j];
            fs[v] = -(a1 * (Fxn[v] - Fx[v]) + a2 * (fyu - Fy[v]) + a3 * (fzu - Fz[v]));
        }
        double fd[6];
        if (vol_pow2) {
#pragma unroll
            for (int v = 0; v < 6; ++v) fd[v] = fs[v] * inv_vol;
        } else {
#pragma unroll
            for (int v = 0; v < 6; ++v) fd[v] = fs[v] / vol;
        }
        if (i < BS - 1) store_plane(pl[A], pf, oidx, hidx, has_halo);

        // ---- per-cell RK2 update -------------------------------------------
        {
            const int64_t o = own_at(i);
            double u[6], u1[6], w[6];
            if (STAGE == 1) {
                prim2cons6(wi, u);
#pragma unroll
                for (int v = 0; v < 6; ++v) u1[v] = u[v] + dt * (0.0 + fd[v]);
            } else {
                double us[6];
                prim2cons6(w0n, u);
                prim2cons6(wi, us);
#pragma unroll
                for (int v = 0; v < 6; ++v) {
                    double ut = us[v] + dt * (0.0 + fd[v]);
                    u1[v] = 0.5 * u[v] + 0.5 * ut;
                }
            }
            cons2prim6(u1, w);
#pragma unroll
            for (int v = 0; v < 6; ++v) __builtin_nontemporal_store(w[v], &Wout[v * vs + o]);
            if (STAGE == 2) {
                const double d = cfl_denom(g, w[0], w[1], w[2], w[3], w[4]);
                if (d > cfl_best) cfl_best = d;   // NaN never wins
            }
        }

#pragma unroll
        for (int q = 0; q < 6; ++q) Fx[q] = Fxn[q];
        if (i + 3 <= BS + 1) load_at(i + 3, pf);
        if (STAGE == 2 && i + 1 < BS) {
#pragma unroll
            for (int q = 0; q < 6; ++q) w0n[q] = W0[q * vs + own_at(i + 1)];
        }
    }
    return cfl_best;
}

template <int STAGE, int RS>
__global__ void __launch_bounds__(NT, 2) hydro_stage_kernel(HydroStageArgs a)
{
    __shared__ HydroLds L;
    const Geom& g = a.g;
    const int b = block_of_workgroup(blockIdx.x, gridDim.x);
    const int bz = b % g.nbz;
    const int by = (b / g.nbz) % g.nby;
    const int bx = b / (g.nbz * g.nby);
    const double dt = a.dt_ptr ? *a.dt_ptr : a.dt;
    const double d = hydro_block<STAGE, RS>(L, g, a.Win, a.W0, a.Wout, dt, bx, by, bz);
    if (STAGE == 2 && a.cfl_part != nullptr) {
        const double m = workgroup_max(d, L.v1p[0]);
        if (threadIdx.x == 0) a.cfl_part[b] = m;
    }
}

// One table entry of a batched stage: block (bx, by, bz) of instance `inst`.
template <int STAGE>
__device__ __forceinline__ void ens_entry(HydroLds& L, const EnsInst* __restrict__ inst,
                                          const EnsBlock* __restrict__ blocks, int b)
{
    const EnsBlock e = blocks[b];
    const EnsInst& I = inst[e.inst];
    const double dt = I.dev[0];
    const double d = hydro_block<STAGE>(L, I.g, I.W[STAGE == 1 ? 0 : 1], I.W[0],
                                        I.W[STAGE == 1 ? 1 : 0], dt, e.bx, e.by, e.bz);
    if (STAGE == 2) {
        // dev[2] holds the bits of the instance's largest denominator
        // (non-negative doubles order as their bit patterns); the step clears
        // it before this launch
        // every thread holds the maximum; a lane of wave 1 raises it, so that
        // in the persistent kernel this atomic and the next entry's queue
        // fetch (thread 0, wave 0) never share a wave and the compiler cannot
        // merge the two single-lane regions across the loop's back edge
        const double m = workgroup_max(d, L.v1p[0]);
        if (threadIdx.x == 64)
            atomicMax(reinterpret_cast<unsigned long long*>(I.dev + 2),
                      (unsigned long long)__double_as_longlong(m));
    }
}

// Batched stage over a table of (instance, block) entries, one workgroup per
// entry; XCD_REMAP gives every XCD a contiguous run of the table.
template <int STAGE, int XCD_REMAP>
__global__ void __launch_bounds__(NT, 2) ens_stage_kernel(const EnsInst* __restrict__ inst,
                                                      const EnsBlock* __restrict__ blocks)
{
    __shared__ HydroLds L;
    const int b = XCD_REMAP ? block_of_workgroup(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    ens_entry<STAGE>(L, inst, blocks, b);
}

// Persistent batched stage: a fixed number of workgroups pull table entries
// from atomic counters until the table is exhausted.  With XCD_QUEUES the
// table is cut into 8 contiguous chunks, one counter each, and workgroup w
// (dispatched to XCD w mod 8) drains chunk w mod 8 first, then the others.
// Every workgroup leaves once all counters have passed their chunk ends.
template <int STAGE, int XCD_QUEUES>
__global__ void __launch_bounds__(NT, 2) ens_stage_persistent_kernel(const EnsInst* __restrict__ inst,
                                                                 const EnsBlock* __restrict__ blocks,
                                                                 int nblocks, int* queues)
{
    __shared__ HydroLds L;
    __shared__ int s_b;
    const int nq = XCD_QUEUES ? 8 : 1;
    const int per = (nblocks + nq - 1) / nq;
    for (int r = 0; r < nq; ++r) {
        const int q = XCD_QUEUES ? (int)((blockIdx.x + r) & 7) : 0;
        const int lo = q * per;
        const int hi = min(nblocks, lo + per);
        for (;;) {
            if (threadIdx.x == 0) s_b = lo + atomicAdd(&queues[q], 1);
            __syncthreads();
            const int b = __builtin_amdgcn_readfirstlane(s_b);
            __syncthreads();
            if (b >= hi) break;
            ens_entry<STAGE>(L, inst, blocks, b);
            // The entry's stage-2 epilogue raises the CFL maximum from wave 1
            // (thread 64), never from the fetching lane; this barrier also
            // keeps the two apart.  (An earlier build had both on thread 0 and
            // no barrier here: the compiler merged the two thread-0 regions
            // across the back edge and parked lane 0 outside the loop while
            // the rest of its wave went round again -- a hang.)
            __syncthreads();
        }
    }
}

__global__ void queue_reset_kernel(int* q)
{
    if (threadIdx.x < 8) q[threadIdx.x] = 0;
}

}  // namespace

hipError_t launch_hydro_stage(const HydroStageArgs& a, int stage, hipStream_t stream)
{
    const int nblocks = a.g.nbx * a.g.nby * a.g.nbz;
    const bool hlle = a.riemann == FX_RIEMANN_HLLE;
    if (stage == 1) {
        if (hlle) hipLaunchKernelGGL((hydro_stage_kernel<1, RS_HLLE>), dim3(nblocks), dim3(NT), 0, stream, a);
        else hipLaunchKernelGGL((hydro_stage_kernel<1, RS_HLLC>), dim3(nblocks), dim3(NT), 0, stream, a);
    } else {
        if (hlle) hipLaunchKernelGGL((hydro_stage_kernel<2, RS_HLLE>), dim3(nblocks), dim3(NT), 0, stream, a);
        else hipLaunchKernelGGL((hydro_stage_kernel<2, RS_HLLC>), dim3(nblocks), dim3(NT), 0, stream, a);
    }
    return hipGetLastError();
}


hipError_t launch_ens_stage(const EnsInst* inst, const EnsBlock* blocks, int nblocks, int stage,
                            int persistent_wgs, int xcd_remap, int xcd_queues, int* queues,
                            hipStream_t s)
{
    if (nblocks <= 0) return hipSuccess;
    if (persistent_wgs <= 0) {
        if (stage == 1) {
            if (xcd_remap) hipLaunchKernelGGL((ens_stage_kernel<1, 1>), dim3(nblocks), dim3(NT), 0, s, inst, blocks);
            else hipLaunchKernelGGL((ens_stage_kernel<1, 0>), dim3(nblocks), dim3(NT), 0, s, inst, blocks);
        } else {
            if (xcd_remap) hipLaunchKernelGGL((ens_stage_kernel<2, 1>), dim3(nblocks), dim3(NT), 0, s, inst, blocks);
            else hipLaunchKernelGGL((ens_stage_kernel<2, 0>), dim3(nblocks), dim3(NT), 0, s, inst, blocks);
        }
        return hipGetLastError();
    }
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, s, queues);
    const dim3 grid(persistent_wgs);
    if (stage == 1) {
        if (xcd_queues) hipLaunchKernelGGL((ens_stage_persistent_kernel<1, 1>), grid, dim3(NT), 0, s, inst, blocks, nblocks, queues);
        else hipLaunchKernelGGL((ens_stage_persistent_kernel<1, 0>), grid, dim3(NT), 0, s, inst, blocks, nblocks, queues);
    } else {
        if (xcd_queues) hipLaunchKernelGGL((ens_stage_persistent_kernel<2, 1>), grid, dim3(NT), 0, s, inst, blocks, nblocks, queues);
        else hipLaunchKernelGGL((ens_stage_persistent_kernel<2, 0>), grid, dim3(NT), 0, s, inst, blocks, nblocks, queues);
    }
    return hipGetLastError();
}

}  // namespace fx
