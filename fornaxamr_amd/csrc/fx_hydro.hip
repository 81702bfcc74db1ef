// Hydro RK2 stage on gfx950: MC reconstruction + HLLC(theta) + flux
// divergence + per-cell RK2 update, fused in one kernel, one workgroup per
// 16^3 block.
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
constexpr int NQ = 7;              // rho v1 v2 v3 e ye p
constexpr int NHALO = PC - BS * BS; // 144
constexpr int NT = 256;

struct PlaneRegs {
    double own[6];
    double halo[6];
};

__device__ __forceinline__ void load_plane(const double* __restrict__ W, int64_t vs,
                                           int64_t off_own, int64_t off_halo, bool has_halo,
                                           PlaneRegs& r)
{
#pragma unroll
    for (int q = 0; q < 6; ++q) r.own[q] = W[q * vs + off_own];
    if (has_halo) {
#pragma unroll
        for (int q = 0; q < 6; ++q) r.halo[q] = W[q * vs + off_halo];
    }
}

__device__ __forceinline__ void store_plane(double (*S)[PC], const PlaneRegs& r, double gm1,
                                            int oidx, int hidx, bool has_halo)
{
#pragma unroll
    for (int q = 0; q < 6; ++q) S[q][oidx] = r.own[q];
    S[6][oidx] = gm1 * r.own[0] * r.own[4];
    if (has_halo) {
#pragma unroll
        for (int q = 0; q < 6; ++q) S[q][hidx] = r.halo[q];
        S[6][hidx] = gm1 * r.halo[0] * r.halo[4];
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

template <int STAGE>
__global__ void __launch_bounds__(NT) hydro_stage_kernel(HydroStageArgs a)
{
    __shared__ double pl[2][NQ][PC];
    __shared__ double v1p[PC];
    __shared__ double exy[6][(BS + 1) * BS];    // y-face fluxes of plane i
    __shared__ double exz[6][BS * (BS + 1)];    // z-face fluxes of plane i

    const Geom& g = a.g;
    const int t = threadIdx.x;
    const int j = t >> 4;
    const int k = t & 15;
    const int b = block_of_workgroup(blockIdx.x, gridDim.x);
    const int bz = b % g.nbz;
    const int by = (b / g.nbz) % g.nby;
    const int bx = b / (g.nbz * g.nby);
    const int x0 = bx * BS, y0 = by * BS, z0 = bz * BS;
    const int64_t vs = g.vstride;
    const double gm1 = g.gamma - 1.0;
    const double gam = g.gamma;
    const double dt = a.dt_ptr ? *a.dt_ptr : a.dt;

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
    const int64_t col_own = cell_off(g, x0, y0 + j, z0 + k);
    const int64_t col_halo = cell_off(g, x0, y0 + hj - NG, z0 + hk - NG);

    auto own_at = [&](int p) { return col_own + (int64_t)p * g.sx; };
    auto halo_at = [&](int p) { return col_halo + (int64_t)p * g.sx; };

    // ---- prologue: planes -2..2 -----------------------------------------
    double cm[NQ], c0[NQ];     // column cells i and i+1 (q incl. p)
    double Ls[NQ];             // left state of x-face i+1
    double Fx[6];              // flux through x-face i
    PlaneRegs pf;              // prefetched plane i+2

    {
        double cmm[NQ];
        PlaneRegs r;
        // plane -1 -> slot 1 (full) and v1p; plane 0 -> slot 0
        load_plane(a.Win, vs, own_at(-1), halo_at(-1), has_halo, r);
        store_plane(pl[1], r, gm1, oidx, hidx, has_halo);
        v1p[oidx] = r.own[1];
        if (has_halo) v1p[hidx] = r.halo[1];
#pragma unroll
        for (int q = 0; q < 6; ++q) cm[q] = r.own[q];   // cm temporarily = cell -1
        cm[6] = gm1 * cm[0] * cm[4];
        load_plane(a.Win, vs, own_at(0), halo_at(0), has_halo, r);
        store_plane(pl[0], r, gm1, oidx, hidx, has_halo);
#pragma unroll
        for (int q = 0; q < 6; ++q) c0[q] = r.own[q];   // c0 temporarily = cell 0
        c0[6] = gm1 * c0[0] * c0[4];
        double c1[NQ];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            cmm[q] = a.Win[q * vs + own_at(-2)];
            c1[q] = a.Win[q * vs + own_at(1)];
        }
        cmm[6] = gm1 * cmm[0] * cmm[4];
        c1[6] = gm1 * c1[0] * c1[4];
        __syncthreads();

        // x-face 0: L from cell -1, R from cell 0
        double WL[6], WR[6], pL, pR;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            WL[q] = cm[q] + mc_half(cmm[q], cm[q], c0[q]);
            double h0 = mc_half(cm[q], c0[q], c1[q]);
            WR[q] = c0[q] - h0;
            Ls[q] = c0[q] + h0;
        }
        pL = cm[6] + mc_half(cmm[6], cm[6], c0[6]);
        {
            double h0 = mc_half(cm[6], c0[6], c1[6]);
            pR = c0[6] - h0;
            Ls[6] = c0[6] + h0;
        }
        // deltas: L cell in plane -1 (slot 1), R cell in plane 0 (slot 0)
        const double* A2 = pl[1][2];
        const double* A3 = pl[1][3];
        const double* B2 = pl[0][2];
        const double* B3 = pl[0][3];
        const int c = oidx;
        double du = c0[1] - cm[1];
        double dv = dmin(dmin(A2[c] - A2[c - PW], A2[c + PW] - A2[c]),
                         dmin(B2[c] - B2[c - PW], B2[c + PW] - B2[c]));
        double dw = dmin(dmin(A3[c] - A3[c - 1], A3[c + 1] - A3[c]),
                         dmin(B3[c] - B3[c - 1], B3[c + 1] - B3[c]));
        double vf[3];
        hllc(WL, WR, pL, pR, gam, gam, du, dv, dw, Fx, vf);

        // window for step 0: cm = cell 0, c0 = cell 1
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            cm[q] = c0[q];
            c0[q] = c1[q];
        }
        __syncthreads();
        load_plane(a.Win, vs, own_at(1), halo_at(1), has_halo, r);
        store_plane(pl[1], r, gm1, oidx, hidx, has_halo);
        load_plane(a.Win, vs, own_at(2), halo_at(2), has_halo, pf);
        __syncthreads();
    }

    const double a1 = g.area[0], a2 = g.area[1], a3 = g.area[2], vol = g.vol;
    constexpr int EX = (BS + 1) * BS;        // 272 faces per direction and plane

    for (int i = 0; i < BS; ++i) {
        const int A = i & 1;
        const int B = A ^ 1;

        // ---- faces: 0 = own y-face (low side), 1 = own z-face (low side),
        //      2 = extra high-side face (one wave), 3 = x-face i+1.  y/z fluxes
        //      go straight to the LDS exchange buffers. ----------------------
        double Fxn[6];
        const int xw = i & 3;                  // wave solving the 32 extra faces
        const int lane = t & 63;
        const bool extra = (t >> 6) == xw && lane < 32;
        for (int f = 0; f < 4; ++f) {
            if (f == 2 && !extra) continue;
            double WL[6], WR[6], pL, pR, du, dv, dw;
            int dir, fj = j, fk = k;
            if (f == 3) {
                // x-face between own cells i (cm) and i+1 (c0); cell i+2 is pf
                double cp[NQ];
#pragma unroll
                for (int q = 0; q < 6; ++q) cp[q] = pf.own[q];
                cp[6] = gm1 * cp[0] * cp[4];
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    double h = mc_half(cm[q], c0[q], cp[q]);
                    double r = c0[q] - h;
                    double l = c0[q] + h;
                    if (q < 6) { WL[q] = Ls[q]; WR[q] = r; }
                    else { pL = Ls[6]; pR = r; }
                    Ls[q] = l;
                }
                const double* A2 = pl[A][2];
                const double* A3 = pl[A][3];
                const double* B2 = pl[B][2];
                const double* B3 = pl[B][3];
                const int c = oidx;
                du = c0[1] - cm[1];
                dv = dmin(dmin(A2[c] - A2[c - PW], A2[c + PW] - A2[c]),
                          dmin(B2[c] - B2[c - PW], B2[c + PW] - B2[c]));
                dw = dmin(dmin(A3[c] - A3[c - 1], A3[c + 1] - A3[c]),
                          dmin(B3[c] - B3[c - 1], B3[c + 1] - B3[c]));
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
                // local order: rho, vn, vt1, vt2, e, ye
                const int qn = dir == 1 ? 2 : 3;
                const int qt1 = dir == 1 ? 3 : 1;
                const int qt2 = dir == 1 ? 1 : 2;
                const int qmap[6] = {0, qn, qt1, qt2, 4, 5};
#pragma unroll
                for (int lv = 0; lv < 6; ++lv) {
                    const double* P = pl[A][qmap[lv]];
                    double q0 = P[c - 2 * S], q1 = P[cl], q2 = P[c], q3 = P[c + S];
                    WL[lv] = q1 + mc_half(q0, q1, q2);
                    WR[lv] = q2 - mc_half(q1, q2, q3);
                }
                {
                    const double* P = pl[A][6];
                    double q0 = P[c - 2 * S], q1 = P[cl], q2 = P[c], q3 = P[c + S];
                    pL = q1 + mc_half(q0, q1, q2);
                    pR = q2 - mc_half(q1, q2, q3);
                }
                // velocity deltas (src/prj_flux.c:1967-2036)
                const double* Vn = pl[A][qn];
                du = Vn[c] - Vn[cl];
                const double* Xm = v1p;
                const double* Xc = pl[A][1];
                const double* Xp = pl[B][1];
                const double* Vt = pl[A][dir == 1 ? 3 : 2];   // in-plane transverse velocity
                double din = dmin(dmin(Vt[cl] - Vt[cl - T], Vt[cl + T] - Vt[cl]),
                                  dmin(Vt[c] - Vt[c - T], Vt[c + T] - Vt[c]));
                double dx1 = dmin(dmin(Xc[cl] - Xm[cl], Xp[cl] - Xc[cl]),
                                  dmin(Xc[c] - Xm[c], Xp[c] - Xc[c]));
                // y-face: dv = v3 along z (in plane), dw = v1 along x;
                // z-face: dv = v1 along x, dw = v2 along y (in plane)
                dv = dir == 1 ? din : dx1;
                dw = dir == 1 ? dx1 : din;
            }
            double Fl[6], vf[3];
            hllc(WL, WR, pL, pR, gam, gam, du, dv, dw, Fl, vf);
            // back to the global variable order (src/prj_flux.c:609-690)
            double G[6];
            G[0] = Fl[0];
            G[1] = dir == 0 ? Fl[1] : (dir == 1 ? Fl[3] : Fl[2]);
            G[2] = dir == 0 ? Fl[2] : (dir == 1 ? Fl[1] : Fl[3]);
            G[3] = dir == 0 ? Fl[3] : (dir == 1 ? Fl[2] : Fl[1]);
            G[4] = Fl[4];
            G[5] = Fl[5];
            if (f == 3) {
#pragma unroll
                for (int v = 0; v < 6; ++v) Fxn[v] = G[v];
            } else if (dir == 1) {
#pragma unroll
                for (int v = 0; v < 6; ++v) exy[v][fj * BS + fk] = G[v];
            } else {
#pragma unroll
                for (int v = 0; v < 6; ++v) exz[v][fj * (BS + 1) + fk] = G[v];
            }
        }
        __syncthreads();   // fluxes written; all reads of plane i and v1p done

        // ---- divergence in the reference's x, y, z order ---------------------
        double fd[6];
#pragma unroll
        for (int v = 0; v < 6; ++v) {
            fd[v] = -(a1 * (Fxn[v] - Fx[v]) + a2 * (exy[v][(j + 1) * BS + k] - exy[v][j * BS + k]) +
                      a3 * (exz[v][j * (BS + 1) + k + 1] - exz[v][j * (BS + 1) + k])) / vol;
        }
        v1p[oidx] = pl[A][1][oidx];
        if (has_halo) v1p[hidx] = pl[A][1][hidx];
        __syncthreads();   // v1 of plane i saved; slot A is free for plane i+2
        if (i < BS - 1) store_plane(pl[A], pf, gm1, oidx, hidx, has_halo);

        // ---- per-cell RK2 update -------------------------------------------
        {
            const int64_t o = own_at(i);
            double u[6], u1[6], w[6];
            if (STAGE == 1) {
                prim2cons6(cm, u);
#pragma unroll
                for (int v = 0; v < 6; ++v) u1[v] = u[v] + dt * (0.0 + fd[v]);
            } else {
                double us[6], w0[6];
#pragma unroll
                for (int q = 0; q < 6; ++q) w0[q] = a.W0[q * vs + o];
                prim2cons6(w0, u);
                prim2cons6(cm, us);
#pragma unroll
                for (int v = 0; v < 6; ++v) {
                    double ut = us[v] + dt * (0.0 + fd[v]);
                    u1[v] = 0.5 * u[v] + 0.5 * ut;
                }
            }
            cons2prim6(u1, w);
#pragma unroll
            for (int v = 0; v < 6; ++v) __builtin_nontemporal_store(w[v], &a.Wout[v * vs + o]);
        }

        // ---- advance the column window and prefetch plane i+3 ----------------
#pragma unroll
        for (int q = 0; q < 6; ++q) Fx[q] = Fxn[q];
#pragma unroll
        for (int q = 0; q < NQ; ++q) cm[q] = c0[q];
#pragma unroll
        for (int q = 0; q < 6; ++q) c0[q] = pf.own[q];
        c0[6] = gm1 * c0[0] * c0[4];
        if (i + 3 <= BS + 1) load_plane(a.Win, vs, own_at(i + 3), halo_at(i + 3), has_halo, pf);
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_hydro_stage(const HydroStageArgs& a, int stage, hipStream_t stream)
{
    const int nblocks = a.g.nbx * a.g.nby * a.g.nbz;
    if (stage == 1) {
        hipLaunchKernelGGL(hydro_stage_kernel<1>, dim3(nblocks), dim3(NT), 0, stream, a);
    } else {
        hipLaunchKernelGGL(hydro_stage_kernel<2>, dim3(nblocks), dim3(NT), 0, stream, a);
    }
    return hipGetLastError();
}

}  // namespace fx
