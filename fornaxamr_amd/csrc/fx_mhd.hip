// MHD RK2 stage on gfx950: MC reconstruction + HLLD fluxes, upwind CT edge
// EMFs, CT face update and the per-cell update.
//
// Reference path (src/prj_timeint.c:2732-2973): prj_flux_update with HLLD for
// every block (src/prj_flux.c:2149-2521, src/prj_riemann.c:321-440), the edge
// EMFs of prj_timeint_mhd_update_emf / prj_mhd_emf_upwind
// (src/prj_timeint.c:133-413, src/prj_mhd.c:627-718), the CT face update
// (src/prj_timeint.c:415-495) and the per-cell update with the cell field taken
// from the new faces (src/prj_timeint.c:498-510, 1255-1350, 1562-1655).
//
// Three launches per stage:
//   mhd_flux_kernel   one workgroup per 16^3 block, marching x-planes like the
//                     hydro kernel: HLLD on the block's faces, flux divergence,
//                     RK combination of the hydro conserved variables and their
//                     conversion to rho, v, Ye and the field-free part of the
//                     internal energy (into the output slot), and the per-face
//                     EMF inputs (Bv1, Bv2, normal velocity);
//   mhd_ghostface_kernel  the same face inputs on the one-face transverse layer
//                     outside the domain that boundary edges read (the
//                     reference computes it from ghost cells, -1..16 loops);
//   mhd_ct_kernel     one workgroup per block, marching x-planes: cell EMFs,
//                     edge EMFs, CT update of the block's faces, cell-centred B
//                     and the magnetic term of the internal energy.
// Faces on block boundaries are computed by both neighbours with identical
// inputs (as the reference does per block); duplicate stores write identical
// values.
#include <hip/hip_runtime.h>

#include "fx_kernels.hpp"
#include "fx_physics.hpp"

namespace fx {

namespace {

constexpr int PW = BS + 2 * NG;     // 20
constexpr int PC = PW * PW;         // 400
constexpr int NQ = 10;              // rho v1 v2 v3 e ye b1 b2 b3 p
constexpr int NHALO = PC - BS * BS; // 144
constexpr int NT = 256;

__device__ __forceinline__ void set_error(int* err, int code, int dir, int x, int y, int z, int stage)
{
    if (atomicCAS(err, 0, code) == 0) {
        err[1] = dir;
        err[2] = x;
        err[3] = y;
        err[4] = z;
        err[5] = stage;
    }
}

__device__ __forceinline__ int block_of_workgroup(int wg, int nwg)
{
    if ((nwg & 7) == 0) {
        int per = nwg >> 3;
        return (wg & 7) * per + (wg >> 3);
    }
    return wg;
}

// prim -> cons with field (src/prj_eos.c:2112-2158)
__device__ __forceinline__ void prim2cons9(const double* w, double* u)
{
    double rho = w[RHO], v1 = w[V1], v2 = w[V2], v3 = w[V3];
    double b1 = w[B1], b2 = w[B2], b3 = w[B3];
    u[0] = rho;
    u[1] = rho * v1;
    u[2] = rho * v2;
    u[3] = rho * v3;
    u[4] = rho * w[EINT] + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3) +
           0.5 * (b1 * b1 + b2 * b2 + b3 * b3);
    u[5] = rho * w[YE];
}

// one HLLD face from 4-cell stencils along the normal; the stencil of
// variable v is rows(v)[c + (s - 2) * S], s = 0..3 (cells c-2S .. c+S, the face
// between c - S and c), pq the pressure of those cells.  Writes the global-order flux of the
// hydro variables and the face EMF inputs.  The rotation to the face frame
// (src/prj_flux.c:274-349) picks the source row per variable.
__device__ __forceinline__ double sel3(int dir, double a, double b, double c)
{
    return dir == 0 ? a : (dir == 1 ? b : c);
}

template <typename Rows>
__device__ __forceinline__ int hlld_face(int dir, const Rows& rows, int c, int S,
                                         const double* pq, double gam, double bn, double* F6,
                                         double* fd3)
{
    double WL[9], WR[9];
#pragma unroll
    for (int lv = 0; lv < 9; ++lv) {
        if (lv == B1) continue;   // the normal field is the face value
        const double* row = rows(gmap(dir, lv));
        const double s0 = row[c - 2 * S], s1 = row[c - S], s2 = row[c], s3 = row[c + S];
        WL[lv] = s1 + mc_half(s0, s1, s2);
        WR[lv] = s2 - mc_half(s1, s2, s3);
    }
    WL[B1] = bn;
    WR[B1] = bn;
    double pL = pq[1] + mc_half(pq[0], pq[1], pq[2]);
    double pR = pq[2] - mc_half(pq[1], pq[2], pq[3]);
    double Fl[9], vf[3], bv1 = 0.0, bv2 = 0.0;
    int rc = hlld(WL, WR, pL, pR, gam, gam, bn, Fl, vf, &bv1, &bv2);
    // global order of the hydro variables (src/prj_flux.c:609-690)
    F6[0] = Fl[0];
    F6[1] = sel3(dir, Fl[1], Fl[3], Fl[2]);
    F6[2] = sel3(dir, Fl[2], Fl[1], Fl[3]);
    F6[3] = sel3(dir, Fl[3], Fl[2], Fl[1]);
    F6[4] = Fl[4];
    F6[5] = Fl[5];
    fd3[0] = bv1;
    fd3[1] = bv2;
    fd3[2] = vf[0];
    return rc;
}

__device__ __forceinline__ void store_fdat(const MhdStageArgs& a, int dir, int x, int y, int z,
                                           const double* fd3)
{
    const int64_t o = cell_off(a.g, x, y, z);
#pragma unroll
    for (int c = 0; c < 3; ++c) a.fdat[(dir * 3 + c) * a.g.vstride + o] = fd3[c];
}

// ---------------------------------------------------------------------------
// Flux kernel.  LDS holds two full planes (x = i and i+1, 20x20 cells with the
// 2-cell halo, 9 primitives + pressure) so that a thread keeps little state in
// registers across the HLLD solves: the left state of its next x-face and the
// flux through its current x-face.  Per plane each thread solves the low-side
// y- and z-face of its own cell and the x-face i+1 (cells i..i+2 along its
// column); half of one wave (rotating) solves the 32 high-side y/z faces on
// the block boundary.  The high-side y/z fluxes of a cell come from the
// neighbouring lanes (shuffles), across wave rows and from the extra faces
// through a small LDS buffer.
template <int STAGE>
__global__ void __launch_bounds__(NT, 2) mhd_flux_kernel(MhdStageArgs a)
{
    __shared__ double pl[2][NQ][PC];
    // y-face fluxes that cross a wave boundary: rows 4, 8, 12 (first rows of
    // waves 1-3) and 16 (the block's high-side faces); z-face fluxes at k = 16.
    // Double-buffered by plane parity, so one barrier per plane suffices: after
    // plane i's barrier a wave finishes plane i (divergence, update of its own
    // cell, plane i+2 into plane i's slot -- its own cell and halo cell, which
    // no other thread reads before the next barrier) and may run on into plane
    // i+1's faces, which read plane i+1's slot and write the other exchange
    // buffer while slower waves still read this one.
    __shared__ double fyb[2][6][5][BS];
    __shared__ double fz16[2][6][BS];

    const Geom& g = a.g;
    const int t = threadIdx.x;
    const int j = t >> 4;
    const int k = t & 15;
    const int lane = t & 63;
    const int b = block_of_workgroup(blockIdx.x, gridDim.x);
    const int bz = b % g.nbz;
    const int by = (b / g.nbz) % g.nby;
    const int bx = b / (g.nbz * g.nby);
    const int x0 = bx * BS, y0 = by * BS, z0 = bz * BS;
    const int64_t vs = g.vstride;
    const double gm1 = g.gamma - 1.0;
    const double gam = g.gamma;
    const double dt = a.dt_ptr ? *a.dt_ptr : a.dt;

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

    // own cell of plane p: 9 primitives and the pressure
    auto load_own = [&](int p, double* c) {
#pragma unroll
        for (int q = 0; q < 9; ++q) c[q] = a.Win[q * vs + own_at(p)];
        c[9] = gm1 * c[0] * c[4];
    };
    // plane p into LDS slot s: the own cell given, the halo cell loaded here
    auto store_plane = [&](int s, const double* own, int p) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) pl[s][q][oidx] = own[q];
        if (has_halo) {
            double h[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) h[q] = a.Win[q * vs + halo_at(p)];
#pragma unroll
            for (int q = 0; q < 9; ++q) pl[s][q][hidx] = h[q];
            pl[s][9][hidx] = gm1 * h[0] * h[4];
        }
    };

    double Ls[NQ];   // left state of the next x-face (rho .. b3, p)
    double Fx[6];    // flux through the current x-face

    // x-face between own cells L = plane l and R = plane l+1 with stencil
    // cells l-1 (cm), l (c0), l+1 (c1), l+2 (c2); Ls holds the left state on
    // entry and the left state of the next face on exit
    auto x_face = [&](int xf, const double* c0, const double* c1, const double* c2, double* F6) {
        double WL[9], WR[9], pL = 0.0, pR = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double h = mc_half(c0[q], c1[q], c2[q]);
            double r = c1[q] - h;
            if (q < 9) { WL[q] = Ls[q]; WR[q] = r; } else { pL = Ls[9]; pR = r; }
            Ls[q] = c1[q] + h;
        }
        double bn = a.bf_in[cell_off(g, x0 + xf, y0 + j, z0 + k)];
        WL[B1] = bn;
        WR[B1] = bn;
        double Fl[9], vf[3], bv1 = 0.0, bv2 = 0.0;
        int rc = hlld(WL, WR, pL, pR, gam, gam, bn, Fl, vf, &bv1, &bv2);
        if (rc) set_error(a.err, rc, 0, x0 + xf, y0 + j, z0 + k, STAGE);
#pragma unroll
        for (int v = 0; v < 6; ++v) F6[v] = Fl[v];
        double fd3[3] = {bv1, bv2, vf[0]};
        store_fdat(a, 0, x0 + xf, y0 + j, z0 + k, fd3);
    };

    // ---- prologue: planes 0, 1 into LDS; x-face 0 -------------------------
    {
        double cmm[NQ], cm[NQ], c0[NQ], c1[NQ];
        load_own(-2, cmm);
        load_own(-1, cm);
        load_own(0, c0);
        load_own(1, c1);
        store_plane(0, c0, 0);
        store_plane(1, c1, 1);
#pragma unroll
        for (int q = 0; q < NQ; ++q) Ls[q] = cm[q] + mc_half(cmm[q], cm[q], c0[q]);
        x_face(0, cm, c0, c1, Fx);
        __syncthreads();
    }

    const double a1 = g.area[0], a2 = g.area[1], a3 = g.area[2], vol = g.vol;
    const double* bfy = a.bf_in + g.fvstride;
    const double* bfz = a.bf_in + 2 * g.fvstride;

    for (int i = 0; i < BS; ++i) {
        const int A = i & 1;       // LDS slot of plane i, and its exchange buffers
        const int B = A ^ 1;       // plane i + 1
        // ---- y/z faces of plane i ------------------------------------------
        const int xw = i & 3;
        const bool extra = (t >> 6) == xw && lane < 32;
        // the wave with the extra faces issues first on its SIMD until the barrier
        if (__builtin_amdgcn_readfirstlane(t >> 6) == xw) __builtin_amdgcn_s_setprio(2);
        double Fy[6], Fz[6];
#pragma unroll 1
        for (int f = 0; f < 3; ++f) {
            if (f == 2 && !extra) continue;
            int dir, fj = j, fk = k;
            if (f == 0) dir = 1;
            else if (f == 1) dir = 2;
            else if (lane < 16) { dir = 1; fj = BS; fk = lane; }
            else { dir = 2; fj = lane - 16; fk = BS; }
            const int c = (fj + NG) * PW + (fk + NG);
            const int S = dir == 1 ? PW : 1;
            const int gx = x0 + i, gy = y0 + fj, gz = z0 + fk;
            double bn = (dir == 1 ? bfy : bfz)[cell_off(g, gx, gy, gz)];
            double F6[6], fd3[3];
            const double pq[4] = {pl[A][9][c - 2 * S], pl[A][9][c - S], pl[A][9][c],
                                  pl[A][9][c + S]};
            int rc = hlld_face(dir, [&](int v) { return &pl[A][v][0]; }, c, S, pq, gam, bn, F6,
                               fd3);
            if (rc) set_error(a.err, rc, dir, gx, gy, gz, STAGE);
            store_fdat(a, dir, gx, gy, gz, fd3);
            if (f == 0) {
#pragma unroll
                for (int v = 0; v < 6; ++v) Fy[v] = F6[v];
                if ((j & 3) == 0 && j > 0) {
#pragma unroll
                    for (int v = 0; v < 6; ++v) fyb[A][v][j >> 2][k] = F6[v];
                }
            } else if (f == 1) {
#pragma unroll
                for (int v = 0; v < 6; ++v) Fz[v] = F6[v];
            } else if (dir == 1) {
#pragma unroll
                for (int v = 0; v < 6; ++v) fyb[A][v][4][fk] = F6[v];
            } else {
#pragma unroll
                for (int v = 0; v < 6; ++v) fz16[A][v][fj] = F6[v];
            }
        }
        // ---- x-face i+1 from cells i, i+1 (LDS) and i+2 (loaded) ------------
        double c2[NQ];
        load_own(i + 2, c2);
        double Fxn[6];
        {
            double c0[NQ], c1[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                c0[q] = pl[A][q][oidx];
                c1[q] = pl[B][q][oidx];
            }
            x_face(i + 1, c0, c1, c2, Fxn);
        }
        __syncthreads();
        __builtin_amdgcn_s_setprio(0);
        // halo cell of plane i+2, loaded here so that its latency overlaps the
        // divergence and update below (it goes to LDS at the end of the plane)
        double hh[9];
        if (has_halo && i < BS - 1) {
#pragma unroll
            for (int q = 0; q < 9; ++q) hh[q] = a.Win[q * vs + halo_at(i + 2)];
        }
        // stage 2: W^n of the own cell, for the RK combination below
        double w0[9];
        if (STAGE == 2) {
#pragma unroll
            for (int q = 0; q < 9; ++q) w0[q] = a.W0[q * vs + own_at(i)];
        }

        // ---- divergence (x, y, z order) and the RK combination -------------
        double fd[6];
#pragma unroll
        for (int v = 0; v < 6; ++v) {
            const double fy_dn = __shfl_down(Fy[v], 16);
            const double fz_dn = __shfl_down(Fz[v], 1);
            const double fy_hi = (j & 3) != 3 ? fy_dn : fyb[A][v][(j + 1) >> 2][k];
            const double fz_hi = k != BS - 1 ? fz_dn : fz16[A][v][j];
            fd[v] = -(a1 * (Fxn[v] - Fx[v]) + a2 * (fy_hi - Fy[v]) + a3 * (fz_hi - Fz[v])) / vol;
        }
        {
            double cm[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) cm[q] = pl[A][q][oidx];
            double u[6], u1[6];
            if (STAGE == 1) {
                prim2cons9(cm, u);
#pragma unroll
                for (int v = 0; v < 6; ++v) u1[v] = u[v] + dt * (0.0 + fd[v]);
            } else {
                double us[6];
                prim2cons9(w0, u);
                prim2cons9(cm, us);
#pragma unroll
                for (int v = 0; v < 6; ++v) {
                    double ut = us[v] + dt * (0.0 + fd[v]);
                    u1[v] = 0.5 * u[v] + 0.5 * ut;
                }
            }
            // cons -> prim of the parts that do not depend on the new field
            // (src/prj_eos.c:2160-2220): rho, v, Ye, and u4/rho - kinetic, the
            // first two terms of the specific internal energy, which the CT
            // kernel completes by subtracting the magnetic term
            double w[6];
            if (u1[0] == 0.0) {
#pragma unroll
                for (int v = 0; v < 6; ++v) w[v] = 0.0;
            } else {
                const double rho = u1[0];
                const double v1 = u1[1] / rho, v2 = u1[2] / rho, v3 = u1[3] / rho;
                const double kinetic = 0.5 * (v1 * v1 + v2 * v2 + v3 * v3);
                w[0] = rho;
                w[1] = v1;
                w[2] = v2;
                w[3] = v3;
                w[4] = u1[4] / rho - kinetic;
                w[5] = u1[5] / rho;
            }
            const int64_t o = own_at(i);
#pragma unroll
            for (int v = 0; v < 6; ++v) a.Wout[v * vs + o] = w[v];
        }
#pragma unroll
        for (int v = 0; v < 6; ++v) Fx[v] = Fxn[v];
        // ---- plane i+2 into the slot of plane i ----------------------------
        if (i < BS - 1) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) pl[A][q][oidx] = c2[q];
            if (has_halo) {
#pragma unroll
                for (int q = 0; q < 9; ++q) pl[A][q][hidx] = hh[q];
                pl[A][9][hidx] = gm1 * hh[0] * hh[4];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Face EMF inputs on the transverse ghost layer of the domain: faces of
// direction d whose transverse coordinates include -1 or N (the reference's
// -1..16 face loops on boundary blocks, src/prj_flux.c:2241-2251).
__device__ __forceinline__ bool decode_ghostface(const Geom& g, int64_t id64, int& dir, int& x,
                                                 int& y, int& z)
{
    // per direction: normal extent N_d + 1, transverse box (N_t + 2) x (N_u + 2)
    // minus the inner N_t x N_u.  32-bit unsigned arithmetic: the layer has
    // about 12 N^2 faces, far below 2^31 for any mesh the C ABI accepts, and
    // 64-bit division by a run-time divisor is a long software sequence.
    const int n[3] = {g.nx, g.ny, g.nz};
    unsigned id = (unsigned)id64;
    for (int d = 0; d < 3; ++d) {
        const int t1 = d == 2 ? 0 : d + 1, t2 = d == 0 ? 2 : (d == 1 ? 0 : 1);
        const unsigned w2 = (unsigned)(n[t2] + 2);
        const unsigned outer = (unsigned)(n[t1] + 2) * w2;
        const unsigned ring = outer - (unsigned)n[t1] * (unsigned)n[t2];
        const unsigned cnt = (unsigned)(n[d] + 1) * ring;
        if (id < cnt) {
            // z-faces: the normal index fastest, so that a wave's stencils
            // (along z) are contiguous; x- and y-faces: the ring position
            // fastest
            unsigned nd, r;
            if (d == 2) {
                r = id / (unsigned)(n[d] + 1);
                nd = id - r * (unsigned)(n[d] + 1);
            } else {
                nd = id / ring;
                r = id - nd * ring;
            }
            int c1, c2;
            // ring enumeration: two full rows (c1 = -1, N) then two side columns
            const unsigned rows = 2u * w2;
            if (r < rows) {
                const unsigned hi = r >= w2 ? 1u : 0u;
                c1 = hi ? n[t1] : -1;
                c2 = (int)(r - hi * w2) - 1;
            } else {
                r -= rows;
                c1 = (int)(r >> 1);
                c2 = (r & 1) ? n[t2] : -1;
            }
            int c[3];
            c[d] = (int)nd;
            c[t1] = c1;
            c[t2] = c2;
            dir = d;
            x = c[0];
            y = c[1];
            z = c[2];
            return true;
        }
        id -= cnt;
    }
    return false;
}

__device__ __forceinline__ double bf_clamped(const Geom& g, const double* bf, int dir, int x,
                                             int y, int z)
{
    int hx = g.nx - 1 + (dir == 0), hy = g.ny - 1 + (dir == 1), hz = g.nz - 1 + (dir == 2);
    x = x < 0 ? 0 : (x > hx ? hx : x);
    y = y < 0 ? 0 : (y > hy ? hy : y);
    z = z < 0 ? 0 : (z > hz ? hz : z);
    return bf[dir * g.fvstride + cell_off(g, x, y, z)];
}

__global__ void __launch_bounds__(256) mhd_ghostface_kernel(MhdStageArgs a, int64_t total, int stage)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= total) return;
    int dir, x, y, z;
    if (!decode_ghostface(a.g, id, dir, x, y, z)) return;
    const Geom& g = a.g;
    const double gm1 = g.gamma - 1.0;
    // the stencil: base (the right cell), the stride along the face normal and
    // between variables -- in the padded arrays, or, for x- and y-faces on the
    // ghost planes z = -1 and z = nz, in their compact copies (launch_boundary_fill)
    int S = dir == 0 ? (int)g.sx : (dir == 1 ? (int)g.sy : 1);
    const double* base = a.Win + cell_off(g, x, y, z);
    int64_t vs = g.vstride;
    if (dir != 2 && (z == -1 || z == g.nz) && a.zring != nullptr) {
        const int64_t pl = zring_plane(g);
        const int zc = z == -1 ? 0 : 1;
        vs = 2 * pl;
        if (dir == 0) {
            S = g.ny + 4;
            base = a.zring + zc * pl + (int64_t)(x + 2) * (g.ny + 4) + (y + 2);
        } else {
            S = g.nx + 4;
            base = a.zring + 18 * pl + zc * pl + (int64_t)(y + 2) * (g.nx + 4) + (x + 2);
        }
    }
    double pq[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const double* cell = base + (s - 2) * S;
        pq[s] = gm1 * cell[0] * cell[4 * vs];
    }
    double bn = bf_clamped(g, a.bf_in, dir, x, y, z);
    double F6[6], fd3[3];
    int rc = hlld_face(dir, [&](int v) { return base + v * vs; }, 0, S, pq, g.gamma, bn, F6, fd3);
    if (rc) set_error(a.err, rc, dir, x, y, z, stage);
    store_fdat(a, dir, x, y, z, fd3);
}

// ---------------------------------------------------------------------------
// CT: edge EMFs, face update, cell-centred field and cons->prim.
constexpr int CW = BS + 2;           // 18: cells -1..16
constexpr int CC = CW * CW;          // 324
constexpr int EW = BS + 1;           // 17

__device__ __forceinline__ double fdat_at(const MhdStageArgs& a, int dir, int comp, int x, int y,
                                          int z)
{
    return a.fdat[(dir * 3 + comp) * a.g.vstride + cell_off(a.g, x, y, z)];
}

// resistive curl B at an edge (src/prj_mhd.c:689-711), faces of the stage input
__device__ double edge_curl_b(const MhdStageArgs& a, int dir, int i, int j, int k)
{
    const Geom& g = a.g;
    const double* bf = a.bf_in;
    if (dir == 0) {
        return (bf_clamped(g, bf, 2, i, j, k) - bf_clamped(g, bf, 2, i, j - 1, k)) / g.dx[1] -
               (bf_clamped(g, bf, 1, i, j, k) - bf_clamped(g, bf, 1, i, j, k - 1)) / g.dx[2];
    } else if (dir == 1) {
        return (bf_clamped(g, bf, 0, i, j, k) - bf_clamped(g, bf, 0, i, j, k - 1)) / g.dx[2] -
               (bf_clamped(g, bf, 2, i, j, k) - bf_clamped(g, bf, 2, i - 1, j, k)) / g.dx[0];
    }
    return (bf_clamped(g, bf, 1, i, j, k) - bf_clamped(g, bf, 1, i - 1, j, k)) / g.dx[0] -
           (bf_clamped(g, bf, 0, i, j, k) - bf_clamped(g, bf, 0, i, j - 1, k)) / g.dx[1];
}

template <int STAGE>
__global__ void __launch_bounds__(NT, 4) mhd_ct_kernel(MhdStageArgs a)
{
    // cell EMF components and eta mask of planes (ring of 2), cells -1..16
    __shared__ double ce[2][4][CC];
    // edge EMFs: Ex of plane i, Ey/Ez at x = i and x = i + 1 (ring of 2)
    __shared__ double ex_e[EW * EW];
    __shared__ double ey_e[2][BS * EW];
    __shared__ double ez_e[2][EW * BS];
    // new By / Bz of plane i for the cell-centred field
    __shared__ double nby[EW * BS];
    __shared__ double nbz[BS * EW];

    const Geom& g = a.g;
    const int t = threadIdx.x;
    const int j = t >> 4;
    const int k = t & 15;
    // workgroup -> (x-part, block): a block's x-planes may be split over
    // ct_split workgroups, each marching np planes from x0 with its own
    // prologue (the face between two parts is formed by both, from the same
    // edges: identical values)
    const int wg = block_of_workgroup(blockIdx.x, gridDim.x);
    const int np = BS / a.ct_split;
    const int part = wg % a.ct_split;
    const int b = wg / a.ct_split;
    const int bz = b % g.nbz;
    const int by = (b / g.nbz) % g.nby;
    const int bx = b / (g.nbz * g.nby);
    const int x0 = bx * BS + part * np, y0 = by * BS, z0 = bz * BS;
    const int64_t vs = g.vstride;
    const double dt = a.dt_ptr ? *a.dt_ptr : a.dt;
    const double eta0 = g.eta;

    auto cell_fill = [&](int slot, int p) {
        for (int c = t; c < CC; c += NT) {
            const int cj = c / CW - 1, ck = c % CW - 1;
            const int64_t o = cell_off(g, x0 + p, y0 + cj, z0 + ck);
            double v1 = a.Win[1 * vs + o], v2 = a.Win[2 * vs + o], v3 = a.Win[3 * vs + o];
            double b1 = a.Win[6 * vs + o], b2 = a.Win[7 * vs + o], b3 = a.Win[8 * vs + o];
            // src/prj_timeint.c:133-177
            double e0 = b2 * v3 - b3 * v2;
            double e1 = b3 * v1 - b1 * v3;
            double e2 = b1 * v2 - b2 * v1;
            if (!isfinite(e0) || !isfinite(e1) || !isfinite(e2))
                set_error(a.err, 32, -1, x0 + p, y0 + cj, z0 + ck, STAGE);
            ce[slot][0][c] = e0;
            ce[slot][1][c] = e1;
            ce[slot][2][c] = e2;
            ce[slot][3][c] = eta_mask(a.Win[o], a.Win[5 * vs + o]);
        }
    };
    auto cidx = [](int cj, int ck) { return (cj + 1) * CW + (ck + 1); };
    // one edge EMF from gathered inputs
    auto edge = [&](int dir, int gx, int gy, int gz, const double* ef, const double* ec,
                    const double* vn, const double* mk) -> double {
        for (int n = 0; n < 4; ++n) {
            if (!isfinite(ef[n]) || !isfinite(ec[n]) || !isfinite(vn[n])) {
                set_error(a.err, 33, dir, gx, gy, gz, STAGE);
            }
        }
        double emf = emf_upwind(ef, ec, vn);
        double m = 0.0;
        for (int n = 0; n < 4; ++n) m = mk[n] > m ? mk[n] : m;
        double eta = eta0 + m;
        if (eta != 0.0) emf += eta * edge_curl_b(a, dir, gx, gy, gz);
        if (!isfinite(emf)) set_error(a.err, 33, dir, gx, gy, gz, STAGE);
        return emf;
    };
    // Ey at x-position X (local), (j, k) with j 0..15, k 0..16; cell planes X-1 (slot sm), X (slot s0)
    auto edge_y = [&](int X, int jj, int kk, int sm, int s0) -> double {
        const int gx = x0 + X, gy = y0 + jj, gz = z0 + kk;
        double ef[4], ec[4], vn[4], mk[4];
        ef[0] = fdat_at(a, 0, 0, gx, gy, gz);
        ef[1] = fdat_at(a, 2, 1, gx, gy, gz);
        ef[2] = fdat_at(a, 0, 0, gx, gy, gz - 1);
        ef[3] = fdat_at(a, 2, 1, gx - 1, gy, gz);
        vn[0] = fdat_at(a, 0, 2, gx, gy, gz);
        vn[1] = fdat_at(a, 2, 2, gx, gy, gz);
        vn[2] = fdat_at(a, 0, 2, gx, gy, gz - 1);
        vn[3] = fdat_at(a, 2, 2, gx - 1, gy, gz);
        const int c0 = cidx(jj, kk), c1 = cidx(jj, kk - 1);
        ec[0] = ce[s0][1][c0]; mk[0] = ce[s0][3][c0];
        ec[1] = ce[s0][1][c1]; mk[1] = ce[s0][3][c1];
        ec[2] = ce[sm][1][c1]; mk[2] = ce[sm][3][c1];
        ec[3] = ce[sm][1][c0]; mk[3] = ce[sm][3][c0];
        return edge(1, gx, gy, gz, ef, ec, vn, mk);
    };
    // Ez at x-position X, (j, k) with j 0..16, k 0..15
    auto edge_z = [&](int X, int jj, int kk, int sm, int s0) -> double {
        const int gx = x0 + X, gy = y0 + jj, gz = z0 + kk;
        double ef[4], ec[4], vn[4], mk[4];
        ef[0] = fdat_at(a, 1, 0, gx, gy, gz);
        ef[1] = fdat_at(a, 0, 1, gx, gy, gz);
        ef[2] = fdat_at(a, 1, 0, gx - 1, gy, gz);
        ef[3] = fdat_at(a, 0, 1, gx, gy - 1, gz);
        vn[0] = fdat_at(a, 1, 2, gx, gy, gz);
        vn[1] = fdat_at(a, 0, 2, gx, gy, gz);
        vn[2] = fdat_at(a, 1, 2, gx - 1, gy, gz);
        vn[3] = fdat_at(a, 0, 2, gx, gy - 1, gz);
        const int c0 = cidx(jj, kk), c3 = cidx(jj - 1, kk);
        ec[0] = ce[s0][2][c0]; mk[0] = ce[s0][3][c0];
        ec[1] = ce[sm][2][c0]; mk[1] = ce[sm][3][c0];
        ec[2] = ce[sm][2][c3]; mk[2] = ce[sm][3][c3];
        ec[3] = ce[s0][2][c3]; mk[3] = ce[s0][3][c3];
        return edge(2, gx, gy, gz, ef, ec, vn, mk);
    };
    // Ex in plane i (local), (j, k) with j, k 0..16; cells in slot s
    auto edge_x = [&](int i, int jj, int kk, int s) -> double {
        const int gx = x0 + i, gy = y0 + jj, gz = z0 + kk;
        double ef[4], ec[4], vn[4], mk[4];
        ef[0] = fdat_at(a, 2, 0, gx, gy, gz);
        ef[1] = fdat_at(a, 1, 1, gx, gy, gz);
        ef[2] = fdat_at(a, 2, 0, gx, gy - 1, gz);
        ef[3] = fdat_at(a, 1, 1, gx, gy, gz - 1);
        vn[0] = fdat_at(a, 2, 2, gx, gy, gz);
        vn[1] = fdat_at(a, 1, 2, gx, gy, gz);
        vn[2] = fdat_at(a, 2, 2, gx, gy - 1, gz);
        vn[3] = fdat_at(a, 1, 2, gx, gy, gz - 1);
        const int c0 = cidx(jj, kk), c1 = cidx(jj - 1, kk), c2 = cidx(jj - 1, kk - 1),
                  c3 = cidx(jj, kk - 1);
        ec[0] = ce[s][0][c0]; mk[0] = ce[s][3][c0];
        ec[1] = ce[s][0][c1]; mk[1] = ce[s][3][c1];
        ec[2] = ce[s][0][c2]; mk[2] = ce[s][3][c2];
        ec[3] = ce[s][0][c3]; mk[3] = ce[s][3][c3];
        return edge(0, gx, gy, gz, ef, ec, vn, mk);
    };
    // CT value of one face (src/prj_timeint.c:415-495); e_hi/e_lo pairs in
    // the reference's order, `dn` the two cell widths; the face's stage-input
    // field is `pre` where `use_pre`, else loaded here
    auto ct = [&](int dir, int gx, int gy, int gz, double ea1, double ea0, double da, double eb1,
                  double eb0, double db, bool use_pre = false, double pre = 0.0) -> double {
        const int64_t o = dir * g.fvstride + cell_off(g, gx, gy, gz);
        double value = use_pre ? pre : a.bf_in[o];
        value += -dt * (ea1 - ea0) / da;
        value += dt * (eb1 - eb0) / db;
        if (!isfinite(value)) set_error(a.err, 34, dir, gx, gy, gz, STAGE);
        if (STAGE == 2) value = 0.5 * a.bf0[o] + 0.5 * value;
        a.bf_out[o] = value;
        return value;
    };

    // ---- prologue: cell planes -1, 0; Ey/Ez at x = 0; Bx at x = 0 ---------
    cell_fill(1, -1);
    cell_fill(0, 0);
    __syncthreads();
    for (int e = t; e < BS * EW; e += NT) {
        int jj = e / EW, kk = e % EW;
        ey_e[0][e] = edge_y(0, jj, kk, 1, 0);
    }
    for (int e = t; e < EW * BS; e += NT) {
        int jj = e / BS, kk = e % BS;
        ez_e[0][e] = edge_z(0, jj, kk, 1, 0);
    }
    __syncthreads();
    // Bx(0, j, k): -dt (Ez(j+1) - Ez(j))/dy + dt (Ey(k+1) - Ey(k))/dz
    double bx_lo = ct(0, x0, y0 + j, z0 + k, ez_e[0][(j + 1) * BS + k], ez_e[0][j * BS + k],
                      g.dx[1], ey_e[0][j * EW + k + 1], ey_e[0][j * EW + k], g.dx[2]);

    // stage 2: each thread's running CFL maximum, kept in LDS (a register
    // live across the plane loop would cost the kernel its fourth wave)
    __shared__ double cfl_acc[NT];
    if (STAGE == 2) cfl_acc[t] = 0.0;
    for (int i = 0; i < np; ++i) {
        const int s0 = i & 1;         // cell slot of plane i
        const int s1 = s0 ^ 1;        // plane i + 1
        const int e0 = i & 1;         // edge slot of x = i
        const int e1 = e0 ^ 1;        // x = i + 1
        // cell plane i+1 into the slot of plane i-1: that slot's last readers
        // (the edges of the previous plane, or of the prologue) are behind the
        // barrier that preceded the previous CT step
        cell_fill(s1, i + 1);
        // stage 1: loads of the last phase for this thread's own cell, issued
        // here so that their latency overlaps the edge phase -- rho and
        // u4/rho - kinetic, which the flux kernel left in the output slot
        // (stage 2 has no VGPRs to spare for them across the edges)
        double rho_c = 0.0, t_c = 0.0;
        if (STAGE == 1) {
            const int64_t oc = cell_off(g, x0 + i, y0 + j, z0 + k);
            rho_c = a.Wout[oc];
            t_c = a.Wout[4 * vs + oc];
        }
        __syncthreads();
        // Each loop has a remainder beyond NT items (33, 16, 16); rotating
        // the thread -> item map by whole waves gives each remainder to a
        // different wave (0, 3, 2), so no wave carries all three extra passes.
        for (int e = t; e < EW * EW; e += NT) {
            int jj = e / EW, kk = e % EW;
            ex_e[e] = edge_x(i, jj, kk, s0);
        }
        for (int e = (t + 64) & (NT - 1); e < BS * EW; e += NT) {
            int jj = e / EW, kk = e % EW;
            ey_e[e1][e] = edge_y(i + 1, jj, kk, s0, s1);
        }
        for (int e = (t + 128) & (NT - 1); e < EW * BS; e += NT) {
            int jj = e / BS, kk = e % BS;
            ez_e[e1][e] = edge_z(i + 1, jj, kk, s0, s1);
        }
        // stage 2: the stage-input fields of this thread's first face in each
        // CT loop below (Bx(i+1) of its own cell, By / Bz of plane i), loaded
        // before the barrier so that their latency overlaps the wait (stage 1,
        // which holds rho and t of its cell here, has no VGPRs for them)
        const int fy0 = (t + 64) & (NT - 1), fz0 = (t + 128) & (NT - 1);
        double bx_in = 0.0, by_in = 0.0, bz_in = 0.0;
        if (STAGE == 2) {
            const int64_t fv = g.fvstride;
            bx_in = a.bf_in[cell_off(g, x0 + i + 1, y0 + j, z0 + k)];
            by_in = a.bf_in[fv + cell_off(g, x0 + i, y0 + fy0 / BS, z0 + fy0 % BS)];
            bz_in = a.bf_in[2 * fv + cell_off(g, x0 + i, y0 + fz0 / EW, z0 + fz0 % EW)];
        }
        __syncthreads();
        // CT of Bx(i+1), By(plane i), Bz(plane i)
        double bx_hi = ct(0, x0 + i + 1, y0 + j, z0 + k, ez_e[e1][(j + 1) * BS + k],
                          ez_e[e1][j * BS + k], g.dx[1], ey_e[e1][j * EW + k + 1],
                          ey_e[e1][j * EW + k], g.dx[2], STAGE == 2, bx_in);
        for (int f = fy0; f < EW * BS; f += NT) {   // remainder: wave 3
            int jj = f / BS, kk = f % BS;   // By(i, jj 0..16, kk 0..15)
            nby[f] = ct(1, x0 + i, y0 + jj, z0 + kk, ex_e[jj * EW + kk + 1], ex_e[jj * EW + kk],
                        g.dx[2], ez_e[e1][jj * BS + kk], ez_e[e0][jj * BS + kk], g.dx[0],
                        STAGE == 2 && f == fy0, by_in);
        }
        for (int f = fz0; f < BS * EW; f += NT) {  // remainder: wave 2
            int jj = f / EW, kk = f % EW;   // Bz(i, jj 0..15, kk 0..16)
            nbz[f] = ct(2, x0 + i, y0 + jj, z0 + kk, ey_e[e1][jj * EW + kk], ey_e[e0][jj * EW + kk],
                        g.dx[0], ex_e[(jj + 1) * EW + kk], ex_e[jj * EW + kk], g.dx[1],
                        STAGE == 2 && f == fz0, bz_in);
        }
        __syncthreads();
        // cell-centred field from the new faces, and the rest of cons ->
        // prim (cons2prim9): the flux kernel left rho, v, Ye and
        // t = u4/rho - kinetic in Wout; eint = t - magnetic, with
        // magnetic = 0.5 |B|^2 / rho, is the reference's left-to-right
        // u4/rho - kinetic - magnetic
        {
            double bc[3];
            bc[0] = 0.5 * (bx_lo + bx_hi);
            bc[1] = 0.5 * (nby[j * BS + k] + nby[(j + 1) * BS + k]);
            bc[2] = 0.5 * (nbz[j * EW + k] + nbz[j * EW + k + 1]);
            const int64_t o = cell_off(g, x0 + i, y0 + j, z0 + k);
            if (STAGE == 2) {
                rho_c = a.Wout[o];
                t_c = a.Wout[4 * vs + o];
            }
            const double rho = rho_c;
            double eint = 0.0;
            if (rho != 0.0) {
                const double b2 = bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2];
                const double magnetic = 0.5 * b2 / rho;
                eint = t_c - magnetic;
            } else {
                bc[0] = bc[1] = bc[2] = 0.0;   // cons2prim9 zeroes every prim of an empty cell
            }
            a.Wout[4 * vs + o] = eint;
#pragma unroll
            for (int c = 0; c < 3; ++c) a.Wout[(6 + c) * vs + o] = bc[c];
            if (STAGE == 2) {
                // the next step's CFL maximum, from the new state
                const double d = cfl_denom(g, rho, a.Wout[1 * vs + o], a.Wout[2 * vs + o],
                                           a.Wout[3 * vs + o], eint);
                if (d > cfl_acc[t]) cfl_acc[t] = d;   // NaN never wins
            }
        }
        bx_lo = bx_hi;
    }
    if (STAGE == 2 && a.cfl_part != nullptr) {
        __syncthreads();   // last plane's nby reads done
        const double m = workgroup_max(cfl_acc[t], nby);
        if (t == 0) a.cfl_part[blockIdx.x] = m;
    }
}

__global__ void bf2bc_kernel(Geom g, double* W, const double* bf)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    int z = (int)(id % g.nz);
    int64_t r = id / g.nz;
    int y = (int)(r % g.ny);
    int x = (int)(r / g.ny);
    const int64_t o = cell_off(g, x, y, z);
    const double* bxf = bf;
    const double* byf = bf + g.fvstride;
    const double* bzf = bf + 2 * g.fvstride;
    W[6 * g.vstride + o] = 0.5 * (bxf[o] + bxf[o + g.sx]) / 1.0;
    W[7 * g.vstride + o] = 0.5 * (byf[o] + byf[o + g.sy]) / 1.0;
    W[8 * g.vstride + o] = 0.5 * (bzf[o] + bzf[o + 1]) / 1.0;
}

int64_t ghostface_count(const Geom& g)
{
    const int n[3] = {g.nx, g.ny, g.nz};
    int64_t tot = 0;
    for (int d = 0; d < 3; ++d) {
        const int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
        const int64_t ring = (int64_t)(n[t1] + 2) * (n[t2] + 2) - (int64_t)n[t1] * n[t2];
        tot += (int64_t)(n[d] + 1) * ring;
    }
    return tot;
}

}  // namespace

int mhd_ct_split(const Geom& g, int ncu)
{
    // the CT kernel holds 4 workgroups per CU (36 KB LDS each)
    const int64_t nblocks = (int64_t)g.nbx * g.nby * g.nbz;
    int split = 1;
    while (split < 4 && nblocks * split * 2 <= 4 * (int64_t)ncu) split *= 2;
    return split;
}

hipError_t launch_mhd_stage(const MhdStageArgs& a, int stage, hipStream_t s, hipStream_t side,
                            hipEvent_t fork, hipEvent_t join)
{
    const int nblocks = a.g.nbx * a.g.nby * a.g.nbz;
    if (a.ct_split != 1 && a.ct_split != 2 && a.ct_split != 4) return hipErrorInvalidValue;
    if (side != nullptr && (fork == nullptr || join == nullptr)) return hipErrorInvalidValue;
    const int nct = nblocks * a.ct_split;
    const int64_t ngf = ghostface_count(a.g);
    if (ngf >= ((int64_t)1 << 31)) return hipErrorInvalidValue;   // 32-bit decode
    // The ghost-layer faces and the flux kernel's faces are disjoint sets of
    // fdat entries and both only read the state and the faces: the ghost-layer
    // kernel (latency-bound, a fraction of the chip) can run beside the flux
    // kernel (issue-bound at two workgroups per CU).
    hipStream_t gs = s;
    if (side != nullptr) {
        if (hipError_t e = hipEventRecord(fork, s)) return e;
        if (hipError_t e = hipStreamWaitEvent(side, fork, 0)) return e;
        gs = side;
    }
    hipLaunchKernelGGL(mhd_ghostface_kernel, dim3((unsigned)((ngf + 255) / 256)), dim3(256), 0, gs,
                       a, ngf, stage);
    if (side != nullptr) {
        if (hipError_t e = hipEventRecord(join, side)) return e;
    }
    if (stage == 1) hipLaunchKernelGGL(mhd_flux_kernel<1>, dim3(nblocks), dim3(NT), 0, s, a);
    else hipLaunchKernelGGL(mhd_flux_kernel<2>, dim3(nblocks), dim3(NT), 0, s, a);
    if (side != nullptr) {
        if (hipError_t e = hipStreamWaitEvent(s, join, 0)) return e;
    }
    if (stage == 1) hipLaunchKernelGGL(mhd_ct_kernel<1>, dim3(nct), dim3(NT), 0, s, a);
    else hipLaunchKernelGGL(mhd_ct_kernel<2>, dim3(nct), dim3(NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bf2bc(const Geom& g, double* W, const double* bf, hipStream_t s)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(bf2bc_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, W, bf);
    return hipGetLastError();
}

}  // namespace fx
