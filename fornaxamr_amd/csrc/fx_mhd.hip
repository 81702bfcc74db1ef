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
//                     RK combination of the hydro conserved variables (U1),
//                     and the per-face EMF inputs (Bv1, Bv2, normal velocity);
//   mhd_ghostface_kernel  the same face inputs on the one-face transverse layer
//                     outside the domain that boundary edges read (the
//                     reference computes it from ghost cells, -1..16 loops);
//   mhd_ct_kernel     one workgroup per block, marching x-planes: cell EMFs,
//                     edge EMFs, CT update of the block's faces, cell-centred B
//                     and cons->prim of the block's cells.
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
constexpr int EX = (BS + 1) * BS;   // 272

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

// cons -> prim with field (src/prj_eos.c:2160-2220)
__device__ __forceinline__ void cons2prim9(const double* u, const double* b, double* w)
{
    double rho = u[0];
    if (rho == 0.0) {
#pragma unroll
        for (int v = 0; v < 9; ++v) w[v] = 0.0;
        return;
    }
    double v1 = u[1] / rho, v2 = u[2] / rho, v3 = u[3] / rho;
    double kinetic = 0.5 * (v1 * v1 + v2 * v2 + v3 * v3);
    double magnetic = 0.5 * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]) / rho;
    w[0] = rho;
    w[1] = v1;
    w[2] = v2;
    w[3] = v3;
    w[4] = u[4] / rho - kinetic - magnetic;
    w[5] = u[5] / rho;
    w[6] = b[0];
    w[7] = b[1];
    w[8] = b[2];
}

// one HLLD face from 4-cell stencils along the normal: q[s][gv] for cells
// c-2..c+1 in the global variable order, pressure p[s]; writes the
// global-order flux (first 6 variables) and the face EMF inputs.  The
// rotation to the face frame (src/prj_flux.c:274-349) is done with selects so
// that no register array is indexed at run time.
__device__ __forceinline__ double sel3(int dir, double a, double b, double c)
{
    return dir == 0 ? a : (dir == 1 ? b : c);
}

__device__ __forceinline__ int hlld_face(int dir, const double (*q)[9], const double* p,
                                         double gam, double bn, double* F6, double* fd3)
{
    double WL[9], WR[9];
#pragma unroll
    for (int lv = 0; lv < 9; ++lv) {
        if (lv == B1) continue;   // the normal field is the face value
        double s[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double* r = q[c];
            switch (lv) {
            case 1: s[c] = sel3(dir, r[1], r[2], r[3]); break;
            case 2: s[c] = sel3(dir, r[2], r[3], r[1]); break;
            case 3: s[c] = sel3(dir, r[3], r[1], r[2]); break;
            case 7: s[c] = sel3(dir, r[7], r[8], r[6]); break;
            case 8: s[c] = sel3(dir, r[8], r[6], r[7]); break;
            default: s[c] = r[lv]; break;
            }
        }
        WL[lv] = s[1] + mc_half(s[0], s[1], s[2]);
        WR[lv] = s[2] - mc_half(s[1], s[2], s[3]);
    }
    WL[B1] = bn;
    WR[B1] = bn;
    double pL = p[1] + mc_half(p[0], p[1], p[2]);
    double pR = p[2] - mc_half(p[1], p[2], p[3]);
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
template <int STAGE>
__global__ void __launch_bounds__(NT) mhd_flux_kernel(MhdStageArgs a)
{
    __shared__ double pl[NQ][PC];
    __shared__ double exy[6][EX];
    __shared__ double exz[6][EX];

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

    // column window: q[0..8] prims, q[9] pressure
    double cm[NQ], c0[NQ], cp[NQ], Ls[NQ], Fx[6];
    double hpf[9];   // prefetched halo cell of plane i+1

    auto load_own = [&](int p, double* c) {
#pragma unroll
        for (int q = 0; q < 9; ++q) c[q] = a.Win[q * vs + own_at(p)];
        c[9] = gm1 * c[0] * c[4];
    };
    auto load_halo = [&](int p) {
        if (has_halo) {
#pragma unroll
            for (int q = 0; q < 9; ++q) hpf[q] = a.Win[q * vs + halo_at(p)];
        }
    };
    auto store_plane = [&](const double* own) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) pl[q][oidx] = own[q];
        if (has_halo) {
#pragma unroll
            for (int q = 0; q < 9; ++q) pl[q][hidx] = hpf[q];
            pl[9][hidx] = gm1 * hpf[0] * hpf[4];
        }
    };
    // x-face between column cells L (cm) and R (c0) with right stencil cell cp;
    // Ls holds the left state on entry and the next face's left state on exit.
    auto x_face = [&](int xf, double* F6) {
        double WL[9], WR[9], pL, pR;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double h = mc_half(cm[q], c0[q], cp[q]);
            double r = c0[q] - h;
            if (q < 9) { WL[q] = Ls[q]; WR[q] = r; } else { pL = Ls[9]; pR = r; }
            Ls[q] = c0[q] + h;
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

    // ---- prologue ------------------------------------------------------
    {
        double cmm[NQ];
        load_own(-2, cmm);
        load_own(-1, cm);
        load_own(0, c0);
        load_own(1, cp);
        // left state of x-face 0 from cell -1
#pragma unroll
        for (int q = 0; q < NQ; ++q) Ls[q] = cm[q] + mc_half(cmm[q], cm[q], c0[q]);
        x_face(0, Fx);
        // plane 0 into LDS
        load_halo(0);
        // window for step 0: cm = cell 0, c0 = cell 1, cp = cell 2
#pragma unroll
        for (int q = 0; q < NQ; ++q) { cm[q] = c0[q]; c0[q] = cp[q]; }
        store_plane(cm);
        load_own(2, cp);
        load_halo(1);
        __syncthreads();
    }

    const double a1 = g.area[0], a2 = g.area[1], a3 = g.area[2], vol = g.vol;
    const double* bfy = a.bf_in + g.fvstride;
    const double* bfz = a.bf_in + 2 * g.fvstride;

    for (int i = 0; i < BS; ++i) {
        double w0[9];
        if (STAGE == 2) {
#pragma unroll
            for (int q = 0; q < 9; ++q) w0[q] = a.W0[q * vs + own_at(i)];
        }
        // y-face and z-face on the low side of the own cell; extra high-side
        // faces by half of one wave
        const int xw = i & 3;
        const int lane = t & 63;
        const bool extra = (t >> 6) == xw && lane < 32;
        for (int f = 0; f < 3; ++f) {
            if (f == 2 && !extra) continue;
            int dir, fj = j, fk = k;
            if (f == 0) dir = 1;
            else if (f == 1) dir = 2;
            else if (lane < 16) { dir = 1; fj = BS; fk = lane; }
            else { dir = 2; fj = lane - 16; fk = BS; }
            const int c = (fj + NG) * PW + (fk + NG);
            const int S = dir == 1 ? PW : 1;
            double q[4][9], p[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int cc = c + (s - 2) * S;
#pragma unroll
                for (int v = 0; v < 9; ++v) q[s][v] = pl[v][cc];
                p[s] = pl[9][cc];
            }
            const int gx = x0 + i, gy = y0 + fj, gz = z0 + fk;
            double bn = (dir == 1 ? bfy : bfz)[cell_off(g, gx, gy, gz)];
            double F6[6], fd3[3];
            int rc = hlld_face(dir, q, p, gam, bn, F6, fd3);
            if (rc) set_error(a.err, rc, dir, gx, gy, gz, STAGE);
            store_fdat(a, dir, gx, gy, gz, fd3);
            if (dir == 1) {
#pragma unroll
                for (int v = 0; v < 6; ++v) exy[v][fj * BS + fk] = F6[v];
            } else {
#pragma unroll
                for (int v = 0; v < 6; ++v) exz[v][fj * (BS + 1) + fk] = F6[v];
            }
        }
        double Fxn[6];
        x_face(i + 1, Fxn);
        __syncthreads();

        // divergence in the reference's x, y, z order and the RK combination
        double fd[6];
#pragma unroll
        for (int v = 0; v < 6; ++v) {
            fd[v] = -(a1 * (Fxn[v] - Fx[v]) + a2 * (exy[v][(j + 1) * BS + k] - exy[v][j * BS + k]) +
                      a3 * (exz[v][j * (BS + 1) + k + 1] - exz[v][j * (BS + 1) + k])) / vol;
        }
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
        {
            const int64_t o = own_at(i);
#pragma unroll
            for (int v = 0; v < 6; ++v) a.U1[v * vs + o] = u1[v];
        }
#pragma unroll
        for (int v = 0; v < 6; ++v) Fx[v] = Fxn[v];
        // advance: plane i+1 (= c0 + prefetched halo) into LDS
#pragma unroll
        for (int q = 0; q < NQ; ++q) { cm[q] = c0[q]; c0[q] = cp[q]; }
        __syncthreads();
        if (i < BS - 1) {
            store_plane(cm);
            if (i + 3 <= BS + 1) load_own(i + 3, cp);
            load_halo(i + 2);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Face EMF inputs on the transverse ghost layer of the domain: faces of
// direction d whose transverse coordinates include -1 or N (the reference's
// -1..16 face loops on boundary blocks, src/prj_flux.c:2241-2251).
__device__ __forceinline__ bool decode_ghostface(const Geom& g, int64_t id, int& dir, int& x,
                                                 int& y, int& z)
{
    // per direction: normal extent N_d + 1, transverse box (N_t + 2) x (N_u + 2)
    // minus the inner N_t x N_u
    const int n[3] = {g.nx, g.ny, g.nz};
    for (int d = 0; d < 3; ++d) {
        const int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
        const int64_t outer = (int64_t)(n[t1] + 2) * (n[t2] + 2);
        const int64_t ring = outer - (int64_t)n[t1] * n[t2];
        const int64_t cnt = (int64_t)(n[d] + 1) * ring;
        if (id < cnt) {
            const int nd = (int)(id / ring);
            int64_t r = id % ring;
            int c1, c2;
            // ring enumeration: two full rows (c1 = -1, N) then two side columns
            const int64_t rows = 2 * (int64_t)(n[t2] + 2);
            if (r < rows) {
                c1 = r < n[t2] + 2 ? -1 : n[t1];
                c2 = (int)(r % (n[t2] + 2)) - 1;
            } else {
                r -= rows;
                c1 = (int)(r >> 1);
                c2 = (r & 1) ? n[t2] : -1;
            }
            int c[3];
            c[d] = nd;
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
    const int e[3] = {dir == 0, dir == 1, dir == 2};
    double q[4][9], p[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int o = s - 2;
        const int64_t off = cell_off(g, x + o * e[0], y + o * e[1], z + o * e[2]);
#pragma unroll
        for (int v = 0; v < 9; ++v) q[s][v] = a.Win[v * g.vstride + off];
        p[s] = gm1 * q[s][0] * q[s][4];
    }
    double bn = bf_clamped(g, a.bf_in, dir, x, y, z);
    double F6[6], fd3[3];
    int rc = hlld_face(dir, q, p, g.gamma, bn, F6, fd3);
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
__global__ void __launch_bounds__(NT) mhd_ct_kernel(MhdStageArgs a)
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
    const int b = block_of_workgroup(blockIdx.x, gridDim.x);
    const int bz = b % g.nbz;
    const int by = (b / g.nbz) % g.nby;
    const int bx = b / (g.nbz * g.nby);
    const int x0 = bx * BS, y0 = by * BS, z0 = bz * BS;
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
    // the reference's order, `dn` the two cell widths
    auto ct = [&](int dir, int gx, int gy, int gz, double ea1, double ea0, double da, double eb1,
                  double eb0, double db) -> double {
        const int64_t o = dir * g.fvstride + cell_off(g, gx, gy, gz);
        double value = a.bf_in[o];
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

    for (int i = 0; i < BS; ++i) {
        const int s0 = i & 1;         // cell slot of plane i
        const int s1 = s0 ^ 1;        // plane i + 1
        const int e0 = i & 1;         // edge slot of x = i
        const int e1 = e0 ^ 1;        // x = i + 1
        // cell plane i+1 into the slot of plane i-1
        __syncthreads();
        cell_fill(s1, i + 1);
        __syncthreads();
        for (int e = t; e < EW * EW; e += NT) {
            int jj = e / EW, kk = e % EW;
            ex_e[e] = edge_x(i, jj, kk, s0);
        }
        for (int e = t; e < BS * EW; e += NT) {
            int jj = e / EW, kk = e % EW;
            ey_e[e1][e] = edge_y(i + 1, jj, kk, s0, s1);
        }
        for (int e = t; e < EW * BS; e += NT) {
            int jj = e / BS, kk = e % BS;
            ez_e[e1][e] = edge_z(i + 1, jj, kk, s0, s1);
        }
        __syncthreads();
        // CT of Bx(i+1), By(plane i), Bz(plane i)
        double bx_hi = ct(0, x0 + i + 1, y0 + j, z0 + k, ez_e[e1][(j + 1) * BS + k],
                          ez_e[e1][j * BS + k], g.dx[1], ey_e[e1][j * EW + k + 1],
                          ey_e[e1][j * EW + k], g.dx[2]);
        for (int f = t; f < EW * BS; f += NT) {
            int jj = f / BS, kk = f % BS;   // By(i, jj 0..16, kk 0..15)
            nby[f] = ct(1, x0 + i, y0 + jj, z0 + kk, ex_e[jj * EW + kk + 1], ex_e[jj * EW + kk],
                        g.dx[2], ez_e[e1][jj * BS + kk], ez_e[e0][jj * BS + kk], g.dx[0]);
        }
        for (int f = t; f < BS * EW; f += NT) {
            int jj = f / EW, kk = f % EW;   // Bz(i, jj 0..15, kk 0..16)
            nbz[f] = ct(2, x0 + i, y0 + jj, z0 + kk, ey_e[e1][jj * EW + kk], ey_e[e0][jj * EW + kk],
                        g.dx[0], ex_e[(jj + 1) * EW + kk], ex_e[jj * EW + kk], g.dx[1]);
        }
        __syncthreads();
        // cell-centred field from the new faces and cons -> prim
        {
            double bc[3];
            bc[0] = 0.5 * (bx_lo + bx_hi);
            bc[1] = 0.5 * (nby[j * BS + k] + nby[(j + 1) * BS + k]);
            bc[2] = 0.5 * (nbz[j * EW + k] + nbz[j * EW + k + 1]);
            const int64_t o = cell_off(g, x0 + i, y0 + j, z0 + k);
            double u[6], w[9];
#pragma unroll
            for (int v = 0; v < 6; ++v) u[v] = a.U1[v * vs + o];
            cons2prim9(u, bc, w);
#pragma unroll
            for (int v = 0; v < 9; ++v) a.Wout[v * vs + o] = w[v];
        }
        bx_lo = bx_hi;
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

hipError_t launch_mhd_stage(const MhdStageArgs& a, int stage, hipStream_t s)
{
    const int nblocks = a.g.nbx * a.g.nby * a.g.nbz;
    const int64_t ngf = ghostface_count(a.g);
    hipLaunchKernelGGL(mhd_ghostface_kernel, dim3((unsigned)((ngf + 255) / 256)), dim3(256), 0, s,
                       a, ngf, stage);
    if (stage == 1) {
        hipLaunchKernelGGL(mhd_flux_kernel<1>, dim3(nblocks), dim3(NT), 0, s, a);
        hipLaunchKernelGGL(mhd_ct_kernel<1>, dim3(nblocks), dim3(NT), 0, s, a);
    } else {
        hipLaunchKernelGGL(mhd_flux_kernel<2>, dim3(nblocks), dim3(NT), 0, s, a);
        hipLaunchKernelGGL(mhd_ct_kernel<2>, dim3(nblocks), dim3(NT), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_bf2bc(const Geom& g, double* W, const double* bf, hipStream_t s)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(bf2bc_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, W, bf);
    return hipGetLastError();
}

}  // namespace fx
