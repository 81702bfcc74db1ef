// Domain ghost fill, CFL time step and the per-face / per-edge batched entry
// points.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fx_kernels.hpp"
#include "fx_physics.hpp"

namespace fx {

namespace {

inline unsigned grid_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }


// outflow face-field ghost value: nearest storage face inside the domain
// (src/prj_boundary.c:902-995)
__device__ __forceinline__ double bf_clamped(const Geom& g, const double* bf, int dir, int x,
                                             int y, int z)
{
    int hx = g.nx - 1 + (dir == 0), hy = g.ny - 1 + (dir == 1), hz = g.nz - 1 + (dir == 2);
    x = x < 0 ? 0 : (x > hx ? hx : x);
    y = y < 0 ? 0 : (y > hy ? hy : y);
    z = z < 0 ? 0 : (z > hz ? hz : z);
    return bf[dir * g.fvstride + cell_off(g, x, y, z)];
}

// Ghost value of shell cell (x, y, z) from its source cell: 6 hydro variables
// (normal velocities negated across reflecting faces) and, for MHD, the cell
// field from the (ghost) faces.
__device__ __forceinline__ void fill_one(const Geom& g, double* __restrict__ W,
                                         const double* __restrict__ bf,
                                         double* __restrict__ zring, int x, int y, int z)
{
    int flip = 0;
    int mx = map_axis(g, 0, g.nx, x, flip);
    int my = map_axis(g, 1, g.ny, y, flip);
    int mz = map_axis(g, 2, g.nz, z, flip);
    const int64_t src = cell_off(g, mx, my, mz);
    const int64_t dst = cell_off(g, x, y, z);
    const int64_t vs = g.vstride;
    double val[9];
#pragma unroll
    for (int v = 0; v < 6; ++v) {
        val[v] = W[v * vs + src];
        if (v >= 1 && v <= 3 && (flip & (1 << (v - 1)))) val[v] = -val[v];
        W[v * vs + dst] = val[v];
    }
    if (g.mhd) {
        // cell B of a ghost cell is the average of its (ghost) faces, as
        // prj_mhd_bf2bc_all_mesh does after the ghost exchange
        // (src/prj_boundary.c:1234-1239; src/prj_mhd.c:555-603)
        double b1 = 0.5 * (bf_clamped(g, bf, 0, x, y, z) + bf_clamped(g, bf, 0, x + 1, y, z));
        double b2 = 0.5 * (bf_clamped(g, bf, 1, x, y, z) + bf_clamped(g, bf, 1, x, y + 1, z));
        double b3 = 0.5 * (bf_clamped(g, bf, 2, x, y, z) + bf_clamped(g, bf, 2, x, y, z + 1));
        val[6] = b1 / 1.0;
        val[7] = b2 / 1.0;
        val[8] = b3 / 1.0;
        W[6 * vs + dst] = val[6];
        W[7 * vs + dst] = val[7];
        W[8 * vs + dst] = val[8];
        if (zring != nullptr && (z == -1 || z == g.nz)) {
            // the compact copies of the planes z = -1 and z = nz
            const int64_t pl = zring_plane(g), vstr = 2 * pl;
            const int zc = z == -1 ? 0 : 1;
            const int64_t oy = zc * pl + (int64_t)(x + 2) * (g.ny + 4) + (y + 2);
            const int64_t ox = 18 * pl + zc * pl + (int64_t)(y + 2) * (g.nx + 4) + (x + 2);
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                zring[v * vstr + oy] = val[v];
                zring[v * vstr + ox] = val[v];
            }
        }
    }
}

// One item of the ghost shell, id in [0, shell_cell_count(g)): the x-part
// (planes x = -2, -1, nx, nx+1 incl. their y/z ghosts) and the y-part (rows
// y = -2, -1, ny, ny+1 of the interior x range incl. z ghosts) one cell per
// item, z fastest; the z-part the four ghost cells of an interior (x, y) row
// (z = -2, -1, nz, nz+1) as four adjacent items: one cell per item keeps
// every item's loads independent (two cells per item serialise, since the
// stores of the first may alias the loads of the second, and took the 256^3
// MHD fill from 80 to 154 us).  The MHD stage reads this shell; the hydro
// stage maps its out-of-domain reads onto their sources instead
// (fx_hydro.hip).
// The decomposition uses 32-bit unsigned arithmetic (a shell has < 2^31
// items): 64-bit division by a run-time divisor is a long software sequence.
__device__ __forceinline__ void fill_shell_cell(const Geom& g, double* __restrict__ W,
                                                const double* __restrict__ bf,
                                                double* __restrict__ zring, int64_t id64)
{
    const unsigned nyp = (unsigned)(g.ny + 2 * NG), nzp = (unsigned)(g.nz + 2 * NG);
    const unsigned nx_part = 4u * nyp * nzp;
    const unsigned ny_part = 4u * (unsigned)g.nx * nzp;
    const unsigned id = (unsigned)id64;
    if (id < nx_part) {
        unsigned r = id;
        const int z = (int)(r % nzp) - NG; r /= nzp;
        const int y = (int)(r % nyp) - NG; r /= nyp;
        const int s = (int)r;                     // 0..3
        fill_one(g, W, bf, zring, s < 2 ? s - NG : g.nx + (s - 2), y, z);
        return;
    }
    if (id < nx_part + ny_part) {
        unsigned r = id - nx_part;
        const int z = (int)(r % nzp) - NG; r /= nzp;
        const int s = (int)(r & 3u); r >>= 2;
        fill_one(g, W, bf, zring, (int)r, s < 2 ? s - NG : g.ny + (s - 2), z);
        return;
    }
    unsigned r = id - nx_part - ny_part;
    const int c = (int)(r & 3u); r >>= 2;         // 0, 1: z = -2, -1; 2, 3: nz, nz+1
    const int y = (int)(r % (unsigned)g.ny);
    const int x = (int)(r / (unsigned)g.ny);
    fill_one(g, W, bf, zring, x, y, c < 2 ? c - NG : g.nz + (c - 2));
}

__global__ void boundary_fill_kernel(Geom g, double* __restrict__ W, const double* __restrict__ bf,
                                     double* __restrict__ zring, int64_t total)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= total) return;
    fill_shell_cell(g, W, bf, zring, id);
}

// CFL denominator of one active cell (src/prj_timeint.c:2163-2182).  The
// reference takes min(1e99, min over cells of cfl / denom); correctly rounded
// division by a positive number is monotone, so that is min(1e99, cfl / max
// denom) exactly: the reductions carry the largest denominator (NaN never
// wins a comparison, as in the reference) and divide once.
__device__ __forceinline__ double cell_denom(const Geom& g, const double* __restrict__ W,
                                             int64_t o)
{
    const int64_t vs = g.vstride;
    return cfl_denom(g, W[o], W[vs + o], W[2 * vs + o], W[3 * vs + o], W[4 * vs + o]);
}

__device__ __forceinline__ double dt_from_denom(double cfl, double max_denom)
{
    const double dt = cfl / max_denom;
    return dt < 1.0e99 ? dt : 1.0e99;
}

// largest denominator of the 16^3 block (bx, by, bz); thread (j, k) walks x
__device__ __forceinline__ double block_max_denom(const Geom& g, const double* __restrict__ W,
                                                  int bx, int by, int bz, double* red)
{
    const int j = threadIdx.x >> 4, k = threadIdx.x & 15;
    const int64_t o0 = cell_off(g, bx * BS, by * BS + j, bz * BS + k);
    double best = 0.0;
    for (int i = 0; i < BS; ++i) {
        const double d = cell_denom(g, W, o0 + (int64_t)i * g.sx);
        if (d > best) best = d;
    }
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double o = red[threadIdx.x + w];
            if (o > red[threadIdx.x]) red[threadIdx.x] = o;
        }
        __syncthreads();
    }
    return red[0];
}

// src/prj_timeint.c:2144-2256 (sound speed only, also for MHD), in two
// launches without atomics: each workgroup streams a stretch of the cells in
// storage order (x, y, z) and writes its largest denominator to
// partial[blockIdx.x]; one workgroup then reduces the partials and forms dt
__global__ void __launch_bounds__(256) dt_reduce_kernel(Geom g, const double* __restrict__ W,
                                                        double* __restrict__ partial)
{
    __shared__ double red[256];
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    double best = 0.0;
    for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < n;
         id += (int64_t)gridDim.x * blockDim.x) {
        const unsigned u = (unsigned)id;     // cells < 2^31
        int z = (int)(u % (unsigned)g.nz);
        unsigned r = u / (unsigned)g.nz;
        int y = (int)(r % (unsigned)g.ny);
        int x = (int)(r / (unsigned)g.ny);
        const double d = cell_denom(g, W, cell_off(g, x, y, z));
        if (d > best) best = d;
    }
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && red[threadIdx.x + w] > red[threadIdx.x])
            red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// dt = min(dt_new, factor * dt_prev) after the first step (src/main.c:517-522);
// dt_new is the CFL minimum times the integrator's SSP scale (1 for RK2, 6 for
// eSSPRK(10,4): src/prj_timeint.c:2246-2253)
__device__ __forceinline__ void dt_finalize(double dt_new, double* dt_out, double* dt_prev,
                                            double factor, double scale)
{
    if (scale != 1.0) dt_new *= scale;
    double dt = dt_new;
    if (dt_prev != nullptr) {
        double prev = dt_prev[0];
        if (prev > 0.0) {
            double lim = factor * prev;
            dt = dt_new <= lim ? dt_new : lim;
        }
        dt_prev[0] = dt;
    }
    dt_out[0] = dt;
}

__global__ void __launch_bounds__(256) dt_final_kernel(const double* __restrict__ partial, int n,
                                                       double cfl, double* dt_out, double* dt_prev,
                                                       double factor, double scale)
{
    __shared__ double red[256];
    double best = 0.0;
    for (int b = threadIdx.x; b < n; b += blockDim.x)
        if (partial[b] > best) best = partial[b];
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && red[threadIdx.x + w] > red[threadIdx.x])
            red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) dt_finalize(dt_from_denom(cfl, red[0]), dt_out, dt_prev, factor, scale);
}

// ensembles: dev[2] of each instance accumulates the bits of its largest
// denominator (non-negative doubles order as their bit patterns)
__global__ void ens_dt_init_kernel(const EnsInst* __restrict__ inst, int ninst)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ninst) inst[i].dev[2] = 0.0;
}

// one workgroup per block of the table: 256 threads x 16 cells along x
__global__ void __launch_bounds__(256) ens_dt_reduce_kernel(const EnsInst* __restrict__ inst,
                                                            const EnsBlock* __restrict__ blocks)
{
    __shared__ double red[256];
    const EnsBlock e = blocks[blockIdx.x];
    const double m = block_max_denom(inst[e.inst].g, inst[e.inst].W[0], e.bx, e.by, e.bz, red);
    if (threadIdx.x == 0)
        atomicMax(reinterpret_cast<unsigned long long*>(inst[e.inst].dev + 2),
                  (unsigned long long)__double_as_longlong(m));
}

__global__ void ens_dt_finalize_kernel(const EnsInst* __restrict__ inst, int ninst, double cfl,
                                       double factor)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ninst) {
        double* d = inst[i].dev;
        dt_finalize(dt_from_denom(cfl, d[2]), d, d + 1, factor, 1.0);
    }
}

// prj_reconstruct_mc_face_slope over n stencils (qm, q0, qp)
__global__ void mc_slope_batch_kernel(int n, const double* qm, const double* q0, const double* qp,
                                      double* slope)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    slope[f] = mc_slope(qm[f], q0[f], qp[f]);
}

__global__ void hllc_batch_kernel(int n, const double* WL, const double* WR, const double* pL,
                                  const double* pR, const double* gL, const double* gR,
                                  const double* du, const double* dv, const double* dw,
                                  double* flux, double* vface)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    double wl[6], wr[6], F[6], vf[3];
#pragma unroll
    for (int v = 0; v < 6; ++v) {
        wl[v] = WL[(int64_t)v * n + f];
        wr[v] = WR[(int64_t)v * n + f];
    }
    hllc(wl, wr, pL[f], pR[f], gL[f], gR[f], du[f], dv[f], dw[f], F, vf);
#pragma unroll
    for (int v = 0; v < 6; ++v) flux[(int64_t)v * n + f] = F[v];
#pragma unroll
    for (int v = 0; v < 3; ++v) vface[(int64_t)v * n + f] = vf[v];
}

__global__ void hlld_batch_kernel(int n, const double* WL, const double* WR, const double* pL,
                                  const double* pR, const double* gL, const double* gR,
                                  const double* bn, double* flux, double* vface, double* bv1,
                                  double* bv2, int* status)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    double wl[9], wr[9], F[9], vf[3], e1 = 0.0, e2 = 0.0;
#pragma unroll
    for (int v = 0; v < 9; ++v) {
        wl[v] = WL[(int64_t)v * n + f];
        wr[v] = WR[(int64_t)v * n + f];
    }
    wl[B1] = bn[f];
    wr[B1] = bn[f];
#pragma unroll
    for (int v = 0; v < 9; ++v) F[v] = 0.0;
    vf[0] = vf[1] = vf[2] = 0.0;
    int rc = hlld(wl, wr, pL[f], pR[f], gL[f], gR[f], bn[f], F, vf, &e1, &e2);
    status[f] = rc;
#pragma unroll
    for (int v = 0; v < 9; ++v) flux[(int64_t)v * n + f] = F[v];
#pragma unroll
    for (int v = 0; v < 3; ++v) vface[(int64_t)v * n + f] = vf[v];
    bv1[f] = e1;
    bv2[f] = e2;
}

__global__ void emf_batch_kernel(int n, const double* ef, const double* ec, const double* vn,
                                 double* emf)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    double a[4], c[4], v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        a[q] = ef[(int64_t)q * n + f];
        c[q] = ec[(int64_t)q * n + f];
        v[q] = vn[(int64_t)q * n + f];
    }
    emf[f] = emf_upwind(a, c, v);
}

// reference block face index (src/prj_defs.h:394-399; 20^3 block, 2 ghosts)
__device__ __forceinline__ int block_face_idx(int d, int i, int j, int k)
{
    constexpr int P = BS + 2 * NG;
    if (d == 0) return ((i + NG) * P + (j + NG)) * P + (k + NG);
    if (d == 1) return ((i + NG) * (P + 1) + (j + NG)) * P + (k + NG);
    return ((i + NG) * P + (j + NG)) * (P + 1) + (k + NG);
}

// reference block edge index (src/prj_defs.h:400-405)
__device__ __forceinline__ int block_edge_idx(int d, int i, int j, int k)
{
    constexpr int P = BS + 2 * NG;
    if (d == 0) return ((i + NG) * (P + 1) + (j + NG)) * (P + 1) + (k + NG);
    if (d == 1) return ((i + NG) * P + (j + NG)) * (P + 1) + (k + NG);
    return ((i + NG) * (P + 1) + (j + NG)) * P + (k + NG);
}

__device__ __forceinline__ bool block_face_ok(int d, int i, int j, int k)
{
    constexpr int P = BS + 2 * NG;
    const int c[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (c[a] < -NG || c[a] > (a == d ? P - NG : P - NG - 1)) return false;
    return true;
}

// prj_mhd_emf_upwind with the resistive term (src/prj_mhd.c:627-718), one
// edge of one block per thread.  status: 0, 37 edge outside the block's edge
// storage (src/prj_mhd.c:636-640), 31 non-finite input, 38 a face array
// missing for a nonzero eta, 35 a curl read outside the block's face storage
// (out of bounds in the reference), 36 non-finite curl B, 39 non-finite result
// (src/prj_mhd.c:707-715).
__global__ void emf_block_batch_kernel(EmfBlockArgs a)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.n) return;
    const int dir = a.edge[4 * f], i = a.edge[4 * f + 1], j = a.edge[4 * f + 2],
              k = a.edge[4 * f + 3];
    int st = 0;
    double emf = __builtin_nan("");
    if (dir < 0 || dir > 2 || i < -NG || i >= BS + NG || j < -NG || j >= BS + NG || k < -NG ||
        k >= BS + NG) {
        st = 37;
    } else {
        double ef[4], ec[4], vn[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            ef[q] = a.ef[(int64_t)q * a.n + f];
            ec[q] = a.ec[(int64_t)q * a.n + f];
            vn[q] = a.vn[(int64_t)q * a.n + f];
            if (!isfinite(ef[q]) || !isfinite(ec[q]) || !isfinite(vn[q])) st = 31;
        }
        if (st == 0) {
            double e = emf_upwind(ef, ec, vn);
            const double eta = a.eta[f];
            if (eta != 0.0) {
                // components and shifts of the two differences (src/prj_mhd.c:694-706)
                const int fa = dir == 0 ? 2 : (dir == 1 ? 0 : 1);
                const int fb = dir == 0 ? 1 : (dir == 1 ? 2 : 0);
                const int ia = i - (dir == 2), ja = j - (dir == 0), ka = k - (dir == 1);
                const int ib = i - (dir == 1), jb = j - (dir == 2), kb = k - (dir == 0);
                const double wa = a.dx[fa == 2 ? 1 : (fa == 0 ? 2 : 0)];
                const double wb = a.dx[fb == 1 ? 2 : (fb == 2 ? 0 : 1)];
                if (a.bf[fa] == nullptr || a.bf[fb] == nullptr) {
                    st = 38;
                } else if (!block_face_ok(fa, i, j, k) || !block_face_ok(fa, ia, ja, ka) ||
                           !block_face_ok(fb, i, j, k) || !block_face_ok(fb, ib, jb, kb)) {
                    st = 35;
                } else {
                    const double curl_b = (a.bf[fa][block_face_idx(fa, i, j, k)] -
                                           a.bf[fa][block_face_idx(fa, ia, ja, ka)]) / wa -
                                          (a.bf[fb][block_face_idx(fb, i, j, k)] -
                                           a.bf[fb][block_face_idx(fb, ib, jb, kb)]) / wb;
                    if (!isfinite(curl_b)) st = 36;
                    else e += eta * curl_b;
                }
            }
            if (st == 0 && !isfinite(e)) st = 39;
            if (st == 0) {
                emf = e;
                if (a.block_emf[dir] != nullptr) a.block_emf[dir][block_edge_idx(dir, i, j, k)] = e;
            }
        }
    }
    a.emf[f] = emf;
    a.status[f] = st;
}

}  // namespace

// number of items of the ghost-shell enumeration of fill_shell_cell
static int64_t shell_cell_count(const Geom& g)
{
    // items: one cell each (x-part, y-part, z-part)
    const int64_t nyp = g.ny + 2 * NG, nzp = g.nz + 2 * NG;
    return 4 * nyp * nzp + 4 * (int64_t)g.nx * nzp + 4 * (int64_t)g.nx * g.ny;
}

// eSSPRK convex blend of the active cells (src/prj_timeint.c:2382-2417, hydro):
// u = w_s u(W_saved) + (1 - w_s) u(W), W <- prim(u)
__global__ void __launch_bounds__(256) blend_kernel(Geom g, double* __restrict__ W,
                                                    const double* __restrict__ Ws,
                                                    double saved_weight)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const unsigned cid = (unsigned)id;        // cells < 2^31
    const int z = (int)(cid % (unsigned)g.nz);
    const unsigned r = cid / (unsigned)g.nz;
    const int y = (int)(r % (unsigned)g.ny);
    const int x = (int)(r / (unsigned)g.ny);
    const int64_t o = cell_off(g, x, y, z);
    const int64_t vs = g.vstride;
    const double current_weight = 1.0 - saved_weight;
    double w[6], ws[6];
#pragma unroll
    for (int v = 0; v < 6; ++v) {
        w[v] = W[v * vs + o];
        ws[v] = Ws[v * vs + o];
    }
    double u[6], us[6];
    u[0] = w[0];
    u[1] = w[0] * w[1];
    u[2] = w[0] * w[2];
    u[3] = w[0] * w[3];
    u[4] = w[0] * w[4] + 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
    u[5] = w[0] * w[5];
    us[0] = ws[0];
    us[1] = ws[0] * ws[1];
    us[2] = ws[0] * ws[2];
    us[3] = ws[0] * ws[3];
    us[4] = ws[0] * ws[4] + 0.5 * ws[0] * (ws[1] * ws[1] + ws[2] * ws[2] + ws[3] * ws[3]);
    us[5] = ws[0] * ws[5];
#pragma unroll
    for (int v = 0; v < 6; ++v) u[v] = saved_weight * us[v] + current_weight * u[v];
    // cons -> prim (src/prj_eos.c:2160-2220)
    const double rho = u[0];
    if (rho == 0.0) {
#pragma unroll
        for (int v = 0; v < 6; ++v) W[v * vs + o] = 0.0;
        return;
    }
    const double v1 = u[1] / rho, v2 = u[2] / rho, v3 = u[3] / rho;
    const double kinetic = 0.5 * (v1 * v1 + v2 * v2 + v3 * v3);
    W[o] = rho;
    W[vs + o] = v1;
    W[2 * vs + o] = v2;
    W[3 * vs + o] = v3;
    W[4 * vs + o] = u[4] / rho - kinetic;
    W[5 * vs + o] = u[5] / rho;
}

// eSSPRK convex blend of the face fields (src/prj_timeint.c:2347-2371): every
// active face bf = w_s bf_saved + (1 - w_s) bf.  Items enumerate the three
// components' active faces, x-faces first.
__global__ void __launch_bounds__(256) blend_faces_kernel(Geom g, double* __restrict__ bf,
                                                          const double* __restrict__ bfs,
                                                          double saved_weight)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nx1 = (int64_t)(g.nx + 1) * g.ny * g.nz, ny1 = (int64_t)g.nx * (g.ny + 1) * g.nz;
    const int64_t nz1 = (int64_t)g.nx * g.ny * (g.nz + 1);
    if (id >= nx1 + ny1 + nz1) return;
    int d;
    int64_t f;
    if (id < nx1) { d = 0; f = id; }
    else if (id < nx1 + ny1) { d = 1; f = id - nx1; }
    else { d = 2; f = id - nx1 - ny1; }
    const int ez = g.nz + (d == 2), ey = g.ny + (d == 1);
    const int z = (int)(f % ez);
    const int64_t r = f / ez;
    const int y = (int)(r % ey);
    const int x = (int)(r / ey);
    const int64_t o = d * g.fvstride + cell_off(g, x, y, z);
    bf[o] = saved_weight * bfs[o] + (1.0 - saved_weight) * bf[o];
}

// eSSPRK convex blend of the active cells of an MHD mesh
// (src/prj_timeint.c:2382-2417): u = w_s u(W_saved) + (1 - w_s) u(W) with
// both states' own cell field in their energies, then the cell field from the
// (already blended) faces, W <- prim(u).
__global__ void __launch_bounds__(256) blend_mhd_kernel(Geom g, double* __restrict__ W,
                                                        const double* __restrict__ Ws,
                                                        const double* __restrict__ bf,
                                                        double saved_weight)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const unsigned cid = (unsigned)id;        // cells < 2^31
    const int z = (int)(cid % (unsigned)g.nz);
    const unsigned r = cid / (unsigned)g.nz;
    const int y = (int)(r % (unsigned)g.ny);
    const int x = (int)(r / (unsigned)g.ny);
    const int64_t o = cell_off(g, x, y, z);
    const int64_t vs = g.vstride;
    const double current_weight = 1.0 - saved_weight;
    double w[9], ws[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) {
        w[v] = W[v * vs + o];
        ws[v] = Ws[v * vs + o];
    }
    // prim -> cons with field (src/prj_eos.c:2112-2158)
    double u[6], us[6];
    u[0] = w[0];
    u[1] = w[0] * w[1];
    u[2] = w[0] * w[2];
    u[3] = w[0] * w[3];
    u[4] = w[0] * w[4] + 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) +
           0.5 * (w[6] * w[6] + w[7] * w[7] + w[8] * w[8]);
    u[5] = w[0] * w[5];
    us[0] = ws[0];
    us[1] = ws[0] * ws[1];
    us[2] = ws[0] * ws[2];
    us[3] = ws[0] * ws[3];
    us[4] = ws[0] * ws[4] + 0.5 * ws[0] * (ws[1] * ws[1] + ws[2] * ws[2] + ws[3] * ws[3]) +
            0.5 * (ws[6] * ws[6] + ws[7] * ws[7] + ws[8] * ws[8]);
    us[5] = ws[0] * ws[5];
#pragma unroll
    for (int v = 0; v < 6; ++v) u[v] = saved_weight * us[v] + current_weight * u[v];
    // cell field from the blended faces (src/prj_timeint.c:498-510)
    const double* bx = bf;
    const double* by = bf + g.fvstride;
    const double* bz = bf + 2 * g.fvstride;
    const double b1 = 0.5 * (bx[o] + bx[o + g.sx]);
    const double b2 = 0.5 * (by[o] + by[o + g.sy]);
    const double b3 = 0.5 * (bz[o] + bz[o + 1]);
    // cons -> prim with field (src/prj_eos.c:2160-2220)
    const double rho = u[0];
    if (rho == 0.0) {
#pragma unroll
        for (int v = 0; v < 9; ++v) W[v * vs + o] = 0.0;
        return;
    }
    const double v1 = u[1] / rho, v2 = u[2] / rho, v3 = u[3] / rho;
    const double kinetic = 0.5 * (v1 * v1 + v2 * v2 + v3 * v3);
    const double magnetic = 0.5 * (b1 * b1 + b2 * b2 + b3 * b3) / rho;
    W[o] = rho;
    W[vs + o] = v1;
    W[2 * vs + o] = v2;
    W[3 * vs + o] = v3;
    W[4 * vs + o] = u[4] / rho - kinetic - magnetic;
    W[5 * vs + o] = u[5] / rho;
    W[6 * vs + o] = b1;
    W[7 * vs + o] = b2;
    W[8 * vs + o] = b3;
}

__global__ void scale_kernel(const double* src, double* dst, double div)
{
    dst[0] = src[0] / div;
}

hipError_t launch_blend(const Geom& g, double* W, const double* Ws, double saved_weight,
                        hipStream_t s)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(blend_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, g, W, Ws,
                       saved_weight);
    return hipGetLastError();
}

hipError_t launch_blend_mhd(const Geom& g, double* W, const double* Ws, double* bf,
                            const double* bfs, double saved_weight, hipStream_t s)
{
    const int64_t nf = (int64_t)(g.nx + 1) * g.ny * g.nz + (int64_t)g.nx * (g.ny + 1) * g.nz +
                       (int64_t)g.nx * g.ny * (g.nz + 1);
    hipLaunchKernelGGL(blend_faces_kernel, dim3(grid_for(nf, 256)), dim3(256), 0, s, g, bf, bfs,
                       saved_weight);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(blend_mhd_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, g, W, Ws, bf,
                       saved_weight);
    return hipGetLastError();
}

hipError_t launch_div_scalar(const double* src, double* dst, double div, hipStream_t s)
{
    hipLaunchKernelGGL(scale_kernel, dim3(1), dim3(1), 0, s, src, dst, div);
    return hipGetLastError();
}

hipError_t launch_boundary_fill(const Geom& g, double* W, const double* bf, double* zring,
                                hipStream_t s)
{
    const int64_t total = shell_cell_count(g);
    hipLaunchKernelGGL(boundary_fill_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, g, W,
                       bf, zring, total);
    return hipGetLastError();
}

int calc_dt_wgs(const Geom& g)
{
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    return (int)std::min<int64_t>(CALC_DT_WGS, grid_for(n, 256));
}

hipError_t launch_dt_reduce(const Geom& g, const double* W, double* partial, int nwg,
                            hipStream_t s)
{
    hipLaunchKernelGGL(dt_reduce_kernel, dim3(nwg), dim3(256), 0, s, g, W, partial);
    return hipGetLastError();
}

hipError_t launch_dt_final(const double* partial, int n, double cfl, double* dt_out,
                           double* dt_prev, double factor, double scale, hipStream_t s)
{
    hipLaunchKernelGGL(dt_final_kernel, dim3(1), dim3(256), 0, s, partial, n, cfl, dt_out,
                       dt_prev, factor, scale);
    return hipGetLastError();
}

hipError_t launch_calc_dt(const Geom& g, const double* W, double cfl, double* partial,
                          double* dt_out, double* dt_prev, double factor, hipStream_t s,
                          double scale)
{
    const int nwg = calc_dt_wgs(g);
    if (hipError_t e = launch_dt_reduce(g, W, partial, nwg, s)) return e;
    return launch_dt_final(partial, nwg, cfl, dt_out, dt_prev, factor, scale, s);
}

hipError_t launch_ens_dt_reduce(const EnsInst* inst, int ninst, const EnsBlock* blocks,
                                int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(ens_dt_init_kernel, dim3(grid_for(ninst, 64)), dim3(64), 0, s, inst, ninst);
    if (nblocks > 0)
        hipLaunchKernelGGL(ens_dt_reduce_kernel, dim3(nblocks), dim3(256), 0, s, inst, blocks);
    return hipGetLastError();
}

hipError_t launch_ens_dt_clear(const EnsInst* inst, int ninst, hipStream_t s)
{
    hipLaunchKernelGGL(ens_dt_init_kernel, dim3(grid_for(ninst, 64)), dim3(64), 0, s, inst, ninst);
    return hipGetLastError();
}

hipError_t launch_ens_dt_finalize(const EnsInst* inst, int ninst, double cfl, double factor,
                                  hipStream_t s)
{
    hipLaunchKernelGGL(ens_dt_finalize_kernel, dim3(grid_for(ninst, 64)), dim3(64), 0, s, inst,
                       ninst, cfl, factor);
    return hipGetLastError();
}

hipError_t launch_mc_slope_batch(int n, const double* qm, const double* q0, const double* qp,
                                double* slope, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mc_slope_batch_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, qm, q0,
                       qp, slope);
    return hipGetLastError();
}

hipError_t launch_hllc_batch(int n, const double* WL, const double* WR, const double* pL,
                             const double* pR, const double* gL, const double* gR,
                             const double* du, const double* dv, const double* dw,
                             double* flux, double* vface, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(hllc_batch_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, WL, WR,
                       pL, pR, gL, gR, du, dv, dw, flux, vface);
    return hipGetLastError();
}

hipError_t launch_hlld_batch(int n, const double* WL, const double* WR, const double* pL,
                             const double* pR, const double* gL, const double* gR,
                             const double* bn, double* flux, double* vface, double* bv1,
                             double* bv2, int* status, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(hlld_batch_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, WL, WR,
                       pL, pR, gL, gR, bn, flux, vface, bv1, bv2, status);
    return hipGetLastError();
}

hipError_t launch_emf_block_batch(const EmfBlockArgs& a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(emf_block_batch_kernel, dim3(grid_for(a.n, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_emf_batch(int n, const double* ef, const double* ec, const double* vn,
                            double* emf, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(emf_batch_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, ef, ec, vn,
                       emf);
    return hipGetLastError();
}

}  // namespace fx
