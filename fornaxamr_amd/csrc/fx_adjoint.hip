// Differentiable hydro scheme and its hand-written discrete adjoint.
//
// Scheme (not in the reference, which has MC/WENO and HLLC/HLLD): PLM with the
// smooth van Albada slope  (sl (sr^2+e) + sr (sl^2+e)) / (sl^2+sr^2+2e),
// reconstructing rho, v, p = (gamma-1) rho eint and Ye; HLLE fluxes with the
// Davis speeds of src/prj_riemann.c:1518-1519 in the HLL form of
// src/prj_riemann.c:393-407; flux divergence and per-cell RK2 update as in
// src/prj_flux.c:2523-2533 and src/prj_timeint.c:1255-1350, 1562-1655; fixed
// dt; periodic box.  The forward kernels round exactly as the CPU oracle of
// the same scheme.
//
// Layout: compact cell fields [var][X][Y][Z] (Z fastest), var = rho, v1, v2,
// v3, eint, Ye; N = X*Y*Z.  Periodic neighbours by index wrap, no ghosts.
//
// Adjoint.  A stage maps  Wout = prim(a cons(Wa) + b (cons(Wb) + dt L(Wb))).
// The cotangent is carried in conserved variables, Ubar = Jprim(W)^T Wbar;
// then for RK2 (stage 1: a = 0, b = 1 from W0; stage 2: a = b = 1/2 from W0,
// W1), with dL(W)^T taken through the per-face HLLE and per-cell slope VJPs:
//   Ubar1 = 1/2 Ubar2 + Jprim(W1)^T dL(W1)^T (dt/2 Ubar2)
//   Ubar0 = 1/2 Ubar2 + Ubar1 + Jprim(W0)^T dL(W0)^T (dt Ubar1)
// (Jprim(W)^T Jcons(W)^T = I removes the cons() terms).  At an exact tie of a
// min/max the cotangent is split evenly between both operands.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fx_kernels.hpp"

namespace fx {
namespace {

constexpr int NT = 256;

__device__ __forceinline__ int wrap(int i, int n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

__device__ __forceinline__ int64_t acell(const AdjGeom& g, int x, int y, int z)
{
    return ((int64_t)x * g.n[1] + y) * g.n[2] + z;
}

// cell (x,y,z) shifted by s along d
__device__ __forceinline__ int64_t ashift(const AdjGeom& g, int x, int y, int z, int d, int s)
{
    if (d == 0) x = wrap(x + s, g.n[0]);
    else if (d == 1) y = wrap(y + s, g.n[1]);
    else z = wrap(z + s, g.n[2]);
    return acell(g, x, y, z);
}

__device__ __forceinline__ void cell_q(const AdjGeom& g, const double* __restrict__ W, int64_t c,
                                       double* q)
{
    const int64_t N = g.N;
    q[0] = W[c];
    q[1] = W[N + c];
    q[2] = W[2 * N + c];
    q[3] = W[3 * N + c];
    q[4] = ((g.gamma - 1.0) * W[c]) * W[4 * N + c];
    q[5] = W[5 * N + c];
}

__device__ __forceinline__ double va(double sl, double sr, double e)
{
    const double sl2 = sl * sl, sr2 = sr * sr;
    const double num = sl * (sr2 + e) + sr * (sl2 + e);
    const double den = (sl2 + sr2) + 2.0 * e;
    return num / den;
}

// value and partial derivatives of the van Albada slope
__device__ __forceinline__ void va_d(double sl, double sr, double e, double& dl, double& dr)
{
    const double sl2 = sl * sl, sr2 = sr * sr;
    const double num = sl * (sr2 + e) + sr * (sl2 + e);
    const double den = (sl2 + sr2) + 2.0 * e;
    const double v = num / den;
    const double two_lr = 2.0 * sl * sr;
    dl = ((sr2 + e) + two_lr - v * (2.0 * sl)) / den;
    dr = (two_lr + (sl2 + e) - v * (2.0 * sr)) / den;
}

// rotated variable r of direction d: rho, v_n, v_t1, v_t2, p, Ye
__device__ __forceinline__ int rmap(int d, int r)
{
    if (r == 0) return 0;
    if (r >= 4) return r;
    return 1 + (d + r - 1) % 3;
}

struct HlleState {
    double U[6], F[6];
};

__device__ __forceinline__ void hlle_side(const double* a, double gm1, HlleState& s)
{
    const double m = a[0] * a[1];
    const double ke = (0.5 * a[0]) * ((a[1] * a[1] + a[2] * a[2]) + a[3] * a[3]);
    const double E = a[4] / gm1 + ke;
    s.U[0] = a[0];
    s.U[1] = m;
    s.U[2] = a[0] * a[2];
    s.U[3] = a[0] * a[3];
    s.U[4] = E;
    s.U[5] = a[0] * a[5];
    s.F[0] = m;
    s.F[1] = m * a[1] + a[4];
    s.F[2] = m * a[2];
    s.F[3] = m * a[3];
    s.F[4] = (E + a[4]) * a[1];
    s.F[5] = s.U[5] * a[1];
}

__device__ __forceinline__ void hlle(const double* qL, const double* qR, double gamma, double* F)
{
    const double gm1 = gamma - 1.0;
    const double cL = sqrt((gamma * qL[4]) / qL[0]);
    const double cR = sqrt((gamma * qR[4]) / qR[0]);
    const double a1 = qL[1] - cL, a2 = qR[1] - cR;
    const double b1 = qL[1] + cL, b2 = qR[1] + cR;
    const double SL = a1 < a2 ? a1 : a2;
    const double SR = b1 > b2 ? b1 : b2;
    const double bm = SL < 0.0 ? SL : 0.0;
    const double bp = SR > 0.0 ? SR : 0.0;
    HlleState L, R;
    hlle_side(qL, gm1, L);
    hlle_side(qR, gm1, R);
    const double bpm = bp * bm;
    const double den = bp - bm;
#pragma unroll
    for (int v = 0; v < 6; ++v) {
        const double num = (bp * L.F[v] - bm * R.F[v]) + bpm * (R.U[v] - L.U[v]);
        F[v] = num / den;
    }
}

// weight of the first operand of min/max in the cotangent (1, 0 or 1/2 at a tie)
__device__ __forceinline__ double wfirst_min(double a, double b) { return a < b ? 1.0 : (b < a ? 0.0 : 0.5); }
__device__ __forceinline__ double wfirst_max(double a, double b) { return a > b ? 1.0 : (b > a ? 0.0 : 0.5); }

// cotangent of one side's U and F back to its state a = (rho, u, v, w, p, Ye)
__device__ __forceinline__ void hlle_side_vjp(const double* a, double gm1, const double* Ub,
                                              const double* Fb, double* ab)
{
    const double rho = a[0], u = a[1], v = a[2], w = a[3], p = a[4], ye = a[5];
    const double m = rho * u;
    const double v2 = (u * u + v * v) + w * w;
    const double E = p / gm1 + (0.5 * rho) * v2;
    const double mb = Ub[1] + Fb[0] + Fb[1] * u + Fb[2] * v + Fb[3] * w;
    const double Eb = Ub[4] + Fb[4] * u;
    const double U5b = Ub[5] + Fb[5] * u;
    ab[0] += Ub[0] + Ub[2] * v + Ub[3] * w + U5b * ye + mb * u + Eb * (0.5 * v2);
    ab[1] += mb * rho + Fb[1] * m + Fb[4] * (E + p) + Fb[5] * (rho * ye) + Eb * (rho * u);
    ab[2] += Ub[2] * rho + Fb[2] * m + Eb * (rho * v);
    ab[3] += Ub[3] * rho + Fb[3] * m + Eb * (rho * w);
    ab[4] += Fb[1] + Fb[4] * u + Eb / gm1;
    ab[5] += U5b * rho;
}

// HLLE VJP: given Fbar, accumulate qLbar, qRbar (rotated order)
__device__ __forceinline__ void hlle_vjp(const double* qL, const double* qR, double gamma,
                                         const double* Fb, double* qLb, double* qRb)
{
    const double gm1 = gamma - 1.0;
    const double cL = sqrt((gamma * qL[4]) / qL[0]);
    const double cR = sqrt((gamma * qR[4]) / qR[0]);
    const double a1 = qL[1] - cL, a2 = qR[1] - cR;
    const double b1 = qL[1] + cL, b2 = qR[1] + cR;
    const double SL = a1 < a2 ? a1 : a2;
    const double SR = b1 > b2 ? b1 : b2;
    const double bm = SL < 0.0 ? SL : 0.0;
    const double bp = SR > 0.0 ? SR : 0.0;
    HlleState L, R;
    hlle_side(qL, gm1, L);
    hlle_side(qR, gm1, R);
    const double bpm = bp * bm;
    const double den = bp - bm;
    double bpb = 0.0, bmb = 0.0, bpmb = 0.0, denb = 0.0;
    double ULb[6], URb[6], FLb[6], FRb[6];
    const double inv_den = 1.0 / den;
#pragma unroll
    for (int v = 0; v < 6; ++v) {
        const double num = (bp * L.F[v] - bm * R.F[v]) + bpm * (R.U[v] - L.U[v]);
        const double F = num * inv_den;
        const double nb = Fb[v] * inv_den;
        denb -= nb * F;
        bpb += nb * L.F[v];
        bmb -= nb * R.F[v];
        bpmb += nb * (R.U[v] - L.U[v]);
        FLb[v] = nb * bp;
        FRb[v] = -nb * bm;
        URb[v] = nb * bpm;
        ULb[v] = -nb * bpm;
    }
    bpb += bpmb * bm + denb;
    bmb += bpmb * bp - denb;
    const double SRb = bpb * wfirst_max(SR, 0.0);
    const double SLb = bmb * wfirst_min(SL, 0.0);
    const double wa = wfirst_min(a1, a2), wb = wfirst_max(b1, b2);
    const double a1b = SLb * wa, a2b = SLb * (1.0 - wa);
    const double b1b = SRb * wb, b2b = SRb * (1.0 - wb);
    hlle_side_vjp(qL, gm1, ULb, FLb, qLb);
    hlle_side_vjp(qR, gm1, URb, FRb, qRb);
    // u -/+ c
    qLb[1] += a1b + b1b;
    qRb[1] += a2b + b2b;
    const double cLb = b1b - a1b, cRb = b2b - a2b;
    // c = sqrt(gamma p / rho): dc/dp = gamma / (2 c rho), dc/drho = -c / (2 rho)
    qLb[4] += cLb * (gamma / (2.0 * cL * qL[0]));
    qLb[0] -= cLb * (cL / (2.0 * qL[0]));
    qRb[4] += cRb * (gamma / (2.0 * cR * qR[0]));
    qRb[0] -= cRb * (cR / (2.0 * qR[0]));
}

// face states of the low face of cell c along d from cells c-2..c+1, in the
// face frame (rho, v_n, v_t1, v_t2, p, Ye); the cells are read in that order so
// that no register array is indexed by the direction at run time
__device__ __forceinline__ void face_states(const AdjGeom& g, const double* __restrict__ W, int d,
                                            int x, int y, int z, double* qL, double* qR)
{
    const int64_t N = g.N;
    const int64_t o1 = (int64_t)rmap(d, 1) * N, o2 = (int64_t)rmap(d, 2) * N,
                  o3 = (int64_t)rmap(d, 3) * N;
    double q[4][6];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int64_t c = ashift(g, x, y, z, d, s - 2);
        q[s][0] = W[c];
        q[s][1] = W[o1 + c];
        q[s][2] = W[o2 + c];
        q[s][3] = W[o3 + c];
        q[s][4] = ((g.gamma - 1.0) * W[c]) * W[4 * N + c];
        q[s][5] = W[5 * N + c];
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        const double dl = va(q[1][r] - q[0][r], q[2][r] - q[1][r], g.eps2);
        const double dr = va(q[2][r] - q[1][r], q[3][r] - q[2][r], g.eps2);
        qL[r] = q[1][r] + 0.5 * dl;
        qR[r] = q[2][r] - 0.5 * dr;
    }
}

__global__ void __launch_bounds__(NT) adj_flux_kernel(AdjGeom g, const double* __restrict__ W,
                                                     double* __restrict__ F)
{
    const int64_t c = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (c >= g.N) return;
    const int z = (int)(c % g.n[2]);
    const int y = (int)((c / g.n[2]) % g.n[1]);
    const int x = (int)(c / ((int64_t)g.n[2] * g.n[1]));
    for (int d = 0; d < 3; ++d) {
        double qL[6], qR[6], Fr[6];
        face_states(g, W, d, x, y, z, qL, qR);
        hlle(qL, qR, g.gamma, Fr);
#pragma unroll
        for (int r = 0; r < 6; ++r) F[((int64_t)d * 6 + rmap(d, r)) * g.N + c] = Fr[r];
    }
}

__device__ __forceinline__ void cons6(const double* w, double* u)
{
    u[0] = w[0];
    u[1] = w[0] * w[1];
    u[2] = w[0] * w[2];
    u[3] = w[0] * w[3];
    const double v2 = (w[1] * w[1] + w[2] * w[2]) + w[3] * w[3];
    u[4] = w[0] * w[4] + (0.5 * w[0]) * v2;
    u[5] = w[0] * w[5];
}

__device__ __forceinline__ void prim6(const double* u, double* w)
{
    w[0] = u[0];
    w[1] = u[1] / u[0];
    w[2] = u[2] / u[0];
    w[3] = u[3] / u[0];
    const double m2 = (u[1] * u[1] + u[2] * u[2]) + u[3] * u[3];
    w[4] = (u[4] - (0.5 * m2) / u[0]) / u[0];
    w[5] = u[5] / u[0];
}

template <int STAGE>
__global__ void __launch_bounds__(NT) adj_update_kernel(AdjGeom g, double dt,
                                                       const double* __restrict__ Win,
                                                       const double* W0,
                                                       const double* __restrict__ F, double* Wout)
{
    const int64_t c = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (c >= g.N) return;
    const int64_t N = g.N;
    const int z = (int)(c % g.n[2]);
    const int y = (int)((c / g.n[2]) % g.n[1]);
    const int x = (int)(c / ((int64_t)g.n[2] * g.n[1]));
    const int64_t hi[3] = {ashift(g, x, y, z, 0, 1), ashift(g, x, y, z, 1, 1),
                           ashift(g, x, y, z, 2, 1)};
    double w[6], u[6], u0[6], u1[6];
#pragma unroll
    for (int v = 0; v < 6; ++v) w[v] = Win[v * N + c];
    cons6(w, u);
    if (STAGE == 2) {
        double w0[6];
#pragma unroll
        for (int v = 0; v < 6; ++v) w0[v] = W0[v * N + c];
        cons6(w0, u0);
    }
#pragma unroll
    for (int v = 0; v < 6; ++v) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double* Fd = F + ((int64_t)d * 6 + v) * N;
            s = s + g.area[d] * (Fd[hi[d]] - Fd[c]);
        }
        const double fd = (0.0 - s) / g.vol;
        const double ut = u[v] + dt * fd;
        u1[v] = STAGE == 1 ? ut : 0.5 * u0[v] + 0.5 * ut;
    }
    double wn[6];
    prim6(u1, wn);
#pragma unroll
    for (int v = 0; v < 6; ++v) Wout[v * N + c] = wn[v];
}

// ---- reverse ---------------------------------------------------------------

// Ubar = Jprim(W)^T Wbar, W = (rho, v, eint, Ye)
__device__ __forceinline__ void jprim_t(const double* w, const double* Wb, double* Ub)
{
    const double rho = w[0], e = w[4], ye = w[5];
    const double ir = 1.0 / rho;
    const double v2 = (w[1] * w[1] + w[2] * w[2]) + w[3] * w[3];
    Ub[0] = Wb[0] + ir * (-(w[1] * Wb[1] + w[2] * Wb[2] + w[3] * Wb[3]) + (0.5 * v2 - e) * Wb[4] -
                          ye * Wb[5]);
    Ub[1] = ir * (Wb[1] - w[1] * Wb[4]);
    Ub[2] = ir * (Wb[2] - w[2] * Wb[4]);
    Ub[3] = ir * (Wb[3] - w[3] * Wb[4]);
    Ub[4] = ir * Wb[4];
    Ub[5] = ir * Wb[5];
}

// Wbar = Jcons(W)^T Ubar
__device__ __forceinline__ void jcons_t(const double* w, const double* Ub, double* Wb)
{
    const double rho = w[0], e = w[4], ye = w[5];
    const double v2 = (w[1] * w[1] + w[2] * w[2]) + w[3] * w[3];
    Wb[0] = Ub[0] + w[1] * Ub[1] + w[2] * Ub[2] + w[3] * Ub[3] + (e + 0.5 * v2) * Ub[4] + ye * Ub[5];
    Wb[1] = rho * (Ub[1] + w[1] * Ub[4]);
    Wb[2] = rho * (Ub[2] + w[2] * Ub[4]);
    Wb[3] = rho * (Ub[3] + w[3] * Ub[4]);
    Wb[4] = rho * Ub[4];
    Wb[5] = rho * Ub[5];
}

__global__ void __launch_bounds__(NT) adj_w2u_kernel(int64_t N, const double* __restrict__ W,
                                                    const double* __restrict__ Wb, double* Ub)
{
    const int64_t c = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (c >= N) return;
    double w[6], wb[6], ub[6];
#pragma unroll
    for (int v = 0; v < 6; ++v) {
        w[v] = W[v * N + c];
        wb[v] = Wb[v * N + c];
    }
    jprim_t(w, wb, ub);
#pragma unroll
    for (int v = 0; v < 6; ++v) Ub[v * N + c] = ub[v];
}

__global__ void __launch_bounds__(NT) adj_u2w_kernel(int64_t N, const double* __restrict__ W,
                                                    const double* __restrict__ Ub, double* Wb)
{
    const int64_t c = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (c >= N) return;
    double w[6], wb[6], ub[6];
#pragma unroll
    for (int v = 0; v < 6; ++v) {
        w[v] = W[v * N + c];
        ub[v] = Ub[v * N + c];
    }
    jcons_t(w, ub, wb);
#pragma unroll
    for (int v = 0; v < 6; ++v) Wb[v * N + c] = wb[v];
}

// Per low face of every cell and direction: Fbar = coef (A/vol) (lambda_c -
// lambda_{c-1}) with lambda = Ubar, then the HLLE VJP; writes the face-state
// cotangents in global variable order: G[(d*12 + v)*N + c] (left state),
// G[(d*12 + 6 + v)*N + c] (right state).
__global__ void __launch_bounds__(NT, 2) adj_face_vjp_kernel(AdjGeom g, const double* __restrict__ W,
                                                         const double* __restrict__ Ub,
                                                         double coef, double* __restrict__ G)
{
    const int64_t c = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (c >= g.N) return;
    const int64_t N = g.N;
    const int z = (int)(c % g.n[2]);
    const int y = (int)((c / g.n[2]) % g.n[1]);
    const int x = (int)(c / ((int64_t)g.n[2] * g.n[1]));
    for (int d = 0; d < 3; ++d) {
        const int64_t cm = ashift(g, x, y, z, d, -1);
        const double k = coef * (g.area[d] / g.vol);
        double qL[6], qR[6], Fb[6], qLb[6], qRb[6];
        face_states(g, W, d, x, y, z, qL, qR);
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const int v = rmap(d, r);
            Fb[r] = k * (Ub[v * N + c] - Ub[v * N + cm]);
            qLb[r] = 0.0;
            qRb[r] = 0.0;
        }
        hlle_vjp(qL, qR, g.gamma, Fb, qLb, qRb);
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const int v = rmap(d, r);
            G[((int64_t)d * 12 + v) * N + c] = qLb[r];
            G[((int64_t)d * 12 + 6 + v) * N + c] = qRb[r];
        }
    }
}

// Per cell: gather the face-state cotangents through the reconstruction
// (qL(i+1/2) = q_i + Delta_i/2, qR(i-1/2) = q_i - Delta_i/2) into qbar, then
// Wbar (p = (gamma-1) rho eint) and Ubar_out = a A + b B + Jprim(W)^T Wbar.
__global__ void __launch_bounds__(NT) adj_cell_vjp_kernel(AdjGeom g, const double* __restrict__ W,
                                                         const double* __restrict__ G,
                                                         double a, const double* A, double b,
                                                         const double* B, double* Uout)
{
    const int64_t c = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (c >= g.N) return;
    const int64_t N = g.N;
    const int z = (int)(c % g.n[2]);
    const int y = (int)((c / g.n[2]) % g.n[1]);
    const int x = (int)(c / ((int64_t)g.n[2] * g.n[1]));
    double qb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int d = 0; d < 3; ++d) {
        int64_t cc[5];
#pragma unroll
        for (int s = 0; s < 5; ++s) cc[s] = ashift(g, x, y, z, d, s - 2);
        double q[5][6];
#pragma unroll
        for (int s = 0; s < 5; ++s) cell_q(g, W, cc[s], q[s]);
        const double* GL = G + (int64_t)d * 12 * N;
        const double* GR = GL + 6 * N;
#pragma unroll
        for (int v = 0; v < 6; ++v) {
            // slot j holds the face j-1/2: GL = cotangent of its left state
            // (from cell j-1), GR = of its right state (from cell j)
            const double gL_c1 = GL[v * N + cc[3]];   // face c+1/2, left  = cell c
            const double gR_c0 = GR[v * N + cc[2]];   // face c-1/2, right = cell c
            double acc = gL_c1 + gR_c0;
            // slope cotangents of cells c-1, c, c+1: 1/2 GL(i+1/2) - 1/2 GR(i-1/2)
            const double gD_m = 0.5 * GL[v * N + cc[2]] - 0.5 * GR[v * N + cc[1]];
            const double gD_0 = 0.5 * gL_c1 - 0.5 * gR_c0;
            const double gD_p = 0.5 * GL[v * N + cc[4]] - 0.5 * GR[v * N + cc[3]];
            double dl, dr;
            va_d(q[1][v] - q[0][v], q[2][v] - q[1][v], g.eps2, dl, dr);   // Delta_{c-1}
            acc += dr * gD_m;
            va_d(q[2][v] - q[1][v], q[3][v] - q[2][v], g.eps2, dl, dr);   // Delta_c
            acc += (dl - dr) * gD_0;
            va_d(q[3][v] - q[2][v], q[4][v] - q[3][v], g.eps2, dl, dr);   // Delta_{c+1}
            acc -= dl * gD_p;
            qb[v] += acc;
        }
    }
    double w[6], wb[6], ub[6];
#pragma unroll
    for (int v = 0; v < 6; ++v) w[v] = W[v * N + c];
    const double gm1 = g.gamma - 1.0;
    wb[0] = qb[0] + (gm1 * w[4]) * qb[4];
    wb[1] = qb[1];
    wb[2] = qb[2];
    wb[3] = qb[3];
    wb[4] = (gm1 * w[0]) * qb[4];
    wb[5] = qb[5];
    jprim_t(w, wb, ub);
#pragma unroll
    for (int v = 0; v < 6; ++v) {
        double o = ub[v] + a * A[v * N + c];
        if (B) o += b * B[v * N + c];
        Uout[v * N + c] = o;
    }
}

inline dim3 grid_cells(int64_t N) { return dim3((unsigned)((N + NT - 1) / NT)); }

}  // namespace

hipError_t launch_adj_stage(const AdjGeom& g, int stage, double dt, const double* Win,
                            const double* W0, double* F, double* Wout, hipStream_t s)
{
    hipLaunchKernelGGL(adj_flux_kernel, grid_cells(g.N), dim3(NT), 0, s, g, Win, F);
    if (stage == 1)
        hipLaunchKernelGGL(adj_update_kernel<1>, grid_cells(g.N), dim3(NT), 0, s, g, dt, Win, W0, F, Wout);
    else
        hipLaunchKernelGGL(adj_update_kernel<2>, grid_cells(g.N), dim3(NT), 0, s, g, dt, Win, W0, F, Wout);
    return hipGetLastError();
}

hipError_t launch_adj_w2u(const AdjGeom& g, const double* W, const double* Wb, double* Ub,
                          hipStream_t s)
{
    hipLaunchKernelGGL(adj_w2u_kernel, grid_cells(g.N), dim3(NT), 0, s, g.N, W, Wb, Ub);
    return hipGetLastError();
}

hipError_t launch_adj_u2w(const AdjGeom& g, const double* W, const double* Ub, double* Wb,
                          hipStream_t s)
{
    hipLaunchKernelGGL(adj_u2w_kernel, grid_cells(g.N), dim3(NT), 0, s, g.N, W, Ub, Wb);
    return hipGetLastError();
}

// One reverse RK2 step: W0, W1 the stage inputs of the step, U2b the cotangent
// of its output (conserved form, overwritten by that of W0), U1b and G scratch.
hipError_t launch_adj_step_reverse(const AdjGeom& g, double dt, const double* W0, const double* W1,
                                   double* U2b, double* U1b, double* G, hipStream_t s)
{
    const dim3 grid = grid_cells(g.N);
    // stage 2: U1b = 1/2 U2b + Jprim(W1)^T dL(W1)^T (dt/2 U2b)
    hipLaunchKernelGGL(adj_face_vjp_kernel, grid, dim3(NT), 0, s, g, W1, (const double*)U2b,
                       0.5 * dt, G);
    hipLaunchKernelGGL(adj_cell_vjp_kernel, grid, dim3(NT), 0, s, g, W1, (const double*)G, 0.5,
                       (const double*)U2b, 0.0, (const double*)nullptr, U1b);
    // stage 1: U0b = 1/2 U2b + U1b + Jprim(W0)^T dL(W0)^T (dt U1b), into U2b
    hipLaunchKernelGGL(adj_face_vjp_kernel, grid, dim3(NT), 0, s, g, W0, (const double*)U1b, dt, G);
    hipLaunchKernelGGL(adj_cell_vjp_kernel, grid, dim3(NT), 0, s, g, W0, (const double*)G, 0.5,
                       (const double*)U2b, 1.0, (const double*)U1b, U2b);
    return hipGetLastError();
}

}  // namespace fx
