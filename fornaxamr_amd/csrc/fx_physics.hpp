// Per-cell / per-face / per-edge physics of the RK2 Godunov stage, as device
// functions for gfx950.  FP64 throughout.
//
// Every expression keeps the operation order and association of the reference
// (cited per function) and the library is built with -ffp-contract=off, so the
// results are bit-identical to the reference CPU path whenever the FP64
// division and square root round correctly (they do on gfx950).  Min/max are
// written as the reference's ternaries, not fmin/fmax, to keep signed-zero and
// NaN behaviour identical.
#pragma once

#include <hip/hip_runtime.h>

#include "fx_log.hpp"

namespace fx {

// Largest of the workgroup's values (none of them NaN; blockDim.x a multiple
// of 64), returned to every thread.  lds: one double per wave of LDS that no
// thread touches until the next barrier after this call.
__device__ __forceinline__ double workgroup_max(double v, double* lds)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off);
        if (o > v) v = o;
    }
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = lds[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
        if (lds[w] > r) r = lds[w];
    return r;
}

// primitive / conserved variable slots (src/prj_defs.h:466-494)
enum { RHO = 0, V1 = 1, V2 = 2, V3 = 3, EINT = 4, YE = 5, B1 = 6, B2 = 7, B3 = 8 };
enum { MOM1 = 1, MOM2 = 2, MOM3 = 3, ETOT = 4 };

// local (face-frame) -> global primitive / conserved slot of direction dir
// (src/prj_flux.c:274-349): cyclic permutation of the vector components
__device__ __forceinline__ int gmap(int dir, int lv)
{
    if (lv == 0 || lv == 4 || lv == 5) return lv;
    if (lv <= 3) return 1 + (lv - 1 + dir) % 3;
    return 6 + (lv - 6 + dir) % 3;
}

__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }

// Half slope as the flux assembly uses it (src/prj_flux.c:410-418):
// h = 0.5 * prj_reconstruct_mc_face_slope(qm, q0, qp)
// (src/prj_reconstruct.h:355-375), evaluated as
//   (sl*sr > 0) ? 0.5 * copysign(min(|2 sl|, |c1|, |2 sr|), sl) : 0.
// When sl*sr > 0 the three candidates c0 = 2 sl, c1 = sl + (sr - sl)/2 and
// c2 = 2 sr are nonzero with the sign of sl, so min(c0,c1,c2) for sl > 0 and
// max(c0,c1,c2) for sl < 0 both equal copysign(min|c|, sl) exactly (negation
// and scaling by 2 are exact); otherwise both forms return +0.  Inputs with
// NaN give sl*sr > 0 false and 0 in both.  min(|2 sl|, |2 sr|) is taken as
// 2 min(|sl|, |sr|): doubling is exact (or overflows alike) and monotone, and
// a NaN c1 (from infinite slopes) is ignored by fmin in either order.  The
// same bits, half the instructions of the compare/select chain.
__device__ __forceinline__ double mc_half(double qm, double q0, double qp)
{
    double sl = q0 - qm;
    double sr = qp - q0;
    double c1 = sl + 0.5 * (sr - sl);
    double m = fmin(2.0 * fmin(fabs(sl), fabs(sr)), fabs(c1));
    double h = 0.5 * copysign(m, sl);
    return (sl * sr > 0.0) ? h : 0.0;
}

// The full MC face slope, prj_reconstruct_mc_face_slope (src/prj_reconstruct.h:
// 355-375), in the form of mc_half above: mc_half(...) == 0.5 * mc_slope(...)
// bit for bit (the same copysign(min|c|, sl), then the same product by 0.5).
__device__ __forceinline__ double mc_slope(double qm, double q0, double qp)
{
    double sl = q0 - qm;
    double sr = qp - q0;
    double c1 = sl + 0.5 * (sr - sl);
    double m = fmin(2.0 * fmin(fabs(sl), fabs(sr)), fabs(c1));
    double s = copysign(m, sl);
    return (sl * sr > 0.0) ? s : 0.0;
}

// Shock-detector pressure-jump limiter (src/prj_riemann.c:8-17).  The
// velocity deltas enter only through -min(., 0) + cfmax with cfmax > 0, so a
// min that differs from the reference's ternary only in the sign of a zero
// result gives the same theta; fmin (one instruction) is used throughout the
// delta chain.
__device__ __forceinline__ double theta_limiter(double cfmax, double du, double dv, double dw)
{
    double t = (-fmin(du, 0.0) + cfmax) / (-fmin(fmin(dv, dw), 0.0) + cfmax);
    t *= t;
    return fmin(1.0, t * t);
}

// min used for the velocity deltas (see theta_limiter)
__device__ __forceinline__ double dmin_delta(double a, double b) { return fmin(a, b); }

// HLLC with the theta-limited contact speed (src/prj_riemann.c:1433-1599),
// hydro build (6 variables).  W* are face states in the rotated frame.
//
// Written without branches: the reference returns F_L where 0 <= S_L, F_R
// where S_R <= 0, and otherwise the star flux of the side the contact speed
// selects (Toro 2009, eqs. 10.33-10.36).  Here the upwind side K is chosen
// once -- L if 0 <= S_L, R if S_R <= 0, else by the sign of S_M -- and F_K
// (formed from K's state by the same expressions as the reference's
// supersonic fluxes) is taken as it is in the supersonic cases and as the base
// of the star flux otherwise.  The star terms are evaluated in every case and
// discarded by a select where the reference would not have formed them, so
// every returned bit is the reference's; straight-line code lets the compiler
// interleave the independent face solves of a cell.
__device__ __forceinline__ void hllc(const double* WL, const double* WR, double pL, double pR,
                                     double gL, double gR, double du, double dv, double dw,
                                     double* flux, double* vface)
{
    // prj_riemann_state for both sides
    double rL = WL[RHO], uL = WL[V1], vL = WL[V2], wL = WL[V3], yL = WL[YE];
    double rR = WR[RHO], uR = WR[V1], vR = WR[V2], wR = WR[V3], yR = WR[YE];
    double eL = rL * WL[EINT] + 0.5 * rL * (uL * uL + vL * vL + wL * wL);
    double eR = rR * WR[EINT] + 0.5 * rR * (uR * uR + vR * vR + wR * wR);
    double csL = sqrt(gL * pL / rL);
    double csR = sqrt(gR * pR / rR);
    double UL1 = rL * uL, UR1 = rR * uR;

    double SL = dmin(uL - csL, uR - csR);
    double SR = dmax(uL + csL, uR + csR);
    double cfmax = dmax(csL, csR);
    double theta = theta_limiter(cfmax, du, dv, dw);
    double SM = (theta * (pR - pL) + UL1 * (SL - uL) - UR1 * (SR - uR)) /
                (rL * (SL - uL) - rR * (SR - uR));

    const bool sup_l = 0.0 <= SL;
    const bool sup_r = !sup_l && SR <= 0.0;
    const bool left = sup_l || (!sup_r && 0.0 <= SM);
    double r = left ? rL : rR;
    double u = left ? uL : uR;
    double v = left ? vL : vR;
    double w = left ? wL : wR;
    double y = left ? yL : yR;
    double e = left ? eL : eR;
    double p = left ? pL : pR;
    double S = left ? SL : SR;
    double Fr = r * u;
    double Fm1 = r * u * u + p;
    double Fm2 = r * u * v;
    double Fm3 = r * u * w;
    double Fe = (e + p) * u;
    double Fy = r * y * u;
    double rho_star = r * (S - u) / (S - SM);
    double e_over_rho = e / r + (SM - u) * (SM + p / (r * (S - u)));
    // star fluxes formed before the selects, so that each select picks
    // between two computed values (a conditional expression would be emitted
    // as a branch around the star terms)
    const double s0 = Fr + S * (rho_star - r);
    const double s1 = Fm1 + S * (rho_star * SM - r * u);
    const double s2 = Fm2 + S * (rho_star * v - r * v);
    const double s3 = Fm3 + S * (rho_star * w - r * w);
    const double s4 = Fe + S * (rho_star * e_over_rho - e);
    const double s5 = Fy + S * (rho_star * y - r * y);
    const bool sup = sup_l || sup_r;
    flux[0] = sup ? Fr : s0;
    flux[1] = sup ? Fm1 : s1;
    flux[2] = sup ? Fm2 : s2;
    flux[3] = sup ? Fm3 : s3;
    flux[4] = sup ? Fe : s4;
    flux[5] = sup ? Fy : s5;
    vface[0] = sup ? u : SM;
    vface[1] = v;
    vface[2] = w;
}

// HLLE: the HLL flux (Harten, Lax & van Leer 1983) over the Davis wave-speed
// bounds SL = min(uL - cL, uR - cR), SR = max(uL + cL, uR + cR) that the
// reference's HLLC takes (src/prj_riemann.c:1518-1519), on the reference's
// face states (prj_riemann_state, src/prj_riemann.c:1433-1492).  The flux and
// the face velocity (that of the HLL average state) are written in the form of
// the reference's HLL fallback inside HLLD (src/prj_riemann.c:393-407).  Not
// a reference solver: the reference has no hydro HLLE.
__device__ __forceinline__ void hlle_hydro(const double* WL, const double* WR, double pL,
                                           double pR, double gL, double gR, double* flux,
                                           double* vface)
{
    double rL = WL[RHO], uL = WL[V1], vL = WL[V2], wL = WL[V3], yL = WL[YE];
    double rR = WR[RHO], uR = WR[V1], vR = WR[V2], wR = WR[V3], yR = WR[YE];
    double eL = rL * WL[EINT] + 0.5 * rL * (uL * uL + vL * vL + wL * wL);
    double eR = rR * WR[EINT] + 0.5 * rR * (uR * uR + vR * vR + wR * wR);
    double csL = sqrt(gL * pL / rL);
    double csR = sqrt(gR * pR / rR);
    double UL[6] = {rL, rL * uL, rL * vL, rL * wL, eL, rL * yL};
    double UR[6] = {rR, rR * uR, rR * vR, rR * wR, eR, rR * yR};
    double FL[6] = {rL * uL, rL * uL * uL + pL, rL * uL * vL, rL * uL * wL, (eL + pL) * uL,
                    rL * yL * uL};
    double FR[6] = {rR * uR, rR * uR * uR + pR, rR * uR * vR, rR * uR * wR, (eR + pR) * uR,
                    rR * yR * uR};
    double SL = dmin(uL - csL, uR - csR);
    double SR = dmax(uL + csL, uR + csR);
    // supersonic faces take the upwind flux; the fan terms are formed in
    // every case and discarded there (straight-line code, same bits)
    const bool sup_l = 0.0 <= SL;
    const bool sup_r = !sup_l && SR <= 0.0;
    const double inv = 1.0 / (SR - SL);
    double Uh[4];
#pragma unroll
    for (int v = 0; v < 6; ++v) {
        const double hll = (SR * FL[v] - SL * FR[v] + SL * SR * (UR[v] - UL[v])) * inv;
        const double up = sup_l ? FL[v] : FR[v];
        flux[v] = (sup_l || sup_r) ? up : hll;
        if (v < 4) Uh[v] = (SR * UR[v] - SL * UL[v] + FL[v] - FR[v]) * inv;
    }
    const double vh1 = Uh[1] / Uh[0], vh2 = Uh[2] / Uh[0], vh3 = Uh[3] / Uh[0];
    const bool sup = sup_l || sup_r;
    vface[0] = sup ? (sup_l ? uL : uR) : vh1;
    vface[1] = sup ? (sup_l ? vL : vR) : vh2;
    vface[2] = sup ? (sup_l ? wL : wR) : vh3;
}

// ---- HLLD (src/prj_riemann.c:19-440) --------------------------------------

// Error codes mirror the reference abort sites.
enum HlldError {
    HLLD_OK = 0,
    HLLD_BAD_RHO = 11,       // src/prj_riemann.c:159-161
    HLLD_BAD_P = 12,         // :162-164
    HLLD_BAD_GAMMA = 13,     // :165-167
    HLLD_BAD_DISC = 14,      // :126-128
    HLLD_BAD_CF = 15,        // :131-133
    HLLD_BAD_ORDER = 16,     // :361-363
    HLLD_BAD_CONTACT = 17,   // :381-384
    HLLD_COINCIDE = 18,      // :198-200
    HLLD_BAD_STAR_RHO = 19,  // :203-205
    HLLD_BAD_DSS = 20        // :247-249
};

struct HlldSide {
    double rho, vx, vy, vz, by, bz, p, g, ye, e, pt;
};

struct HlldStar {
    double rho, vx, vy, vz, by, bz, ye, e;
};

// conserved vector of a state (src/prj_riemann.c:81-93)
__device__ __forceinline__ void hlld_cons(double rho, double vx, double vy, double vz, double bx,
                                          double by, double bz, double ye, double e, double* U)
{
    U[0] = rho;
    U[1] = rho * vx;
    U[2] = rho * vy;
    U[3] = rho * vz;
    U[4] = e;
    U[5] = rho * ye;
    U[6] = bx;
    U[7] = by;
    U[8] = bz;
}

// physical flux (src/prj_riemann.c:95-115)
__device__ __forceinline__ void hlld_flux(const HlldSide& s, double bx, double* F)
{
    double b2 = bx * bx + s.by * s.by + s.bz * s.bz;
    double pt = s.p + 0.5 * b2;
    double vdotb = s.vx * bx + s.vy * s.by + s.vz * s.bz;
    F[0] = s.rho * s.vx;
    F[1] = s.rho * s.vx * s.vx + pt - bx * bx;
    F[2] = s.rho * s.vx * s.vy - bx * s.by;
    F[3] = s.rho * s.vx * s.vz - bx * s.bz;
    F[4] = (s.e + pt) * s.vx - bx * vdotb;
    F[5] = s.rho * s.ye * s.vx;
    F[6] = 0.0;
    F[7] = s.by * s.vx - bx * s.vy;
    F[8] = s.bz * s.vx - bx * s.vz;
}

// state from primitives (src/prj_riemann.c:142-179)
__device__ __forceinline__ int hlld_side(const double* W, double p, double g, double bn, HlldSide& s)
{
    s.rho = W[RHO];
    s.vx = W[V1];
    s.vy = W[V2];
    s.vz = W[V3];
    s.by = W[B2];
    s.bz = W[B3];
    s.p = p;
    s.g = g;
    s.ye = W[YE];
    if (s.rho <= 0.0) return HLLD_BAD_RHO;
    if (s.p <= 0.0) return HLLD_BAD_P;
    if (s.g <= 0.0) return HLLD_BAD_GAMMA;
    double kinetic = 0.5 * s.rho * (s.vx * s.vx + s.vy * s.vy + s.vz * s.vz);
    double magnetic = 0.5 * (bn * bn + s.by * s.by + s.bz * s.bz);
    s.e = s.rho * W[EINT] + kinetic + magnetic;
    s.pt = s.p + magnetic;
    return HLLD_OK;
}

// fast magnetosonic speed (src/prj_riemann.c:117-140)
__device__ __forceinline__ int hlld_fast(const HlldSide& s, double bx, double& cf)
{
    double a2 = s.g * s.p / s.rho;
    double bx2_rho = bx * bx / s.rho;
    double b2_rho = (bx * bx + s.by * s.by + s.bz * s.bz) / s.rho;
    double disc = (a2 + b2_rho) * (a2 + b2_rho) - 4.0 * a2 * bx2_rho;
    if (disc < 0.0) {
        if (disc < -1.0e-12 * (a2 + b2_rho) * (a2 + b2_rho)) return HLLD_BAD_DISC;
        disc = 0.0;
    }
    double cf2 = 0.5 * (a2 + b2_rho + sqrt(disc));
    if (cf2 < 0.0 || !isfinite(cf2)) return HLLD_BAD_CF;
    cf = sqrt(cf2);
    return HLLD_OK;
}

// outer (single-star) state (src/prj_riemann.c:187-232)
__device__ __forceinline__ int hlld_outer(const HlldSide& s, double bx, double S, double SM,
                                          double pt_star, HlldStar& st)
{
    double alpha = S - s.vx;
    double smdiff = S - SM;
    if (fabs(smdiff) <= 1.0e-14 * (fabs(S) + fabs(SM) + 1.0)) return HLLD_COINCIDE;
    st.rho = s.rho * alpha / smdiff;
    if (st.rho <= 0.0 || !isfinite(st.rho)) return HLLD_BAD_STAR_RHO;
    st.vx = SM;
    st.ye = s.ye;
    double denom = s.rho * alpha * smdiff - bx * bx;
    if (fabs(denom) < 1.0e-8 * pt_star) {
        st.vy = s.vy;
        st.vz = s.vz;
        st.by = s.by;
        st.bz = s.bz;
    } else {
        double q = (SM - s.vx) / denom;
        double mag = (s.rho * alpha * alpha - bx * bx) / denom;
        st.vy = s.vy - bx * s.by * q;
        st.vz = s.vz - bx * s.bz * q;
        st.by = s.by * mag;
        st.bz = s.bz * mag;
    }
    double vdotb = s.vx * bx + s.vy * s.by + s.vz * s.bz;
    double vsdotbs = st.vx * bx + st.vy * st.by + st.vz * st.bz;
    st.e = (alpha * s.e - s.pt * s.vx + pt_star * SM + bx * (vdotb - vsdotbs)) / smdiff;
    return HLLD_OK;
}

// HLLD flux (src/prj_riemann.c:321-440).  WL/WR are rotated face states with
// the normal field slot already set to bn.  Outputs the 9-variable flux, the
// face velocity and the face EMFs Bv1 = bz*vx - bx*vz, Bv2 = bx*vy - by*vx of
// the selected fan state (src/prj_riemann.c:305-319).
__device__ __forceinline__ int hlld(const double* WL, const double* WR, double pL, double pR,
                                 double gL, double gR, double bn, double* flux, double* vface,
                                 double* bv1, double* bv2)
{
    HlldSide L, R;
    int rc;
    if ((rc = hlld_side(WL, pL, gL, bn, L)) != 0) return rc;
    if ((rc = hlld_side(WR, pR, gR, bn, R)) != 0) return rc;
    double cfL, cfR;
    if ((rc = hlld_fast(L, bn, cfL)) != 0) return rc;
    if ((rc = hlld_fast(R, bn, cfR)) != 0) return rc;
    double SL = dmin(L.vx - cfL, R.vx - cfR);
    double SR = dmax(L.vx + cfL, R.vx + cfR);
    if (!(SL < SR)) return HLLD_BAD_ORDER;

    if (0.0 <= SL || SR <= 0.0) {
        const HlldSide& s = 0.0 <= SL ? L : R;
        hlld_flux(s, bn, flux);
        vface[0] = s.vx; vface[1] = s.vy; vface[2] = s.vz;
        *bv1 = s.bz * s.vx - bn * s.vz;
        *bv2 = bn * s.vy - s.by * s.vx;
        return HLLD_OK;
    }
    double denom = R.rho * (SR - R.vx) - L.rho * (SL - L.vx);
    if (fabs(denom) <= 1.0e-14 * (fabs(R.rho * (SR - R.vx)) + fabs(L.rho * (SL - L.vx)) + 1.0))
        return HLLD_BAD_CONTACT;
    double SM = (R.rho * R.vx * (SR - R.vx) - L.rho * L.vx * (SL - L.vx) + (L.pt - R.pt)) / denom;
    double ptl = L.pt + L.rho * (SL - L.vx) * (SM - L.vx);
    double ptr = R.pt + R.rho * (SR - R.vx) * (SM - R.vx);
    double pt_star = 0.5 * (ptl + ptr);
    if (pt_star <= 0.0 || !isfinite(pt_star)) {
        // HLL fallback (src/prj_riemann.c:393-407): both outer states needed
        double UL[9], UR[9], FL[9], FR[9];
        hlld_flux(L, bn, FL);
        hlld_flux(R, bn, FR);
        hlld_cons(L.rho, L.vx, L.vy, L.vz, bn, L.by, L.bz, L.ye, L.e, UL);
        hlld_cons(R.rho, R.vx, R.vy, R.vz, bn, R.by, R.bz, R.ye, R.e, UR);
        double inv = 1.0 / (SR - SL);
        double Uh[9];
        for (int v = 0; v < 9; ++v) {
            flux[v] = (SR * FL[v] - SL * FR[v] + SL * SR * (UR[v] - UL[v])) * inv;
            Uh[v] = (SR * UR[v] - SL * UL[v] + FL[v] - FR[v]) * inv;
        }
        double rho = Uh[0];
        double vx = Uh[1] / rho, vy = Uh[2] / rho, vz = Uh[3] / rho;
        vface[0] = vx; vface[1] = vy; vface[2] = vz;
        *bv1 = Uh[8] * vx - Uh[6] * vz;
        *bv2 = Uh[6] * vy - Uh[7] * vx;
        return HLLD_OK;
    }
    const bool bn_small = (0.5 * bn * bn < 1.0e-8 * pt_star);
    // One side is upwind of the contact; the other side's star state is only
    // needed for the double-star (rotational) states.
    const bool left = 0.0 <= SM;
    const HlldSide& A = left ? L : R;        // upwind outer side
    const double SA = left ? SL : SR;
    double FA[9], UA[9];
    hlld_flux(A, bn, FA);
    hlld_cons(A.rho, A.vx, A.vy, A.vz, bn, A.by, A.bz, A.ye, A.e, UA);
    HlldStar As;
    if ((rc = hlld_outer(A, bn, SA, SM, pt_star, As)) != 0) return rc;
    double sqA = sqrt(As.rho);
    double SAs = left ? SM - fabs(bn) / sqA : SM + fabs(bn) / sqA;
    double UAs[9];
    hlld_cons(As.rho, As.vx, As.vy, As.vz, bn, As.by, As.bz, As.ye, As.e, UAs);
    if ((left ? 0.0 <= SAs : SAs <= 0.0) || bn_small) {
        for (int v = 0; v < 9; ++v) flux[v] = FA[v] + SA * (UAs[v] - UA[v]);
        vface[0] = As.vx; vface[1] = As.vy; vface[2] = As.vz;
        *bv1 = As.bz * As.vx - bn * As.vz;
        *bv2 = bn * As.vy - As.by * As.vx;
        return HLLD_OK;
    }
    // double-star state (src/prj_riemann.c:234-282)
    const HlldSide& Bo = left ? R : L;
    const double SB = left ? SR : SL;
    HlldStar Bs;
    if ((rc = hlld_outer(Bo, bn, SB, SM, pt_star, Bs)) != 0) return rc;
    const HlldStar& Ls = left ? As : Bs;
    const HlldStar& Rs = left ? Bs : As;
    double sl = sqrt(Ls.rho);
    double sr = sqrt(Rs.rho);
    if (sl + sr <= 0.0) return HLLD_BAD_DSS;
    double inv_sum = 1.0 / (sl + sr);
    double sgn = bn < 0.0 ? -1.0 : 1.0;
    double vy = (sl * Ls.vy + sr * Rs.vy + (Rs.by - Ls.by) * sgn) * inv_sum;
    double vz = (sl * Ls.vz + sr * Rs.vz + (Rs.bz - Ls.bz) * sgn) * inv_sum;
    double by = (sl * Rs.by + sr * Ls.by + sl * sr * (Rs.vy - Ls.vy) * sgn) * inv_sum;
    double bz = (sl * Rs.bz + sr * Ls.bz + sl * sr * (Rs.vz - Ls.vz) * sgn) * inv_sum;
    double vbss = (left ? Ls.vx : Rs.vx) * bn + vy * by + vz * bz;
    double ess;
    if (left) {
        double vbl = Ls.vx * bn + Ls.vy * Ls.by + Ls.vz * Ls.bz;
        ess = Ls.e - sl * sgn * (vbl - vbss);
    } else {
        double vbr = Rs.vx * bn + Rs.vy * Rs.by + Rs.vz * Rs.bz;
        ess = Rs.e + sr * sgn * (vbr - vbss);
    }
    double Uss[9];
    hlld_cons(As.rho, As.vx, vy, vz, bn, by, bz, As.ye, ess, Uss);
    for (int v = 0; v < 9; ++v) flux[v] = FA[v] + SA * (UAs[v] - UA[v]) + SAs * (Uss[v] - UAs[v]);
    vface[0] = As.vx; vface[1] = vy; vface[2] = vz;
    *bv1 = bz * As.vx - bn * vz;
    *bv2 = bn * vy - by * As.vx;
    return HLLD_OK;
}

// Upwind CT edge EMF without the resistive term (src/prj_mhd.c:651-683).
__device__ __forceinline__ double emf_upwind(const double* ef, const double* ec, const double* vn)
{
    double emf = 0.25 * (ef[0] + ef[1] + ef[2] + ef[3]);
    if (vn[0] > 0.0)      emf += 0.25 * (ef[3] - ec[3]);
    else if (vn[0] < 0.0) emf += 0.25 * (ef[1] - ec[0]);
    else                  emf += 0.125 * (ef[3] - ec[3] + ef[1] - ec[0]);
    if (vn[1] > 0.0)      emf += 0.25 * (ef[2] - ec[1]);
    else if (vn[1] < 0.0) emf += 0.25 * (ef[0] - ec[0]);
    else                  emf += 0.125 * (ef[2] - ec[1] + ef[0] - ec[0]);
    if (vn[2] > 0.0)      emf += 0.25 * (ef[3] - ec[2]);
    else if (vn[2] < 0.0) emf += 0.25 * (ef[1] - ec[1]);
    else                  emf += 0.125 * (ef[3] - ec[2] + ef[1] - ec[1]);
    if (vn[3] > 0.0)      emf += 0.25 * (ef[2] - ec[2]);
    else if (vn[3] < 0.0) emf += 0.25 * (ef[0] - ec[3]);
    else                  emf += 0.125 * (ef[2] - ec[2] + ef[0] - ec[3]);
    return emf;
}

// Anomalous resistivity mask of one cell (src/prj_timeint.c:230-263):
// 1e12 * max(0, log(rho / 1e13)) * max(0, log(Ye / 0.25)).  A log factor is
// positive only when its argument exceeds 1 and is then taken with the
// correctly rounded log_cr (fx_log.hpp); at or below 1 every faithful log is
// <= 0 and the factor is 0, so the log is skipped there.  The reference aborts
// on a non-finite mask; none arises: each factor is at most
// log(DBL_MAX) < 710, so the mask stays below 1e12 * 710^2.  The quotients
// are formed only past the thresholds: rounding is monotone, so rho <= 1e13
// gives rho / 1e13 <= 1 (and ye / 0.25 is exact), and the factors are 0
// there without a division.
__device__ __forceinline__ double eta_mask(double rho, double ye)
{
    double rf = 0.0, yf = 0.0;
    if (rho > 1.0e13 && isfinite(rho)) {
        double a = rho / 1.0e13;
        if (a > 1.0) rf = log_cr(a);
    }
    if (ye > 0.25 && isfinite(ye)) {
        double b = ye / 0.25;
        if (b > 1.0) yf = log_cr(b);
    }
    return 1.0e12 * rf * yf;
}

}  // namespace fx
