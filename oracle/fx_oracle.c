/*
 * CPU restatement of FornaxAMR's per-block Godunov RK2 stage (hydro and MHD).
 *
 * TEST INFRASTRUCTURE ONLY -- see fx_oracle.h.  Every function cites the
 * reference file:line it follows; expressions keep the reference's operation
 * order and association so that, built without floating-point contraction,
 * results are bit-identical to a serial build of the reference.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   - whole-stage and multi-step outputs against the serial reference build
 *     (oracle/_ref, built by oracle/Makefile from the reference sources), hydro
 *     periodic/outflow/reflect and MHD outflow: bit-exact;
 *   - committed golden fixtures in tests/golden/ made from that build;
 *   - closed forms: MC slope on monotone/extremum/linear data, HLLC and HLLD
 *     consistency F(U,U) = F_phys(U), supersonic upwinding, free-stream
 *     preservation, conservation in a periodic box, div B preservation by CT.
 */
#include "fx_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define BSZ 16
#define NHMAX 9

#define FXO_MIN(a, b) (((a) < (b)) ? (a) : (b))
#define FXO_MAX(a, b) (((a) > (b)) ? (a) : (b))

/* prim: rho, v1, v2, v3, eint, ye [, b1, b2, b3]; cons: rho, m1, m2, m3, E, rho*ye [, B]
 * (src/prj_defs.h:466-494) */
enum { RHO = 0, V1 = 1, V2 = 2, V3 = 3, EINT = 4, YE = 5, B1 = 6, B2 = 7, B3 = 8 };
enum { MOM1 = 1, MOM2 = 2, MOM3 = 3, ETOT = 4 };

static char g_err[256];

static int fxo_fail(int code, const char *msg)
{
    snprintf(g_err, sizeof(g_err), "%s", msg);
    return code;
}

const char *fxo_last_error(void)
{
    return g_err;
}

static int nh(const fxo_grid *g)
{
    return g->mhd ? 9 : 6;
}

int fxo_grid_init(fxo_grid *g, int nx, int ny, int nz, const double *xlim,
    const int *bc, double gamma, int mhd, double eta)
{
    int d;
    int n[3];
    n[0] = nx;
    n[1] = ny;
    n[2] = nz;
    memset(g, 0, sizeof(*g));
    for (d = 0; d < 3; ++d) {
        double block_dx;
        if (n[d] <= 0 || n[d] % BSZ != 0) {
            return fxo_fail(1, "grid extent must be a positive multiple of 16");
        }
        g->n[d] = n[d];
        g->bc[2 * d] = bc[2 * d];
        g->bc[2 * d + 1] = bc[2 * d + 1];
        if ((bc[2 * d] == FXO_BC_PERIODIC) != (bc[2 * d + 1] == FXO_BC_PERIODIC)) {
            return fxo_fail(2, "periodic boundaries must be paired");
        }
        if (bc[2 * d] == FXO_BC_PERIODIC && n[d] < 2 * BSZ) {
            return fxo_fail(3, "a periodic axis needs at least two blocks");
        }
        if (mhd && bc[2 * d] != FXO_BC_OUTFLOW) {
            return fxo_fail(4, "MHD supports outflow boundaries only");
        }
        if (mhd && bc[2 * d + 1] != FXO_BC_OUTFLOW) {
            return fxo_fail(4, "MHD supports outflow boundaries only");
        }
        /* src/prj_mesh.c:1553-1573 */
        block_dx = (xlim[2 * d + 1] - xlim[2 * d]) / (double)(n[d] / BSZ);
        g->dx[d] = block_dx / (double)BSZ;
    }
    /* src/prj_mesh.c:675-687 */
    g->vol = g->dx[0] * g->dx[1] * g->dx[2];
    g->area[0] = g->dx[1] * g->dx[2];
    g->area[1] = g->dx[0] * g->dx[2];
    g->area[2] = g->dx[0] * g->dx[1];
    g->gamma = gamma;
    g->mhd = mhd;
    g->eta = eta;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Indexing                                                                  */

static size_t ncell(const fxo_grid *g)
{
    return (size_t)g->n[0] * g->n[1] * g->n[2];
}

static size_t cidx(const fxo_grid *g, int x, int y, int z)
{
    return ((size_t)x * g->n[1] + (size_t)y) * g->n[2] + (size_t)z;
}

static size_t fidx(const fxo_grid *g, int dir, int x, int y, int z)
{
    size_t ny = (size_t)g->n[1] + (dir == 1);
    size_t nz = (size_t)g->n[2] + (dir == 2);
    return ((size_t)x * ny + (size_t)y) * nz + (size_t)z;
}

/* Cell ghost mapping of one axis: periodic wrap through the wrapped neighbour
 * (src/prj_mesh.c:1597-1645) or the mirror copy of the physical boundary,
 * src_a = -a-1 / 2N-1-a, with the normal velocity negated for reflect
 * (src/prj_boundary.c:413-487). */
static int map_axis(const fxo_grid *g, int axis, int c, int *flip)
{
    int n = g->n[axis];
    int t;
    if (c >= 0 && c < n) {
        return c;
    }
    t = c < 0 ? g->bc[2 * axis] : g->bc[2 * axis + 1];
    if (t == FXO_BC_PERIODIC) {
        return c < 0 ? c + n : c - n;
    }
    if (t == FXO_BC_REFLECT) {
        *flip ^= 1 << axis;
    }
    return c < 0 ? -c - 1 : 2 * n - 1 - c;
}

/* Face ghost mapping for outflow magnetic fields: copy of the nearest storage
 * face inside the domain (src/prj_boundary.c:902-995). */
static double bf_at(const fxo_grid *g, const double *const *bf, int dir, int x, int y, int z)
{
    int c[3];
    int d;
    c[0] = x;
    c[1] = y;
    c[2] = z;
    for (d = 0; d < 3; ++d) {
        int hi = g->n[d] - 1 + (dir == d);
        if (c[d] < 0) {
            c[d] = 0;
        } else if (c[d] > hi) {
            c[d] = hi;
        }
    }
    return bf[dir][fidx(g, dir, c[0], c[1], c[2])];
}

/* Cell-centred B from faces (src/prj_mhd.c:555-603; sqrt_gamma = 1). */
static void bf2bc_cell(const fxo_grid *g, const double *const *bf, int x, int y, int z,
    double *b)
{
    b[0] = 0.5 * (bf_at(g, bf, 0, x, y, z) + bf_at(g, bf, 0, x + 1, y, z));
    b[1] = 0.5 * (bf_at(g, bf, 1, x, y, z) + bf_at(g, bf, 1, x, y + 1, z));
    b[2] = 0.5 * (bf_at(g, bf, 2, x, y, z) + bf_at(g, bf, 2, x, y, z + 1));
    b[0] /= 1.0;
    b[1] /= 1.0;
    b[2] /= 1.0;
}

void fxo_bf2bc(const fxo_grid *g, const double *const *bf, double *W)
{
    size_t N = ncell(g);
    int x;
#pragma omp parallel for schedule(static)
    for (x = 0; x < g->n[0]; ++x) {
        int y;
        int z;
        for (y = 0; y < g->n[1]; ++y) {
            for (z = 0; z < g->n[2]; ++z) {
                double b[3];
                size_t c = cidx(g, x, y, z);
                bf2bc_cell(g, bf, x, y, z, b);
                W[(size_t)B1 * N + c] = b[0];
                W[(size_t)B2 * N + c] = b[1];
                W[(size_t)B3 * N + c] = b[2];
            }
        }
    }
}

/* Primitive value at any cell of the ghost-padded stencil.  Hydro variables
 * come through the ghost mapping; for MHD, ghost-cell B is the face average
 * of the ghost faces, as prj_mhd_bf2bc_all_mesh fills it after the ghost
 * exchange (src/prj_boundary.c:1234-1239). */
static double cell_val(const fxo_grid *g, const double *W, const double *const *bf,
    int v, int x, int y, int z)
{
    int flip = 0;
    int mx = map_axis(g, 0, x, &flip);
    int my = map_axis(g, 1, y, &flip);
    int mz = map_axis(g, 2, z, &flip);
    int ghost = (mx != x) || (my != y) || (mz != z);
    double val;
    if (g->mhd && v >= B1 && ghost) {
        double b[3];
        bf2bc_cell(g, bf, x, y, z, b);
        return b[v - B1];
    }
    val = W[(size_t)v * ncell(g) + cidx(g, mx, my, mz)];
    if (v >= V1 && v <= V3 && (flip & (1 << (v - V1)))) {
        val = -val;
    }
    return val;
}

/* Ideal-gas eosvar refill: p = (g-1) rho e (src/prj_eos.c:240-261 with the
 * problem EOS of problems/prj_problem_ideal_eos.h:24-32). */
static double cell_pressure(const fxo_grid *g, const double *W, int x, int y, int z)
{
    int flip = 0;
    size_t c = cidx(g, map_axis(g, 0, x, &flip), map_axis(g, 1, y, &flip),
        map_axis(g, 2, z, &flip));
    size_t N = ncell(g);
    return (g->gamma - 1.0) * W[(size_t)RHO * N + c] * W[(size_t)EINT * N + c];
}

/* ------------------------------------------------------------------------ */
/* Reconstruction and Riemann solvers                                        */

/* src/prj_reconstruct.h:355-375 */
double fxo_mc_slope(double qm, double q0, double qp)
{
    double sl = q0 - qm;
    double sr = qp - q0;
    double c0 = 2.0 * sl;
    double c1 = sl + 0.5 * (sr - sl);
    double c2 = 2.0 * sr;
    double mn = c0 < c1 ? c0 : c1;
    double mx = c0 > c1 ? c0 : c1;
    double slope;
    mn = mn < c2 ? mn : c2;
    mx = mx > c2 ? mx : c2;
    slope = sl > 0.0 ? mn : mx;
    return (sl * sr > 0.0) ? slope : 0.0;
}

/* src/prj_riemann.c:8-17 */
static double theta_limiter(double cfmax, double du, double dv, double dw)
{
    double theta_r;
    theta_r = (-FXO_MIN(du, 0.0) + cfmax) / (-FXO_MIN(FXO_MIN(dv, dw), 0.0) + cfmax);
    theta_r *= theta_r;
    return FXO_MIN(1.0, theta_r * theta_r);
}

/* src/prj_riemann.c:1433-1492 (hydro build: 6 variables) */
static void hllc_state(const double *W, double pressure, double gamma, double *U,
    double *F, double *cs)
{
    double rho = W[RHO];
    double v1 = W[V1];
    double v2 = W[V2];
    double v3 = W[V3];
    double eint = W[EINT];
    double etot = rho * eint + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3);
    U[RHO] = rho;
    U[MOM1] = rho * v1;
    U[MOM2] = rho * v2;
    U[MOM3] = rho * v3;
    U[ETOT] = etot;
    U[YE] = rho * W[YE];
    *cs = sqrt(gamma * pressure / rho);
    F[RHO] = rho * v1;
    F[MOM1] = rho * v1 * v1 + pressure;
    F[MOM2] = rho * v1 * v2;
    F[MOM3] = rho * v1 * v3;
    F[ETOT] = (etot + pressure) * v1;
    F[YE] = rho * W[YE] * v1;
}

/* src/prj_riemann.c:1494-1599 */
void fxo_hllc(const double *WL, const double *WR, double pL, double pR,
    double gL, double gR, double *flux, double *v_face,
    double du, double dv, double dw)
{
    double UL[6], UR[6], FL[6], FR[6], Us[6];
    double csL, csR, SL, SR, SM, cfmax, theta, rho_star, e_over_rho;
    int v;

    hllc_state(WL, pL, gL, UL, FL, &csL);
    hllc_state(WR, pR, gR, UR, FR, &csR);
    SL = FXO_MIN(WL[V1] - csL, WR[V1] - csR);
    SR = FXO_MAX(WL[V1] + csL, WR[V1] + csR);
    cfmax = FXO_MAX(csL, csR);
    theta = theta_limiter(cfmax, du, dv, dw);
    SM = (theta * (pR - pL) + UL[MOM1] * (SL - WL[V1]) - UR[MOM1] * (SR - WR[V1])) /
        (UL[RHO] * (SL - WL[V1]) - UR[RHO] * (SR - WR[V1]));

    if (0.0 <= SL) {
        for (v = 0; v < 6; ++v) flux[v] = FL[v];
        v_face[0] = WL[V1];
        v_face[1] = WL[V2];
        v_face[2] = WL[V3];
        return;
    }
    if (SR <= 0.0) {
        for (v = 0; v < 6; ++v) flux[v] = FR[v];
        v_face[0] = WR[V1];
        v_face[1] = WR[V2];
        v_face[2] = WR[V3];
        return;
    }
    v_face[0] = SM;
    if (0.0 <= SM) {
        v_face[1] = WL[V2];
        v_face[2] = WL[V3];
        for (v = 0; v < 6; ++v) Us[v] = UL[v];
        rho_star = UL[RHO] * (SL - WL[V1]) / (SL - SM);
        Us[RHO] = rho_star;
        Us[MOM1] = rho_star * SM;
        Us[MOM2] = rho_star * WL[V2];
        Us[MOM3] = rho_star * WL[V3];
        e_over_rho = UL[ETOT] / UL[RHO] +
            (SM - WL[V1]) * (SM + pL / (UL[RHO] * (SL - WL[V1])));
        Us[ETOT] = rho_star * e_over_rho;
        Us[YE] = rho_star * WL[YE];
        for (v = 0; v < 6; ++v) flux[v] = FL[v] + SL * (Us[v] - UL[v]);
    } else {
        v_face[1] = WR[V2];
        v_face[2] = WR[V3];
        for (v = 0; v < 6; ++v) Us[v] = UR[v];
        rho_star = UR[RHO] * (SR - WR[V1]) / (SR - SM);
        Us[RHO] = rho_star;
        Us[MOM1] = rho_star * SM;
        Us[MOM2] = rho_star * WR[V2];
        Us[MOM3] = rho_star * WR[V3];
        e_over_rho = UR[ETOT] / UR[RHO] +
            (SM - WR[V1]) * (SM + pR / (UR[RHO] * (SR - WR[V1])));
        Us[ETOT] = rho_star * e_over_rho;
        Us[YE] = rho_star * WR[YE];
        for (v = 0; v < 6; ++v) flux[v] = FR[v] + SR * (Us[v] - UR[v]);
    }
}

/* ---- HLLD (src/prj_riemann.c:19-440) ---------------------------------- */

#define HLLD_SMALL 1.0e-8

typedef struct {
    double rho, vx, vy, vz, bx, by, bz, p, gamma, ye, e, pt;
    double U[9], F[9];
} hlld_state;

typedef struct {
    double rho, vx, vy, vz, bx, by, bz, ye, e;
    double U[9];
} hlld_star;

/* error codes mirror the reference abort sites */
enum {
    HLLD_OK = 0,
    HLLD_BAD_RHO = 11,       /* src/prj_riemann.c:159-161 */
    HLLD_BAD_P = 12,         /* :162-164 */
    HLLD_BAD_GAMMA = 13,     /* :165-167 */
    HLLD_BAD_DISC = 14,      /* :126-128 */
    HLLD_BAD_CF = 15,        /* :131-133 */
    HLLD_BAD_ORDER = 16,     /* :361-363 */
    HLLD_BAD_CONTACT = 17,   /* :381-384 */
    HLLD_COINCIDE = 18,      /* :198-200 */
    HLLD_BAD_STAR_RHO = 19,  /* :203-205 */
    HLLD_BAD_DSS = 20        /* :247-249 */
};

static const char *hlld_msg(int code)
{
    switch (code) {
    case HLLD_BAD_RHO: return "prj_riemann_hlld: non-positive density";
    case HLLD_BAD_P: return "prj_riemann_hlld: non-positive gas pressure";
    case HLLD_BAD_GAMMA: return "prj_riemann_hlld: non-positive adiabatic index";
    case HLLD_BAD_DISC: return "prj_riemann_hlld: negative fast-speed discriminant";
    case HLLD_BAD_CF: return "prj_riemann_hlld: invalid fast speed";
    case HLLD_BAD_ORDER: return "prj_riemann_hlld: invalid fast-wave ordering";
    case HLLD_BAD_CONTACT: return "prj_riemann_hlld: degenerate contact-speed denominator";
    case HLLD_COINCIDE: return "prj_riemann_hlld: outer wave coincides with contact wave";
    case HLLD_BAD_STAR_RHO: return "prj_riemann_hlld: invalid star density";
    case HLLD_BAD_DSS: return "prj_riemann_hlld: invalid double-star density weights";
    default: return "prj_riemann_hlld: error";
    }
}

/* src/prj_riemann.c:81-93 */
static void hlld_fill_cons(double rho, double vx, double vy, double vz, double bx,
    double by, double bz, double ye, double e, double *U)
{
    U[RHO] = rho;
    U[MOM1] = rho * vx;
    U[MOM2] = rho * vy;
    U[MOM3] = rho * vz;
    U[ETOT] = e;
    U[YE] = rho * ye;
    U[B1] = bx;
    U[B2] = by;
    U[B3] = bz;
}

/* src/prj_riemann.c:95-115 */
static void hlld_fill_flux(double rho, double vx, double vy, double vz, double bx,
    double by, double bz, double p, double ye, double e, double *F)
{
    double b2 = bx * bx + by * by + bz * bz;
    double pt = p + 0.5 * b2;
    double vdotb = vx * bx + vy * by + vz * bz;
    F[RHO] = rho * vx;
    F[MOM1] = rho * vx * vx + pt - bx * bx;
    F[MOM2] = rho * vx * vy - bx * by;
    F[MOM3] = rho * vx * vz - bx * bz;
    F[ETOT] = (e + pt) * vx - bx * vdotb;
    F[YE] = rho * ye * vx;
    F[B1] = 0.0;
    F[B2] = by * vx - bx * vy;
    F[B3] = bz * vx - bx * vz;
}

/* src/prj_riemann.c:117-140 */
static int hlld_fast_speed(const hlld_state *s, double *cf)
{
    double a2 = s->gamma * s->p / s->rho;
    double bx2_rho = s->bx * s->bx / s->rho;
    double b2_rho = (s->bx * s->bx + s->by * s->by + s->bz * s->bz) / s->rho;
    double disc = (a2 + b2_rho) * (a2 + b2_rho) - 4.0 * a2 * bx2_rho;
    double cf2;
    if (disc < 0.0) {
        if (disc < -1.0e-12 * (a2 + b2_rho) * (a2 + b2_rho)) {
            return HLLD_BAD_DISC;
        }
        disc = 0.0;
    }
    cf2 = 0.5 * (a2 + b2_rho + sqrt(disc));
    if (cf2 < 0.0 || !isfinite(cf2)) {
        return HLLD_BAD_CF;
    }
    *cf = sqrt(cf2);
    return HLLD_OK;
}

/* src/prj_riemann.c:142-179 */
static int hlld_state_from_prim(const double *W, double p, double gamma, double bn,
    hlld_state *s)
{
    double kinetic;
    double magnetic;
    s->rho = W[RHO];
    s->vx = W[V1];
    s->vy = W[V2];
    s->vz = W[V3];
    s->bx = bn;
    s->by = W[B2];
    s->bz = W[B3];
    s->p = p;
    s->gamma = gamma;
    s->ye = W[YE];
    if (s->rho <= 0.0) return HLLD_BAD_RHO;
    if (s->p <= 0.0) return HLLD_BAD_P;
    if (s->gamma <= 0.0) return HLLD_BAD_GAMMA;
    kinetic = 0.5 * s->rho * (s->vx * s->vx + s->vy * s->vy + s->vz * s->vz);
    magnetic = 0.5 * (s->bx * s->bx + s->by * s->by + s->bz * s->bz);
    s->e = s->rho * W[EINT] + kinetic + magnetic;
    s->pt = s->p + magnetic;
    hlld_fill_cons(s->rho, s->vx, s->vy, s->vz, s->bx, s->by, s->bz, s->ye, s->e, s->U);
    hlld_fill_flux(s->rho, s->vx, s->vy, s->vz, s->bx, s->by, s->bz, s->p, s->ye, s->e, s->F);
    return HLLD_OK;
}

/* src/prj_riemann.c:187-232 */
static int hlld_outer_star(const hlld_state *s, double S, double SM, double pt_star,
    hlld_star *st)
{
    double alpha = S - s->vx;
    double smdiff = S - SM;
    double denom;
    double vdotb;
    double vsdotbs;
    if (fabs(smdiff) <= 1.0e-14 * (fabs(S) + fabs(SM) + 1.0)) {
        return HLLD_COINCIDE;
    }
    st->rho = s->rho * alpha / smdiff;
    if (st->rho <= 0.0 || !isfinite(st->rho)) {
        return HLLD_BAD_STAR_RHO;
    }
    st->vx = SM;
    st->bx = s->bx;
    st->ye = s->ye;
    denom = s->rho * alpha * smdiff - s->bx * s->bx;
    if (fabs(denom) < HLLD_SMALL * pt_star) {
        st->vy = s->vy;
        st->vz = s->vz;
        st->by = s->by;
        st->bz = s->bz;
    } else {
        double sm_diff_vx_over_denom = (SM - s->vx) / denom;
        double mag_factor = (s->rho * alpha * alpha - s->bx * s->bx) / denom;
        st->vy = s->vy - s->bx * s->by * sm_diff_vx_over_denom;
        st->vz = s->vz - s->bx * s->bz * sm_diff_vx_over_denom;
        st->by = s->by * mag_factor;
        st->bz = s->bz * mag_factor;
    }
    vdotb = s->vx * s->bx + s->vy * s->by + s->vz * s->bz;
    vsdotbs = st->vx * st->bx + st->vy * st->by + st->vz * st->bz;
    st->e = (alpha * s->e - s->pt * s->vx + pt_star * SM + s->bx * (vdotb - vsdotbs)) / smdiff;
    hlld_fill_cons(st->rho, st->vx, st->vy, st->vz, st->bx, st->by, st->bz, st->ye, st->e, st->U);
    return HLLD_OK;
}

/* src/prj_riemann.c:234-282 */
static int hlld_double_star(const hlld_star *l, const hlld_star *r, double bx,
    int right_side, hlld_star *ss)
{
    double sl = sqrt(l->rho);
    double sr = sqrt(r->rho);
    double inv_sum;
    double sgn;
    double vy, vz, by, bz, vbss;
    if (sl + sr <= 0.0) {
        return HLLD_BAD_DSS;
    }
    inv_sum = 1.0 / (sl + sr);
    sgn = bx < 0.0 ? -1.0 : 1.0;
    vy = (sl * l->vy + sr * r->vy + (r->by - l->by) * sgn) * inv_sum;
    vz = (sl * l->vz + sr * r->vz + (r->bz - l->bz) * sgn) * inv_sum;
    by = (sl * r->by + sr * l->by + sl * sr * (r->vy - l->vy) * sgn) * inv_sum;
    bz = (sl * r->bz + sr * l->bz + sl * sr * (r->vz - l->vz) * sgn) * inv_sum;
    vbss = (right_side ? r->vx : l->vx) * bx + vy * by + vz * bz;
    if (!right_side) {
        double vbl = l->vx * bx + l->vy * l->by + l->vz * l->bz;
        *ss = *l;
        ss->e = l->e - sl * sgn * (vbl - vbss);
    } else {
        double vbr = r->vx * bx + r->vy * r->by + r->vz * r->bz;
        *ss = *r;
        ss->e = r->e + sr * sgn * (vbr - vbss);
    }
    ss->vy = vy;
    ss->vz = vz;
    ss->by = by;
    ss->bz = bz;
    hlld_fill_cons(ss->rho, ss->vx, ss->vy, ss->vz, ss->bx, ss->by, ss->bz, ss->ye, ss->e, ss->U);
    return HLLD_OK;
}

/* src/prj_riemann.c:305-319 */
static void hlld_outputs(double vx, double vy, double vz, double bx, double by, double bz,
    double *v_face, double *Bv1, double *Bv2)
{
    v_face[0] = vx;
    v_face[1] = vy;
    v_face[2] = vz;
    *Bv1 = bz * vx - bx * vz;
    *Bv2 = bx * vy - by * vx;
}

/* src/prj_riemann.c:321-440 */
int fxo_hlld(const double *WL, const double *WR, double pL, double pR,
    double gL, double gR, double bn, double *flux, double *v_face,
    double *Bv1, double *Bv2)
{
    hlld_state L, R;
    hlld_star Ls, Rs, ss;
    double cfL, cfR, SL, SR, SM, SLs, SRs, denom, ptl, ptr, pt_star;
    int v;
    int rc;
    int bn_small;

    if ((rc = hlld_state_from_prim(WL, pL, gL, bn, &L)) != 0) return rc;
    if ((rc = hlld_state_from_prim(WR, pR, gR, bn, &R)) != 0) return rc;
    if ((rc = hlld_fast_speed(&L, &cfL)) != 0) return rc;
    if ((rc = hlld_fast_speed(&R, &cfR)) != 0) return rc;
    SL = FXO_MIN(L.vx - cfL, R.vx - cfR);
    SR = FXO_MAX(L.vx + cfL, R.vx + cfR);
    if (!(SL < SR)) return HLLD_BAD_ORDER;

    if (0.0 <= SL) {
        for (v = 0; v < 9; ++v) flux[v] = L.F[v];
        hlld_outputs(L.vx, L.vy, L.vz, L.bx, L.by, L.bz, v_face, Bv1, Bv2);
        return HLLD_OK;
    }
    if (SR <= 0.0) {
        for (v = 0; v < 9; ++v) flux[v] = R.F[v];
        hlld_outputs(R.vx, R.vy, R.vz, R.bx, R.by, R.bz, v_face, Bv1, Bv2);
        return HLLD_OK;
    }
    denom = R.rho * (SR - R.vx) - L.rho * (SL - L.vx);
    if (fabs(denom) <= 1.0e-14 * (fabs(R.rho * (SR - R.vx)) + fabs(L.rho * (SL - L.vx)) + 1.0)) {
        return HLLD_BAD_CONTACT;
    }
    SM = (R.rho * R.vx * (SR - R.vx) - L.rho * L.vx * (SL - L.vx) + (L.pt - R.pt)) / denom;
    ptl = L.pt + L.rho * (SL - L.vx) * (SM - L.vx);
    ptr = R.pt + R.rho * (SR - R.vx) * (SM - R.vx);
    pt_star = 0.5 * (ptl + ptr);
    if (pt_star <= 0.0 || !isfinite(pt_star)) {
        double inv = 1.0 / (SR - SL);
        double Uhll[9];
        double rho;
        for (v = 0; v < 9; ++v) {
            flux[v] = (SR * L.F[v] - SL * R.F[v] + SL * SR * (R.U[v] - L.U[v])) * inv;
            Uhll[v] = (SR * R.U[v] - SL * L.U[v] + L.F[v] - R.F[v]) * inv;
        }
        rho = Uhll[RHO];
        hlld_outputs(Uhll[MOM1] / rho, Uhll[MOM2] / rho, Uhll[MOM3] / rho,
            Uhll[B1], Uhll[B2], Uhll[B3], v_face, Bv1, Bv2);
        return HLLD_OK;
    }
    bn_small = (0.5 * bn * bn < HLLD_SMALL * pt_star);
    if (0.0 <= SM) {
        if ((rc = hlld_outer_star(&L, SL, SM, pt_star, &Ls)) != 0) return rc;
        SLs = SM - fabs(bn) / sqrt(Ls.rho);
        if (0.0 <= SLs || bn_small) {
            for (v = 0; v < 9; ++v) flux[v] = L.F[v] + SL * (Ls.U[v] - L.U[v]);
            hlld_outputs(Ls.vx, Ls.vy, Ls.vz, Ls.bx, Ls.by, Ls.bz, v_face, Bv1, Bv2);
        } else {
            if ((rc = hlld_outer_star(&R, SR, SM, pt_star, &Rs)) != 0) return rc;
            if ((rc = hlld_double_star(&Ls, &Rs, bn, 0, &ss)) != 0) return rc;
            for (v = 0; v < 9; ++v) {
                flux[v] = L.F[v] + SL * (Ls.U[v] - L.U[v]) + SLs * (ss.U[v] - Ls.U[v]);
            }
            hlld_outputs(ss.vx, ss.vy, ss.vz, ss.bx, ss.by, ss.bz, v_face, Bv1, Bv2);
        }
    } else {
        if ((rc = hlld_outer_star(&R, SR, SM, pt_star, &Rs)) != 0) return rc;
        SRs = SM + fabs(bn) / sqrt(Rs.rho);
        if (SRs <= 0.0 || bn_small) {
            for (v = 0; v < 9; ++v) flux[v] = R.F[v] + SR * (Rs.U[v] - R.U[v]);
            hlld_outputs(Rs.vx, Rs.vy, Rs.vz, Rs.bx, Rs.by, Rs.bz, v_face, Bv1, Bv2);
        } else {
            if ((rc = hlld_outer_star(&L, SL, SM, pt_star, &Ls)) != 0) return rc;
            if ((rc = hlld_double_star(&Ls, &Rs, bn, 1, &ss)) != 0) return rc;
            for (v = 0; v < 9; ++v) {
                flux[v] = R.F[v] + SR * (Rs.U[v] - R.U[v]) + SRs * (ss.U[v] - Rs.U[v]);
            }
            hlld_outputs(ss.vx, ss.vy, ss.vz, ss.bx, ss.by, ss.bz, v_face, Bv1, Bv2);
        }
    }
    return HLLD_OK;
}

/* ------------------------------------------------------------------------ */
/* Per-face flux assembly (src/prj_flux.c:2149-2521)                          */

/* local->global primitive permutation (src/prj_flux.c:274-349) */
static const int GMAP[3][9] = {
    { 0, 1, 2, 3, 4, 5, 6, 7, 8 },
    { 0, 2, 3, 1, 4, 5, 7, 8, 6 },
    { 0, 3, 1, 2, 4, 5, 8, 6, 7 },
};

/* unit offsets */
static void unit(int axis, int *e)
{
    e[0] = axis == 0;
    e[1] = axis == 1;
    e[2] = axis == 2;
}

/* Face (x,y,z) of direction dir lies between cell c - e_dir and c. */
typedef struct {
    double F[9];      /* global-order flux */
    double vn;        /* normal face velocity (v_riemann normal component) */
    double vface[3];  /* global-order face velocity */
    double bv1, bv2;  /* face EMFs (MHD) */
} face_out;

/* MC face states of one variable: left state from cell c-1, right from cell c
 * (src/prj_flux.c:373-445: wr = q0 - 0.5*slope, wl = q0 + 0.5*slope). */
static void mc_pair(const double *q, double *wl, double *wr)
{
    /* q[0..3] = cells c-2, c-1, c, c+1 */
    double hl = 0.5 * fxo_mc_slope(q[0], q[1], q[2]);
    double hr = 0.5 * fxo_mc_slope(q[1], q[2], q[3]);
    *wl = q[1] + hl;
    *wr = q[2] - hr;
}

/* HLLE: the HLL flux (Harten, Lax & van Leer 1983) over the Davis wave-speed
 * bounds the reference's HLLC takes (src/prj_riemann.c:1518-1519), on its
 * face states (hllc_state above), written in the form of the reference's HLL
 * fallback inside HLLD (src/prj_riemann.c:393-407): flux
 * (SR FL - SL FR + SL SR (UR - UL)) / (SR - SL) and the face velocity of the
 * HLL average state (SR UR - SL UL + FL - FR) / (SR - SL).  The reference has
 * no hydro HLLE: this function is pinned by closed forms (the jump conditions
 * across the two outer waves, consistency, supersonic upwinding, mirror
 * symmetry) and against the differentiable scheme's independent HLLE. */
void fxo_hlle(const double *WL, const double *WR, double pL, double pR,
    double gL, double gR, double *flux, double *v_face)
{
    double UL[6], UR[6], FL[6], FR[6], Uh[6];
    double csL, csR, SL, SR, inv;
    int v;

    hllc_state(WL, pL, gL, UL, FL, &csL);
    hllc_state(WR, pR, gR, UR, FR, &csR);
    SL = FXO_MIN(WL[V1] - csL, WR[V1] - csR);
    SR = FXO_MAX(WL[V1] + csL, WR[V1] + csR);
    if (0.0 <= SL) {
        for (v = 0; v < 6; ++v) flux[v] = FL[v];
        v_face[0] = WL[V1];
        v_face[1] = WL[V2];
        v_face[2] = WL[V3];
        return;
    }
    if (SR <= 0.0) {
        for (v = 0; v < 6; ++v) flux[v] = FR[v];
        v_face[0] = WR[V1];
        v_face[1] = WR[V2];
        v_face[2] = WR[V3];
        return;
    }
    inv = 1.0 / (SR - SL);
    for (v = 0; v < 6; ++v) {
        flux[v] = (SR * FL[v] - SL * FR[v] + SL * SR * (UR[v] - UL[v])) * inv;
        Uh[v] = (SR * UR[v] - SL * UL[v] + FL[v] - FR[v]) * inv;
    }
    v_face[0] = Uh[MOM1] / Uh[RHO];
    v_face[1] = Uh[MOM2] / Uh[RHO];
    v_face[2] = Uh[MOM3] / Uh[RHO];
}

/* Riemann solver of hydro stages: 0 the reference's HLLC, 1 HLLE */
static int g_riemann = 0;

void fxo_set_riemann(int solver)
{
    g_riemann = solver;
}

static int face_flux(const fxo_grid *g, const double *W, const double *const *bf,
    int dir, int x, int y, int z, face_out *out)
{
    int e[3];
    int lv;
    int s;
    double WL[NHMAX], WR[NHMAX], Fl[NHMAX], vloc[3];
    double pL, pR, gL, gR;
    double q[4];
    int n = nh(g);

    unit(dir, e);
    for (lv = 0; lv < n; ++lv) {
        int gv = GMAP[dir][lv];
        for (s = 0; s < 4; ++s) {
            int o = s - 2;
            q[s] = cell_val(g, W, bf, gv, x + o * e[0], y + o * e[1], z + o * e[2]);
        }
        mc_pair(q, &WL[lv], &WR[lv]);
    }
    for (s = 0; s < 4; ++s) {
        int o = s - 2;
        q[s] = cell_pressure(g, W, x + o * e[0], y + o * e[1], z + o * e[2]);
    }
    mc_pair(q, &pL, &pR);
    for (s = 0; s < 4; ++s) q[s] = g->gamma;
    mc_pair(q, &gL, &gR);

    if (!g->mhd) {
        /* src/prj_flux.c:1967-2036 */
        int a1 = (dir + 1) % 3;
        int a2 = (dir + 2) % 3;
        int t1[3], t2[3];
        int lx = x - e[0], ly = y - e[1], lz = z - e[2];
        double du, dv, dw;
        unit(a1, t1);
        unit(a2, t2);
#define CV(v, X, Y, Z) cell_val(g, W, bf, (v), (X), (Y), (Z))
        du = CV(V1 + dir, x, y, z) - CV(V1 + dir, lx, ly, lz);
        dv = FXO_MIN(
            FXO_MIN(CV(V1 + a1, lx, ly, lz) - CV(V1 + a1, lx - t1[0], ly - t1[1], lz - t1[2]),
                    CV(V1 + a1, lx + t1[0], ly + t1[1], lz + t1[2]) - CV(V1 + a1, lx, ly, lz)),
            FXO_MIN(CV(V1 + a1, x, y, z) - CV(V1 + a1, x - t1[0], y - t1[1], z - t1[2]),
                    CV(V1 + a1, x + t1[0], y + t1[1], z + t1[2]) - CV(V1 + a1, x, y, z)));
        dw = FXO_MIN(
            FXO_MIN(CV(V1 + a2, lx, ly, lz) - CV(V1 + a2, lx - t2[0], ly - t2[1], lz - t2[2]),
                    CV(V1 + a2, lx + t2[0], ly + t2[1], lz + t2[2]) - CV(V1 + a2, lx, ly, lz)),
            FXO_MIN(CV(V1 + a2, x, y, z) - CV(V1 + a2, x - t2[0], y - t2[1], z - t2[2]),
                    CV(V1 + a2, x + t2[0], y + t2[1], z + t2[2]) - CV(V1 + a2, x, y, z)));
#undef CV
        if (g_riemann == 1) {
            fxo_hlle(WL, WR, pL, pR, gL, gR, Fl, vloc);
        } else {
            fxo_hllc(WL, WR, pL, pR, gL, gR, Fl, vloc, du, dv, dw);
        }
        out->bv1 = 0.0;
        out->bv2 = 0.0;
    } else {
        int rc;
        double bn = bf_at(g, bf, dir, x, y, z);
        WL[B1] = bn;
        WR[B1] = bn;
        rc = fxo_hlld(WL, WR, pL, pR, gL, gR, bn, Fl, vloc, &out->bv1, &out->bv2);
        if (rc != 0) {
            return rc;
        }
    }
    for (lv = 0; lv < n; ++lv) {
        out->F[GMAP[dir][lv]] = Fl[lv];
    }
    /* src/prj_flux.c:585-607 */
    out->vface[GMAP[dir][1] - 1] = vloc[0];
    out->vface[GMAP[dir][2] - 1] = vloc[1];
    out->vface[GMAP[dir][3] - 1] = vloc[2];
    out->vn = vloc[0];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* CT edge EMF (src/prj_timeint.c:133-413, src/prj_mhd.c:627-718)            */

/* src/prj_mhd.c:627-718 (eta term handled by the caller) */
int fxo_emf_upwind(const double *emf_face, const double *emf_cell,
    const double *v_norm, double *result)
{
    double emf;
    int n;
    for (n = 0; n < 4; ++n) {
        if (!isfinite(emf_face[n]) || !isfinite(emf_cell[n]) || !isfinite(v_norm[n])) {
            return 31;
        }
    }
    emf = 0.25 * (emf_face[0] + emf_face[1] + emf_face[2] + emf_face[3]);
    if (v_norm[0] > 0.0) {
        emf += 0.25 * (emf_face[3] - emf_cell[3]);
    } else if (v_norm[0] < 0.0) {
        emf += 0.25 * (emf_face[1] - emf_cell[0]);
    } else {
        emf += 0.125 * (emf_face[3] - emf_cell[3] + emf_face[1] - emf_cell[0]);
    }
    if (v_norm[1] > 0.0) {
        emf += 0.25 * (emf_face[2] - emf_cell[1]);
    } else if (v_norm[1] < 0.0) {
        emf += 0.25 * (emf_face[0] - emf_cell[0]);
    } else {
        emf += 0.125 * (emf_face[2] - emf_cell[1] + emf_face[0] - emf_cell[0]);
    }
    if (v_norm[2] > 0.0) {
        emf += 0.25 * (emf_face[3] - emf_cell[2]);
    } else if (v_norm[2] < 0.0) {
        emf += 0.25 * (emf_face[1] - emf_cell[1]);
    } else {
        emf += 0.125 * (emf_face[3] - emf_cell[2] + emf_face[1] - emf_cell[1]);
    }
    if (v_norm[3] > 0.0) {
        emf += 0.25 * (emf_face[2] - emf_cell[2]);
    } else if (v_norm[3] < 0.0) {
        emf += 0.25 * (emf_face[0] - emf_cell[3]);
    } else {
        emf += 0.125 * (emf_face[2] - emf_cell[2] + emf_face[0] - emf_cell[3]);
    }
    *result = emf;
    return 0;
}

/* src/prj_timeint.c:133-177: cell EMF component `dir` of E = -v x B */
static double cell_emf(const fxo_grid *g, const double *W, const double *const *bf,
    int dir, int x, int y, int z, int *rc)
{
    double b1 = cell_val(g, W, bf, B1, x, y, z);
    double b2 = cell_val(g, W, bf, B2, x, y, z);
    double b3 = cell_val(g, W, bf, B3, x, y, z);
    double v1 = cell_val(g, W, bf, V1, x, y, z);
    double v2 = cell_val(g, W, bf, V2, x, y, z);
    double v3 = cell_val(g, W, bf, V3, x, y, z);
    double e[3];
    e[0] = b2 * v3 - b3 * v2;
    e[1] = b3 * v1 - b1 * v3;
    e[2] = b1 * v2 - b2 * v1;
    if (!isfinite(e[0]) || !isfinite(e[1]) || !isfinite(e[2])) {
        *rc = 32;
    }
    return e[dir];
}

/* prj_mhd_emf_upwind with the resistive term on one block's face arrays
 * (src/prj_mhd.c:627-718): edge (dir, i, j, k) in the reference's block
 * indices, bf[3] the block's face fields (FACE_IDX, 20^3 block with two ghost
 * layers), dx the block's cell widths.  Returns 0, 37 (edge index outside
 * [-2, 17], src/prj_mhd.c:636-640), 31 (non-finite input), 35 (a face read
 * outside the block's face storage, where the reference reads out of bounds)
 * or 36 (non-finite curl B or result). */
#define FXO_BS 20
#define FXO_FACE_IDX(d, i, j, k) \
    ((d) == 0 ? (((i) + 2) * FXO_BS * FXO_BS + ((j) + 2) * FXO_BS + ((k) + 2)) : \
     (d) == 1 ? (((i) + 2) * (FXO_BS + 1) * FXO_BS + ((j) + 2) * FXO_BS + ((k) + 2)) : \
                (((i) + 2) * FXO_BS * (FXO_BS + 1) + ((j) + 2) * (FXO_BS + 1) + ((k) + 2)))

static int face_in_block(int d, int i, int j, int k)
{
    int c[3], a;
    c[0] = i;
    c[1] = j;
    c[2] = k;
    for (a = 0; a < 3; ++a) {
        if (c[a] < -2 || c[a] > (a == d ? 18 : 17)) return 0;
    }
    return 1;
}

int fxo_emf_upwind_block(int dir, int i, int j, int k, const double *emf_face,
    const double *emf_cell, const double *v_norm, const double *const *bf, const double *dx,
    double eta, double *result)
{
    double emf, curl_b;
    int rc;
    if (dir < 0 || dir > 2 || i < -2 || i > 17 || j < -2 || j > 17 || k < -2 || k > 17) return 37;
    rc = fxo_emf_upwind(emf_face, emf_cell, v_norm, &emf);
    if (rc != 0) return rc;
    if (eta != 0.0) {
        int fa, fb;   /* the two face components the curl differences */
        int ia[3], ib[3], sa[3] = {0, 0, 0}, sb[3] = {0, 0, 0};
        double wa, wb;
        if (dir == 0) {
            fa = 2; fb = 1; sa[1] = 1; sb[2] = 1; wa = dx[1]; wb = dx[2];
        } else if (dir == 1) {
            fa = 0; fb = 2; sa[2] = 1; sb[0] = 1; wa = dx[2]; wb = dx[0];
        } else {
            fa = 1; fb = 0; sa[0] = 1; sb[1] = 1; wa = dx[0]; wb = dx[1];
        }
        ia[0] = i - sa[0]; ia[1] = j - sa[1]; ia[2] = k - sa[2];
        ib[0] = i - sb[0]; ib[1] = j - sb[1]; ib[2] = k - sb[2];
        if (!face_in_block(fa, i, j, k) || !face_in_block(fa, ia[0], ia[1], ia[2]) ||
            !face_in_block(fb, i, j, k) || !face_in_block(fb, ib[0], ib[1], ib[2])) {
            return 35;
        }
        curl_b = (bf[fa][FXO_FACE_IDX(fa, i, j, k)] - bf[fa][FXO_FACE_IDX(fa, ia[0], ia[1], ia[2])]) / wa -
                 (bf[fb][FXO_FACE_IDX(fb, i, j, k)] - bf[fb][FXO_FACE_IDX(fb, ib[0], ib[1], ib[2])]) / wb;
        if (!isfinite(curl_b)) return 36;
        emf += eta * curl_b;
    }
    if (!isfinite(emf)) return 36;
    *result = emf;
    return 0;
}

/* log of the resistivity mask: the host libm's, as the reference build
 * takes it (default), or correctly rounded (fxo_logcr.c) */
double fxo_log_cr(double x);
static int g_log_cr = 0;

void fxo_set_log_mode(int correctly_rounded)
{
    g_log_cr = correctly_rounded != 0;
}

static double mask_log(double a)
{
    if (!g_log_cr) return log(a);
    /* correctly rounded log is <= 0 exactly where a <= 1, and max(0, .)
     * maps that to 0 as it does for the libm value */
    return a > 1.0 ? fxo_log_cr(a) : 0.0;
}

/* src/prj_timeint.c:230-263 */
static double eta_mask(const fxo_grid *g, const double *W, const double *const *bf,
    int x, int y, int z)
{
    double rho = cell_val(g, W, bf, RHO, x, y, z);
    double ye = cell_val(g, W, bf, YE, x, y, z);
    double rf = 0.0;
    double yf = 0.0;
    if (isfinite(rho) && rho > 0.0) {
        rf = FXO_MAX(0.0, mask_log(rho / 1.0e13));
    }
    if (isfinite(ye) && ye > 0.0) {
        yf = FXO_MAX(0.0, mask_log(ye / 0.25));
    }
    return 1.0e12 * rf * yf;
}

/* ------------------------------------------------------------------------ */
/* Stage                                                                     */

/* extended face grid: MHD carries one transverse ghost layer */
typedef struct {
    int lo[3];
    int cnt[3];
} fbox;

static fbox face_box(const fxo_grid *g, int dir)
{
    fbox b;
    int d;
    for (d = 0; d < 3; ++d) {
        if (d == dir) {
            b.lo[d] = 0;
            b.cnt[d] = g->n[d] + 1;
        } else if (g->mhd) {
            b.lo[d] = -1;
            b.cnt[d] = g->n[d] + 2;
        } else {
            b.lo[d] = 0;
            b.cnt[d] = g->n[d];
        }
    }
    return b;
}

static size_t fbox_idx(const fbox *b, int x, int y, int z)
{
    return ((size_t)(x - b->lo[0]) * b->cnt[1] + (size_t)(y - b->lo[1])) * b->cnt[2] +
        (size_t)(z - b->lo[2]);
}

static size_t fbox_size(const fbox *b)
{
    return (size_t)b->cnt[0] * b->cnt[1] * b->cnt[2];
}

/* src/prj_eos.c:2112-2158 */
static void prim2cons(int n, const double *w, double *u)
{
    double rho = w[RHO], v1 = w[V1], v2 = w[V2], v3 = w[V3], eint = w[EINT];
    u[RHO] = rho;
    u[MOM1] = rho * v1;
    u[MOM2] = rho * v2;
    u[MOM3] = rho * v3;
    if (n == 9) {
        double b1 = w[B1], b2 = w[B2], b3 = w[B3];
        u[ETOT] = rho * eint + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3) +
            0.5 * (b1 * b1 + b2 * b2 + b3 * b3);
        u[B1] = b1;
        u[B2] = b2;
        u[B3] = b3;
    } else {
        u[ETOT] = rho * eint + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3);
    }
    u[YE] = rho * w[YE];
}

/* src/prj_eos.c:2160-2220 */
static void cons2prim(int n, const double *u, double *w)
{
    double rho = u[RHO];
    double v1, v2, v3, kinetic;
    int v;
    if (rho == 0.0) {
        for (v = 0; v < n; ++v) w[v] = 0.0;
        return;
    }
    v1 = u[MOM1] / rho;
    v2 = u[MOM2] / rho;
    v3 = u[MOM3] / rho;
    kinetic = 0.5 * (v1 * v1 + v2 * v2 + v3 * v3);
    w[RHO] = rho;
    w[V1] = v1;
    w[V2] = v2;
    w[V3] = v3;
    if (n == 9) {
        double magnetic = 0.5 * (u[B1] * u[B1] + u[B2] * u[B2] + u[B3] * u[B3]) / rho;
        w[EINT] = u[ETOT] / rho - kinetic - magnetic;
        w[B1] = u[B1];
        w[B2] = u[B2];
        w[B3] = u[B3];
    } else {
        w[EINT] = u[ETOT] / rho - kinetic;
    }
    w[YE] = u[YE] / rho;
}

int fxo_stage(const fxo_grid *g, int stage, double dt,
    const double *W0, const double *W1, double *Wout,
    const double *const *bf0, const double *const *bf1, double *const *bfout,
    double *const *emf_out)
{
    const double *Win = stage == 1 ? W0 : W1;
    const double *const *bfin = stage == 1 ? bf0 : bf1;
    int n = nh(g);
    size_t N = ncell(g);
    double *F[3] = { 0, 0, 0 };
    double *BV1[3] = { 0, 0, 0 };
    double *BV2[3] = { 0, 0, 0 };
    double *VN[3] = { 0, 0, 0 };
    double *EMF[3] = { 0, 0, 0 };
    double *bfnew[3] = { 0, 0, 0 };
    fbox fb[3];
    int d;
    int rc = 0;

    if (stage != 1 && stage != 2) {
        return fxo_fail(40, "stage must be 1 or 2");
    }
    for (d = 0; d < 3; ++d) {
        fb[d] = face_box(g, d);
        F[d] = (double *)malloc(fbox_size(&fb[d]) * n * sizeof(double));
        if (g->mhd) {
            BV1[d] = (double *)malloc(fbox_size(&fb[d]) * sizeof(double));
            BV2[d] = (double *)malloc(fbox_size(&fb[d]) * sizeof(double));
            VN[d] = (double *)malloc(fbox_size(&fb[d]) * sizeof(double));
        }
    }

    /* 1. fluxes over all faces (src/prj_flux.c:2224-2521) */
    for (d = 0; d < 3 && rc == 0; ++d) {
        const fbox *b = &fb[d];
        size_t fs = fbox_size(b);
        int xi;
#pragma omp parallel for schedule(static)
        for (xi = 0; xi < b->cnt[0]; ++xi) {
            int yi, zi, v;
            int x = b->lo[0] + xi;
            for (yi = 0; yi < b->cnt[1]; ++yi) {
                for (zi = 0; zi < b->cnt[2]; ++zi) {
                    int y = b->lo[1] + yi;
                    int z = b->lo[2] + zi;
                    face_out fo;
                    size_t f = fbox_idx(b, x, y, z);
                    int frc = face_flux(g, Win, bfin, d, x, y, z, &fo);
                    if (frc != 0) {
#pragma omp critical
                        {
                            if (rc == 0) {
                                rc = frc;
                                snprintf(g_err, sizeof(g_err), "%s (dir %d face %d %d %d)",
                                    hlld_msg(frc), d, x, y, z);
                            }
                        }
                        continue;
                    }
                    for (v = 0; v < n; ++v) F[d][(size_t)v * fs + f] = fo.F[v];
                    if (g->mhd) {
                        BV1[d][f] = fo.bv1;
                        BV2[d][f] = fo.bv2;
                        VN[d][f] = fo.vn;
                    }
                }
            }
        }
    }

    /* 2. MHD: edge EMFs and CT face update */
    if (rc == 0 && g->mhd) {
        for (d = 0; d < 3 && rc == 0; ++d) {
            size_t ecount = (size_t)(g->n[0] + (d != 0)) * (g->n[1] + (d != 1)) * (g->n[2] + (d != 2));
            EMF[d] = emf_out != 0 && emf_out[d] != 0 ? emf_out[d] :
                (double *)malloc(ecount * sizeof(double));
        }
        for (d = 0; d < 3 && rc == 0; ++d) {
            int up = (d + 1) % 3;
            int right = (d + 2) % 3;
            int en[3];
            int ex;
            int eu[3], er[3];
            unit(up, eu);
            unit(right, er);
            en[0] = g->n[0] + (d != 0);
            en[1] = g->n[1] + (d != 1);
            en[2] = g->n[2] + (d != 2);
#pragma omp parallel for schedule(static)
            for (ex = 0; ex < en[0]; ++ex) {
                int ey, ez;
                for (ey = 0; ey < en[1]; ++ey) {
                    for (ez = 0; ez < en[2]; ++ez) {
                        int e[3], c0[3], c1[3], c2[3], c3[3], f2[3], f3[3];
                        double ef[4], ec[4], vn[4], emf, eta;
                        int k, crc = 0, urc;
                        e[0] = ex;
                        e[1] = ey;
                        e[2] = ez;
                        for (k = 0; k < 3; ++k) {
                            c2[k] = e[k] - eu[k] - er[k];
                            c1[k] = c2[k] + er[k];
                            c3[k] = c2[k] + eu[k];
                            c0[k] = c2[k] + eu[k] + er[k];
                            f2[k] = e[k] - eu[k];
                            f3[k] = e[k] - er[k];
                        }
                        /* face EMFs (src/prj_timeint.c:179-199): the `right`
                         * face carries E_dir in Bv1, the `up` face in Bv2 */
                        ef[0] = BV1[right][fbox_idx(&fb[right], e[0], e[1], e[2])];
                        ef[1] = BV2[up][fbox_idx(&fb[up], e[0], e[1], e[2])];
                        ef[2] = BV1[right][fbox_idx(&fb[right], f2[0], f2[1], f2[2])];
                        ef[3] = BV2[up][fbox_idx(&fb[up], f3[0], f3[1], f3[2])];
                        ec[0] = cell_emf(g, Win, bfin, d, c0[0], c0[1], c0[2], &crc);
                        ec[1] = cell_emf(g, Win, bfin, d, c1[0], c1[1], c1[2], &crc);
                        ec[2] = cell_emf(g, Win, bfin, d, c2[0], c2[1], c2[2], &crc);
                        ec[3] = cell_emf(g, Win, bfin, d, c3[0], c3[1], c3[2], &crc);
                        vn[0] = VN[right][fbox_idx(&fb[right], e[0], e[1], e[2])];
                        vn[1] = VN[up][fbox_idx(&fb[up], e[0], e[1], e[2])];
                        vn[2] = VN[right][fbox_idx(&fb[right], f2[0], f2[1], f2[2])];
                        vn[3] = VN[up][fbox_idx(&fb[up], f3[0], f3[1], f3[2])];
                        /* src/prj_timeint.c:265-290, 404-405 */
                        eta = 0.0;
                        {
                            double m[4];
                            int q;
                            m[0] = eta_mask(g, Win, bfin, c0[0], c0[1], c0[2]);
                            m[1] = eta_mask(g, Win, bfin, c1[0], c1[1], c1[2]);
                            m[2] = eta_mask(g, Win, bfin, c2[0], c2[1], c2[2]);
                            m[3] = eta_mask(g, Win, bfin, c3[0], c3[1], c3[2]);
                            for (q = 0; q < 4; ++q) {
                                if (m[q] > eta) eta = m[q];
                            }
                        }
                        eta = g->eta + eta;
                        urc = fxo_emf_upwind(ef, ec, vn, &emf);
                        if (crc == 0 && urc == 0 && eta != 0.0) {
                            /* src/prj_mhd.c:689-711 */
                            double curl_b;
                            int i = e[0], j = e[1], kk = e[2];
#define BF(dd, X, Y, Z) bf_at(g, bfin, (dd), (X), (Y), (Z))
                            if (d == 0) {
                                curl_b = (BF(2, i, j, kk) - BF(2, i, j - 1, kk)) / g->dx[1] -
                                    (BF(1, i, j, kk) - BF(1, i, j, kk - 1)) / g->dx[2];
                            } else if (d == 1) {
                                curl_b = (BF(0, i, j, kk) - BF(0, i, j, kk - 1)) / g->dx[2] -
                                    (BF(2, i, j, kk) - BF(2, i - 1, j, kk)) / g->dx[0];
                            } else {
                                curl_b = (BF(1, i, j, kk) - BF(1, i - 1, j, kk)) / g->dx[0] -
                                    (BF(0, i, j, kk) - BF(0, i, j - 1, kk)) / g->dx[1];
                            }
#undef BF
                            emf += eta * curl_b;
                        }
                        if (crc != 0 || urc != 0 || !isfinite(emf)) {
#pragma omp critical
                            {
                                if (rc == 0) {
                                    rc = 33;
                                    snprintf(g_err, sizeof(g_err),
                                        "prj_mhd_emf_upwind: non-finite input or result (dir %d edge %d %d %d)",
                                        d, ex, ey, ez);
                                }
                            }
                        }
                        EMF[d][((size_t)ex * en[1] + ey) * en[2] + ez] = emf;
                    }
                }
            }
        }
        /* CT face update (src/prj_timeint.c:415-495) */
        for (d = 0; d < 3 && rc == 0; ++d) {
            size_t fcount = (size_t)(g->n[0] + (d == 0)) * (g->n[1] + (d == 1)) * (g->n[2] + (d == 2));
            bfnew[d] = (double *)malloc(fcount * sizeof(double));
        }
        for (d = 0; d < 3 && rc == 0; ++d) {
            int fn[3];
            int xi;
            fn[0] = g->n[0] + (d == 0);
            fn[1] = g->n[1] + (d == 1);
            fn[2] = g->n[2] + (d == 2);
#pragma omp parallel for schedule(static)
            for (xi = 0; xi < fn[0]; ++xi) {
                int y, z;
                for (y = 0; y < fn[1]; ++y) {
                    for (z = 0; z < fn[2]; ++z) {
                        double value = bfin[d][fidx(g, d, xi, y, z)];
                        int x = xi;
#define EE(dd, X, Y, Z) EMF[dd][((size_t)(X) * (g->n[1] + ((dd) != 1)) + (size_t)(Y)) * \
    (g->n[2] + ((dd) != 2)) + (size_t)(Z)]
                        if (d == 0) {
                            value += -dt * (EE(2, x, y + 1, z) - EE(2, x, y, z)) / g->dx[1];
                            value += dt * (EE(1, x, y, z + 1) - EE(1, x, y, z)) / g->dx[2];
                        } else if (d == 1) {
                            value += -dt * (EE(0, x, y, z + 1) - EE(0, x, y, z)) / g->dx[2];
                            value += dt * (EE(2, x + 1, y, z) - EE(2, x, y, z)) / g->dx[0];
                        } else {
                            value += -dt * (EE(1, x + 1, y, z) - EE(1, x, y, z)) / g->dx[0];
                            value += dt * (EE(0, x, y + 1, z) - EE(0, x, y, z)) / g->dx[1];
                        }
#undef EE
                        if (stage == 2) {
                            value = 0.5 * bf0[d][fidx(g, d, xi, y, z)] + 0.5 * value;
                        }
                        bfnew[d][fidx(g, d, xi, y, z)] = value;
                    }
                }
            }
        }
    }

    /* 3. per-cell RK update (stage 1: src/prj_timeint.c:1255-1350; stage 2:
     *    :1562-1655) with the flux divergence of src/prj_flux.c:2523-2533 */
    if (rc == 0) {
        int x;
        const double *const *bfB = g->mhd ? (const double *const *)bfnew : 0;
#pragma omp parallel for schedule(static)
        for (x = 0; x < g->n[0]; ++x) {
            int y, z, v;
            for (y = 0; y < g->n[1]; ++y) {
                for (z = 0; z < g->n[2]; ++z) {
                    double w[NHMAX], u[NHMAX], u1[NHMAX], fd[NHMAX];
                    size_t c = cidx(g, x, y, z);
                    size_t fi[3], fo2[3];
                    fi[0] = fbox_idx(&fb[0], x, y, z);
                    fo2[0] = fbox_idx(&fb[0], x + 1, y, z);
                    fi[1] = fbox_idx(&fb[1], x, y, z);
                    fo2[1] = fbox_idx(&fb[1], x, y + 1, z);
                    fi[2] = fbox_idx(&fb[2], x, y, z);
                    fo2[2] = fbox_idx(&fb[2], x, y, z + 1);
                    for (v = 0; v < n; ++v) {
                        size_t s0 = (size_t)v * fbox_size(&fb[0]);
                        size_t s1 = (size_t)v * fbox_size(&fb[1]);
                        size_t s2 = (size_t)v * fbox_size(&fb[2]);
                        fd[v] = -(g->area[0] * (F[0][s0 + fo2[0]] - F[0][s0 + fi[0]]) +
                            g->area[1] * (F[1][s1 + fo2[1]] - F[1][s1 + fi[1]]) +
                            g->area[2] * (F[2][s2 + fo2[2]] - F[2][s2 + fi[2]])) / g->vol;
                    }
                    for (v = 0; v < n; ++v) w[v] = W0[(size_t)v * N + c];
                    prim2cons(n, w, u);
                    if (stage == 1) {
                        for (v = 0; v < n; ++v) u1[v] = u[v] + dt * (0.0 + fd[v]);
                    } else {
                        double us[NHMAX], ut;
                        for (v = 0; v < n; ++v) w[v] = W1[(size_t)v * N + c];
                        prim2cons(n, w, us);
                        for (v = 0; v < n; ++v) {
                            ut = us[v] + dt * (0.0 + fd[v]);
                            u1[v] = 0.5 * u[v] + 0.5 * ut;
                        }
                    }
                    if (g->mhd) {
                        /* src/prj_timeint.c:498-510 */
                        u1[B1] = 0.5 * (bfB[0][fidx(g, 0, x, y, z)] + bfB[0][fidx(g, 0, x + 1, y, z)]);
                        u1[B2] = 0.5 * (bfB[1][fidx(g, 1, x, y, z)] + bfB[1][fidx(g, 1, x, y + 1, z)]);
                        u1[B3] = 0.5 * (bfB[2][fidx(g, 2, x, y, z)] + bfB[2][fidx(g, 2, x, y, z + 1)]);
                    }
                    cons2prim(n, u1, w);
                    for (v = 0; v < n; ++v) Wout[(size_t)v * N + c] = w[v];
                }
            }
        }
        if (g->mhd) {
            for (d = 0; d < 3; ++d) {
                size_t fcount = (size_t)(g->n[0] + (d == 0)) * (g->n[1] + (d == 1)) * (g->n[2] + (d == 2));
                memcpy(bfout[d], bfnew[d], fcount * sizeof(double));
            }
        }
    }

    for (d = 0; d < 3; ++d) {
        free(F[d]);
        free(BV1[d]);
        free(BV2[d]);
        free(VN[d]);
        free(bfnew[d]);
        if (EMF[d] != 0 && (emf_out == 0 || emf_out[d] != EMF[d])) {
            free(EMF[d]);
        }
    }
    return rc;
}

/* src/prj_timeint.c:2144-2256 (sound speed only, also for MHD) */
double fxo_calc_dt(const fxo_grid *g, const double *W, double cfl)
{
    size_t N = ncell(g);
    double dt_min = 1.0e99;
    int x;
#pragma omp parallel for schedule(static) reduction(min:dt_min)
    for (x = 0; x < g->n[0]; ++x) {
        int y, z;
        for (y = 0; y < g->n[1]; ++y) {
            for (z = 0; z < g->n[2]; ++z) {
                size_t c = cidx(g, x, y, z);
                double rho = W[(size_t)RHO * N + c];
                double p = (g->gamma - 1.0) * rho * W[(size_t)EINT * N + c];
                double cs = sqrt(g->gamma * p / rho);
                double denom = (fabs(W[(size_t)V1 * N + c]) + cs) / g->dx[0] +
                    (fabs(W[(size_t)V2 * N + c]) + cs) / g->dx[1] +
                    (fabs(W[(size_t)V3 * N + c]) + cs) / g->dx[2];
                double dt_cell = cfl / denom;
                if (dt_cell < dt_min) dt_min = dt_cell;
            }
        }
    }
    return dt_min;
}


/* eSSPRK convex blend with face fields (src/prj_timeint.c:2347-2417, MHD):
 * first every face bf = w_s bf_saved + (1 - w_s) bf
 * (prj_timeint_eSSPRK_blend_bf), then per active cell
 * u = w_s u(W_saved) + (1 - w_s) u(W) with the cell field replaced by the
 * face average of the blended faces (prj_timeint_mhd_set_cons_b_from_bf),
 * W <- prim(u) (prj_timeint_eSSPRK_final_cell).                           */
int fxo_blend_mhd(const fxo_grid *g, double *W, const double *Wsaved, double *const *bf,
    const double *const *bfsaved, double saved_weight)
{
    size_t N = ncell(g);
    const double current_weight = 1.0 - saved_weight;
    int d, x;
    if (!g->mhd) {
        return fxo_fail(41, "fxo_blend_mhd: MHD only");
    }
    for (d = 0; d < 3; ++d) {
        size_t fcount = (size_t)(g->n[0] + (d == 0)) * (g->n[1] + (d == 1)) * (g->n[2] + (d == 2));
        size_t f;
        for (f = 0; f < fcount; ++f) {
            bf[d][f] = saved_weight * bfsaved[d][f] + current_weight * bf[d][f];
        }
    }
#pragma omp parallel for schedule(static)
    for (x = 0; x < g->n[0]; ++x) {
        int y, z, v;
        for (y = 0; y < g->n[1]; ++y) {
            for (z = 0; z < g->n[2]; ++z) {
                size_t c = cidx(g, x, y, z);
                double w[9], ws[9], u[9], us[9];
                for (v = 0; v < 9; ++v) {
                    w[v] = W[(size_t)v * N + c];
                    ws[v] = Wsaved[(size_t)v * N + c];
                }
                prim2cons(9, w, u);
                prim2cons(9, ws, us);
                for (v = 0; v < 9; ++v) u[v] = saved_weight * us[v] + current_weight * u[v];
                u[B1] = 0.5 * (bf[0][fidx(g, 0, x, y, z)] + bf[0][fidx(g, 0, x + 1, y, z)]);
                u[B2] = 0.5 * (bf[1][fidx(g, 1, x, y, z)] + bf[1][fidx(g, 1, x, y + 1, z)]);
                u[B3] = 0.5 * (bf[2][fidx(g, 2, x, y, z)] + bf[2][fidx(g, 2, x, y, z + 1)]);
                cons2prim(9, u, w);
                for (v = 0; v < 9; ++v) W[(size_t)v * N + c] = w[v];
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* eSSPRK convex blend (src/prj_timeint.c:2382-2417, hydro): per active cell
 * u = w_s u(W_saved) + (1 - w_s) u(W), W <- prim(u).                       */
int fxo_blend(const fxo_grid *g, double *W, const double *Wsaved, double saved_weight)
{
    size_t N = ncell(g);
    const double current_weight = 1.0 - saved_weight;
    size_t c;
    if (g->mhd) {
        return fxo_fail(41, "fxo_blend: hydro only");
    }
#pragma omp parallel for schedule(static)
    for (c = 0; c < N; ++c) {
        double w[6], ws[6], u[6], us[6];
        int v;
        for (v = 0; v < 6; ++v) {
            w[v] = W[v * N + c];
            ws[v] = Wsaved[v * N + c];
        }
        prim2cons(6, w, u);
        prim2cons(6, ws, us);
        for (v = 0; v < 6; ++v) u[v] = saved_weight * us[v] + current_weight * u[v];
        cons2prim(6, u, w);
        for (v = 0; v < 6; ++v) W[v * N + c] = w[v];
    }
    return 0;
}
