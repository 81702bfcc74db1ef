/*
 * fxa_oracle.c -- CPU oracle of the differentiable hydro scheme used for the
 * discrete adjoint: smooth van Albada PLM + HLLE (Davis speeds), RK2 with a
 * fixed dt, periodic box, ideal gas.  TEST INFRASTRUCTURE: loaded only by
 * tests and by bench.py's CPU leg as the checker, never by the product.
 *
 * Cell fields are compact [var][X][Y][Z] (Z fastest), var = rho, v1, v2, v3,
 * eint, Ye -- the prim order of src/prj_defs.h:466-494.
 *
 * The scheme is not in the reference (no van Albada, no HLLE); its pieces are
 * pinned by closed forms and invariants in tests/test_adjoint_oracle.py:
 * HLLE consistency F(q,q) = F_phys(q), supersonic upwinding, mirror symmetry,
 * van Albada limits, free-stream preservation, exact periodic conservation;
 * the tangent-linear model (the same body on dual numbers) is pinned against
 * central differences of the primal.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct fxa_grid {
    int n[3];
    double dx[3];
    double area[3];
    double vol;
    double gamma;
    double va_eps2;     /* van Albada regularisation eps^2 */
} fxa_grid;

static inline int64_t fxa_cell(const fxa_grid* g, int x, int y, int z)
{
    return ((int64_t)x * g->n[1] + y) * g->n[2] + z;
}

static void fxa_grid_make(fxa_grid* g, const int* n, const double* xlim, double gamma,
                          double va_eps2)
{
    g->va_eps2 = va_eps2;
    for (int d = 0; d < 3; ++d) {
        g->n[d] = n[d];
        g->dx[d] = (xlim[2 * d + 1] - xlim[2 * d]) / (double)n[d];
    }
    g->vol = g->dx[0] * g->dx[1] * g->dx[2];
    g->area[0] = g->dx[1] * g->dx[2];
    g->area[1] = g->dx[0] * g->dx[2];
    g->area[2] = g->dx[0] * g->dx[1];
    g->gamma = gamma;
}

/* ---- primal instantiation ------------------------------------------------ */
#define T double
#define FN(name) p_##name
#define ADD(a, b) ((a) + (b))
#define SUB(a, b) ((a) - (b))
#define MUL(a, b) ((a) * (b))
#define DIV(a, b) ((a) / (b))
#define SQRT(a) sqrt(a)
#define CST(x) ((double)(x))
#define MIN2(a, b) p_min2((a), (b))
#define MAX2(a, b) p_max2((a), (b))
static inline double p_min2(double a, double b) { return a < b ? a : b; }
static inline double p_max2(double a, double b) { return a > b ? a : b; }
#include "fxa_body.inc"
#undef T
#undef FN
#undef ADD
#undef SUB
#undef MUL
#undef DIV
#undef SQRT
#undef CST
#undef MIN2
#undef MAX2

/* ---- tangent instantiation (dual numbers) --------------------------------- */
typedef struct dual {
    double v, d;
} dual;
static inline dual d_add(dual a, dual b) { dual r = {a.v + b.v, a.d + b.d}; return r; }
static inline dual d_sub(dual a, dual b) { dual r = {a.v - b.v, a.d - b.d}; return r; }
static inline dual d_mul(dual a, dual b) { dual r = {a.v * b.v, a.d * b.v + a.v * b.d}; return r; }
static inline dual d_div(dual a, dual b)
{
    double q = a.v / b.v;
    dual r = {q, (a.d - q * b.d) / b.v};
    return r;
}
static inline dual d_sqrt(dual a)
{
    double s = sqrt(a.v);
    dual r = {s, a.d / (2.0 * s)};
    return r;
}
static inline dual d_cst(double x) { dual r = {x, 0.0}; return r; }
static inline dual d_min2(dual a, dual b)
{
    if (a.v < b.v) return a;
    if (b.v < a.v) return b;
    dual r = {a.v, 0.5 * (a.d + b.d)};
    return r;
}
static inline dual d_max2(dual a, dual b)
{
    if (a.v > b.v) return a;
    if (b.v > a.v) return b;
    dual r = {a.v, 0.5 * (a.d + b.d)};
    return r;
}
#define T dual
#define FN(name) t_##name
#define ADD(a, b) d_add((a), (b))
#define SUB(a, b) d_sub((a), (b))
#define MUL(a, b) d_mul((a), (b))
#define DIV(a, b) d_div((a), (b))
#define SQRT(a) d_sqrt(a)
#define CST(x) d_cst((double)(x))
#define MIN2(a, b) d_min2((a), (b))
#define MAX2(a, b) d_max2((a), (b))
#include "fxa_body.inc"

/* ---- exported ------------------------------------------------------------ */

double fxa_va(double sl, double sr, double eps2) { return p_va(sl, sr, eps2); }

void fxa_hlle(const double* qL, const double* qR, double gamma, double* F)
{
    p_hlle(qL, qR, gamma, F);
}

/* nsteps RK2 steps; W0 in, WT out (may alias); returns 0 or -4 (memory) */
int fxa_forward(const int* n, const double* xlim, double gamma, double va_eps2, double dt,
                int nsteps, const double* W0, double* WT)
{
    fxa_grid g;
    fxa_grid_make(&g, n, xlim, gamma, va_eps2);
    const int64_t N = (int64_t)n[0] * n[1] * n[2];
    double* W1 = malloc(sizeof(double) * 6 * N);
    double* F = malloc(sizeof(double) * 18 * N);
    if (!W1 || !F) {
        free(W1);
        free(F);
        return -4;
    }
    if (WT != W0) memcpy(WT, W0, sizeof(double) * 6 * N);
    p_run(&g, dt, nsteps, WT, W1, F);
    free(W1);
    free(F);
    return 0;
}

/* Tangent-linear model: WT = forward(W0), dWT = d forward(W0) . dW0 */
int fxa_tlm(const int* n, const double* xlim, double gamma, double va_eps2, double dt,
            int nsteps, const double* W0, const double* dW0, double* WT, double* dWT)
{
    fxa_grid g;
    fxa_grid_make(&g, n, xlim, gamma, va_eps2);
    const int64_t N = (int64_t)n[0] * n[1] * n[2];
    dual* W = malloc(sizeof(dual) * 6 * N);
    dual* W1 = malloc(sizeof(dual) * 6 * N);
    dual* F = malloc(sizeof(dual) * 18 * N);
    if (!W || !W1 || !F) {
        free(W);
        free(W1);
        free(F);
        return -4;
    }
    for (int64_t i = 0; i < 6 * N; ++i) {
        W[i].v = W0[i];
        W[i].d = dW0[i];
    }
    t_run(&g, dt, nsteps, W, W1, F);
    for (int64_t i = 0; i < 6 * N; ++i) {
        WT[i] = W[i].v;
        dWT[i] = W[i].d;
    }
    free(W);
    free(W1);
    free(F);
    return 0;
}
