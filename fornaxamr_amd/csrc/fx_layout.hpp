// Device data layout of a uniform single-level mesh of 16^3 blocks.
//
// Cell fields are stored per variable as one padded global array
//   [X + 2][Y + 2][Z + PADZ]  over  X in [-2, NX+2), Y in [-2, NY+2), Z in [-PADZ, NZ+PADZ)
// i.e. the whole mesh is one array with a two-cell ghost shell around the
// domain only.  Blocks hold no ghost copies of each other: a block's stencil
// reads its neighbours' cells directly (they are in HBM or L2/MALL), and only
// the domain boundary needs a ghost fill.  The z padding is 16 cells on each
// side so that every 16-cell block row starts on a 128-byte boundary.
//
// Face fields (MHD) use the same strides with one extra slot along every axis
// (face x of Bx at X in [-2, NX+3)).
#pragma once

#include <cmath>
#include <cstdint>

namespace fx {

constexpr int BS = 16;          // block edge (reference PRJ_BLOCK_SIZE built as 16)
constexpr int NG = 2;           // ghost depth (PRJ_NGHOST)
constexpr int PADZ = 16;        // z padding (alignment)

struct Geom {
    int nx, ny, nz;             // active cells
    int nbx, nby, nbz;          // blocks per axis
    int64_t sy;                 // stride of Y (= padded row length)
    int64_t sx;                 // stride of X
    int64_t vstride;            // stride between variables of one cell field
    int64_t fvstride;           // stride between face-field components
    int64_t origin;             // offset of cell (0,0,0)
    double dx[3];
    double area[3];
    double vol;
    double inv_vol;             // 1/vol, exact when vol_pow2
    int vol_pow2;               // vol is a power of two: x/vol == x*inv_vol bit for bit
    double inv_dx[3];           // 1/dx, exact when dx_pow2
    int dx_pow2;                // every dx is a power of two with a finite reciprocal
    double gamma;
    int bc[6];                  // reference codes: 0 outflow, 1 reflect, 3 periodic
    int mhd;
    double eta;                 // uniform CT resistivity
};

__host__ __device__ inline int64_t cell_off(const Geom& g, int x, int y, int z)
{
    return g.origin + (int64_t)x * g.sx + (int64_t)y * g.sy + z;
}

// Ghost mapping of one axis: the active cell a ghost coordinate c stands for
// -- periodic wrap (the wrapped neighbour of src/prj_mesh.c:1597-1645) or the
// mirror copy of the physical boundary (src/prj_boundary.c:413-487: src =
// -c-1 / 2n-1-c) -- with bit `axis` of flip toggled where a reflecting face
// negates the normal velocity.  Active coordinates map to themselves.
// Written with selects only: inside the stage kernels it sits between a
// prefetch and its use, where a branch would split the schedule.
__host__ __device__ inline int map_axis(const Geom& g, int axis, int n, int c, int& flip)
{
    const bool lo = c < 0, out = lo || c >= n;
    const int t = lo ? g.bc[2 * axis] : g.bc[2 * axis + 1];
    const int wrapped = lo ? c + n : c - n;
    const int mirror = lo ? -c - 1 : 2 * n - 1 - c;
    flip ^= (out && t == 1) ? 1 << axis : 0;
    return out ? (t == 3 ? wrapped : mirror) : c;
}

// CFL denominator of one cell from its primitives (src/prj_timeint.c:2163-2182):
// (|v1| + cs)/dx1 + (|v2| + cs)/dx2 + (|v3| + cs)/dx3 with cs^2 = gamma p / rho
// and p = (gamma - 1) rho e.  Dividing by a power of two and multiplying by
// its (exact) reciprocal round the same real number, so the dx_pow2 form is
// the same double.
__host__ __device__ inline double cfl_denom(const Geom& g, double rho, double v1, double v2,
                                            double v3, double e)
{
    const double p = (g.gamma - 1.0) * rho * e;
    const double cs = sqrt(g.gamma * p / rho);
    const double a = fabs(v1) + cs, b = fabs(v2) + cs, c = fabs(v3) + cs;
    if (g.dx_pow2) return a * g.inv_dx[0] + b * g.inv_dx[1] + c * g.inv_dx[2];
    return a / g.dx[0] + b / g.dx[1] + c / g.dx[2];
}

// exact reciprocal of x when x is a power of two whose reciprocal is a
// finite double, else 0
inline double pow2_reciprocal(double x)
{
    int e = 0;
    const double m = std::frexp(x, &e);   // x = m * 2^e, m in [0.5, 1)
    if (m != 0.5) return 0.0;
    const double r = std::ldexp(1.0, 1 - e);
    return std::isfinite(r) ? r : 0.0;
}

inline Geom make_geom(int nx, int ny, int nz, const double* xlim, const int* bc, double gamma,
                      int mhd, double eta)
{
    Geom g{};
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.nbx = nx / BS; g.nby = ny / BS; g.nbz = nz / BS;
    g.sy = nz + 2 * PADZ;
    g.sx = (int64_t)(ny + 2 * NG + 1) * g.sy;
    g.vstride = (int64_t)(nx + 2 * NG + 1) * g.sx;
    g.fvstride = g.vstride;
    g.origin = (int64_t)NG * g.sx + (int64_t)NG * g.sy + PADZ;
    const int n[3] = {nx, ny, nz};
    for (int d = 0; d < 3; ++d) {
        // block extent first, then cell width (src/prj_mesh.c:1553-1573)
        double block_dx = (xlim[2 * d + 1] - xlim[2 * d]) / (double)(n[d] / BS);
        g.dx[d] = block_dx / (double)BS;
    }
    // src/prj_mesh.c:675-687
    g.vol = g.dx[0] * g.dx[1] * g.dx[2];
    g.area[0] = g.dx[1] * g.dx[2];
    g.area[1] = g.dx[0] * g.dx[2];
    g.area[2] = g.dx[0] * g.dx[1];
    {
        const double r = pow2_reciprocal(g.vol);
        g.vol_pow2 = r != 0.0 ? 1 : 0;
        g.inv_vol = g.vol_pow2 ? r : 1.0 / g.vol;
        g.dx_pow2 = 1;
        for (int d = 0; d < 3; ++d) {
            const double rd = pow2_reciprocal(g.dx[d]);
            if (rd == 0.0) g.dx_pow2 = 0;
            g.inv_dx[d] = rd != 0.0 ? rd : 1.0 / g.dx[d];
        }
    }
    g.gamma = gamma;
    for (int i = 0; i < 6; ++i) g.bc[i] = bc[i];
    g.mhd = mhd;
    g.eta = eta;
    return g;
}

}  // namespace fx
