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
        int e = 0;
        double m = std::frexp(g.vol, &e);   // vol = m * 2^e, m in [0.5, 1)
        g.vol_pow2 = (m == 0.5) ? 1 : 0;
        g.inv_vol = g.vol_pow2 ? std::ldexp(1.0, 1 - e) : 1.0 / g.vol;
    }
    g.gamma = gamma;
    for (int i = 0; i < 6; ++i) g.bc[i] = bc[i];
    g.mhd = mhd;
    g.eta = eta;
    return g;
}

}  // namespace fx
