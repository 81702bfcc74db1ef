// Coarse-fine corrections of a refined mesh, on the reference's per-block
// layouts: the local phase of prj_riemann_flux_send (src/prj_riemann.c:
// 1701-1768) and prj_mhd_emf_send (src/prj_mhd.c:779-1029).
//
// The reference walks blocks and neighbour slots on the host every stage.
// Here the walk happens once per mesh topology (fx_cf_plan_create): it
// flattens every correction into a list of device records -- one per coarse
// face cell (flux) or coarse edge write (EMF) -- and each stage replays the
// lists with one launch per dependency level.  The reference's sequential
// order is kept exactly: a record that reads a location written by an
// earlier record, writes a location an earlier record reads, or overwrites an
// earlier write, lands in a later level.  An overwritten write that nobody
// reads in between is dropped (flux) or kept only for its finiteness check
// (EMF: the reference aborts on a non-finite restricted value wherever it
// occurs).  On meshes with 2:1 balance every correction has one level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/fornax_gpu.h"
#include "fx_kernels.hpp"

namespace fx {
namespace {

constexpr int NT = 256;
constexpr int BS = 16;
constexpr int NG = 2;
constexpr int RB = BS + 2 * NG;                       // 20
constexpr int NEDGES = (RB + 1) * (RB + 1) * RB;      // 8820
constexpr int FID_NONE = 0, FID_SAME = 2, FID_FINER = 3;   // src/prj_defs.h:594-599

// BIDX(nc, 0, i, j, k): the v = 0 slot of a cell; component v sits 4 v further
// (AoSoA lane width 4, src/prj_defs.h:419-434)
int vidx0(int nc, int i, int j, int k)
{
    const int kk = k + NG;
    return ((((i + NG) * RB + (j + NG)) * (RB / 4) + kk / 4) * nc) * 4 + kk % 4;
}

// EDGE_IDX (src/prj_defs.h:400-405)
__host__ __device__ inline int eidx(int dir, int i, int j, int k)
{
    const int I = i + NG, J = j + NG, K = k + NG;
    if (dir == 0) return (I * (RB + 1) + J) * (RB + 1) + K;
    if (dir == 1) return (I * RB + J) * (RB + 1) + K;
    return (I * (RB + 1) + J) * RB + K;
}

bool in_block(int i, int j, int k)
{
    return i >= -NG && i < BS + NG && j >= -NG && j < BS + NG && k >= -NG && k < BS + NG;
}

// prj_riemann_face_axis (src/prj_riemann.c:1676-1699)
int face_axis(const fx_cf_block& a, const fx_cf_block& b, int* side)
{
    int axis = -1, s = -1;
    for (int d = 0; d < 3; ++d) {
        const double tol = 1.0e-12 * a.dx[d];
        if (std::fabs(a.xmax[d] - b.xmin[d]) < tol) {
            if (axis >= 0) return -1;
            axis = d;
            s = 1;
        } else if (std::fabs(b.xmax[d] - a.xmin[d]) < tol) {
            if (axis >= 0) return -1;
            axis = d;
            s = 0;
        }
    }
    if (axis < 0) return -1;
    for (int d = 0; d < 3; ++d) {
        if (d == axis) continue;
        const double lo = std::max(a.xmin[d], b.xmin[d]);
        const double hi = std::min(a.xmax[d], b.xmax[d]);
        if (!(hi - lo > 1.0e-12)) return -1;   // prj_blocks_overlap_open
    }
    *side = s;
    return axis;
}

// Dependency levels of a sequence of writes (see the file comment).
struct Leveler {
    std::unordered_map<int64_t, int> last_write;   // level of the last kept write
    std::unordered_map<int64_t, int> last_read;    // highest level reading the location
    static int get(const std::unordered_map<int64_t, int>& m, int64_t k)
    {
        auto it = m.find(k);
        return it == m.end() ? 0 : it->second;
    }
    int place(const int64_t* src, int nsrc, int64_t dst, bool store)
    {
        int lv = 0;
        for (int s = 0; s < nsrc; ++s) lv = std::max(lv, get(last_write, src[s]));
        if (store) lv = std::max({lv, get(last_write, dst), get(last_read, dst)});
        ++lv;
        for (int s = 0; s < nsrc; ++s) {
            int& r = last_read[src[s]];
            r = std::max(r, lv);
        }
        if (store) last_write[dst] = lv;
        return lv;
    }
};

// Writes overwritten before any read: scanning backwards, a write is dead when
// a later write hits its location with no read in between.
std::vector<char> dead_writes(const std::vector<int64_t>& dst, const std::vector<int64_t>& src,
                              int nsrc)
{
    std::vector<char> dead(dst.size(), 0);
    std::unordered_map<int64_t, char> shadowed;
    for (size_t w = dst.size(); w-- > 0;) {
        auto it = shadowed.find(dst[w]);
        dead[w] = it != shadowed.end() && it->second;
        shadowed[dst[w]] = 1;
        for (int s = 0; s < nsrc; ++s) shadowed[src[w * nsrc + s]] = 0;
    }
    return dead;
}

template <class Rec>
void sort_by_level(std::vector<Rec>& recs, const std::vector<int>& level, std::vector<int>& off)
{
    const int nl = recs.empty() ? 0 : *std::max_element(level.begin(), level.end());
    std::vector<std::vector<Rec>> by(nl);
    for (size_t r = 0; r < recs.size(); ++r) by[level[r] - 1].push_back(recs[r]);
    recs.clear();
    off.assign(1, 0);
    for (auto& v : by) {
        recs.insert(recs.end(), v.begin(), v.end());
        off.push_back((int)recs.size());
    }
}

std::string build_flux(int nblocks, const fx_cf_block* blocks, int nlinks, const fx_cf_link* links,
                       int nvar, CfHostPlan& p)
{
    std::vector<CfFlux> recs;
    std::vector<int64_t> dst, src;
    for (int l = 0; l < nlinks; ++l) {
        const fx_cf_link& k = links[l];
        const int b = k.block, nid = k.neighbor;
        if (b < 0 || b >= nblocks || !blocks[b].local) continue;
        if (k.type != 1 || k.rel_level >= 0 || !k.neighbor_local) continue;
        if (nid < 0 || nid >= nblocks) continue;
        int side = 0;
        const int axis = face_axis(blocks[b], blocks[nid], &side);
        if (axis < 0) continue;
        if (blocks[b].flux[axis] == nullptr || blocks[nid].flux[axis] == nullptr)
            return "missing flux storage on block " + std::to_string(blocks[b].flux[axis] ? nid : b);
        const int t0 = (axis + 1) % 3, t1 = (axis + 2) % 3;
        for (int a = 0; a < k.recv_end[t0] - k.recv_start[t0]; ++a) {
            for (int c = 0; c < k.recv_end[t1] - k.recv_start[t1]; ++c) {
                int rc[3], q[4][3];
                rc[axis] = side == 1 ? 0 : BS;
                rc[t0] = k.recv_start[t0] + a;
                rc[t1] = k.recv_start[t1] + c;
                if (!in_block(rc[0], rc[1], rc[2]))
                    return "link " + std::to_string(l) + " writes outside the coarse block";
                CfFlux r{};
                r.dst = nid;
                r.src = b;
                r.axis = axis;
                r.dst_off = vidx0(nvar, rc[0], rc[1], rc[2]);
                for (int m = 0; m < 4; ++m) {
                    q[m][axis] = side == 1 ? BS : 0;
                    q[m][t0] = k.send_start[t0] + 2 * a + (m & 1);
                    q[m][t1] = k.send_start[t1] + 2 * c + (m >> 1);
                    if (!in_block(q[m][0], q[m][1], q[m][2]))
                        return "link " + std::to_string(l) + " reads outside the fine block";
                    r.s[m] = vidx0(nvar, q[m][0], q[m][1], q[m][2]);
                    src.push_back(((int64_t)(3 * b + axis) << 32) | (uint32_t)r.s[m]);
                }
                dst.push_back(((int64_t)(3 * nid + axis) << 32) | (uint32_t)r.dst_off);
                recs.push_back(r);
            }
        }
    }
    const std::vector<char> dead = dead_writes(dst, src, 4);
    Leveler lev;
    std::vector<CfFlux> kept;
    std::vector<int> level;
    for (size_t w = 0; w < recs.size(); ++w) {
        if (dead[w]) continue;
        level.push_back(lev.place(&src[4 * w], 4, dst[w], true));
        kept.push_back(recs[w]);
    }
    sort_by_level(kept, level, p.flux_off);
    p.flux = std::move(kept);
    return "";
}

std::string build_emf(int nblocks, const fx_cf_block* blocks, int nlinks, const fx_cf_link* links,
                      CfHostPlan& p)
{
    std::vector<CfEmf> recs;
    std::vector<int64_t> dst, src;
    for (int b = 0; b < nblocks; ++b)
        if (blocks[b].local && blocks[b].edge_fidelity[0] != nullptr) p.fid_blocks.push_back(b);
    for (int l = 0; l < nlinks; ++l) {
        const fx_cf_link& k = links[l];
        const int b = k.block, nid = k.neighbor;
        if (b < 0 || b >= nblocks || !blocks[b].local) continue;
        if (nid < 0 || nid >= nblocks || k.rel_level >= 0 || !blocks[nid].local) continue;
        const fx_cf_block& F = blocks[b];
        const fx_cf_block& Cb = blocks[nid];
        for (int d = 0; d < 3; ++d)
            if (F.emf[d] == nullptr || Cb.emf[d] == nullptr)
                return "missing emf storage on block " + std::to_string(F.emf[d] ? nid : b);
        // prj_mhd_restrict_emf_to_coarse (src/prj_mhd.c:853-965)
        int ax[3], touch = 0, snd[3], rcv[3];
        for (int d = 0; d < 3; ++d) {
            const double tol = 1.0e-12 * F.dx[d];
            ax[d] = 0;
            snd[d] = rcv[d] = -100;
            if (std::fabs(F.xmax[d] - Cb.xmin[d]) < tol) {
                ax[d] = 1;
                snd[d] = BS;
                rcv[d] = 0;
                ++touch;
            } else if (std::fabs(Cb.xmax[d] - F.xmin[d]) < tol) {
                ax[d] = -1;
                snd[d] = 0;
                rcv[d] = BS;
                ++touch;
            }
        }
        if (touch == 0 || touch > 2) continue;
        for (int dir = 0; dir < 3; ++dir) {
            if (ax[dir] != 0) continue;
            const int t0 = (dir + 1) % 3, t1 = (dir + 2) % 3;
            const int off = (F.xmin[dir] + F.xmax[dir]) < (Cb.xmin[dir] + Cb.xmax[dir]) ? 0 : BS / 2;
            int jt = -1, joff = 0, jn = 1;
            if (!(ax[t0] != 0 && ax[t1] != 0)) {
                jt = ax[t0] != 0 ? t1 : t0;
                joff = (F.xmin[jt] + F.xmax[jt]) < (Cb.xmin[jt] + Cb.xmax[jt]) ? 0 : BS / 2;
                jn = BS / 2 + 1;
            }
            for (int i = 0; i < BS; i += 2) {
                for (int j = 0; j < jn; ++j) {
                    if (jt >= 0) {
                        snd[jt] = 2 * j;
                        rcv[jt] = j + joff;
                    }
                    CfEmf r{};
                    r.dst = nid;
                    r.src = b;
                    r.dir = dir;
                    snd[dir] = i;
                    r.a = eidx(dir, snd[0], snd[1], snd[2]);
                    snd[dir] = i + 1;
                    r.b = eidx(dir, snd[0], snd[1], snd[2]);
                    rcv[dir] = i / 2 + off;
                    r.dst_idx = eidx(dir, rcv[0], rcv[1], rcv[2]);
                    r.store = 1;
                    src.push_back(((int64_t)(3 * b + dir) << 32) | (uint32_t)r.a);
                    src.push_back(((int64_t)(3 * b + dir) << 32) | (uint32_t)r.b);
                    dst.push_back(((int64_t)(3 * nid + dir) << 32) | (uint32_t)r.dst_idx);
                    recs.push_back(r);
                }
            }
        }
    }
    const std::vector<char> dead = dead_writes(dst, src, 2);
    Leveler lev;
    std::vector<int> level(recs.size());
    for (size_t w = 0; w < recs.size(); ++w) {
        recs[w].store = dead[w] ? 0 : 1;
        level[w] = lev.place(&src[2 * w], 2, dst[w], !dead[w]);
    }
    sort_by_level(recs, level, p.emf_off);
    p.emf = std::move(recs);
    return "";
}

// ---- kernels ----------------------------------------------------------------

__global__ void __launch_bounds__(NT) cf_flux_kernel(const CfPtrs* __restrict__ tab,
                                                     const CfFlux* __restrict__ recs, int n,
                                                     int nvar)
{
    const int t = blockIdx.x * NT + threadIdx.x;
    if (t >= n) return;
    const CfFlux r = recs[t];
    const double* F = tab[r.src].flux[r.axis];
    double* G = tab[r.dst].flux[r.axis];
    for (int v = 0; v < nvar; ++v) {
        const int o = 4 * v;
        G[r.dst_off + o] = 0.25 * (F[r.s[0] + o] + F[r.s[1] + o] + F[r.s[2] + o] + F[r.s[3] + o]);
    }
}

// edge fidelity reset of prj_mhd_init_edge_fidelity (src/prj_mhd.c:967-1000):
// NONE everywhere, SAME on the active edges (index 0..15 along the edge,
// 0..16 across it)
__global__ void __launch_bounds__(NT) cf_fid_init_kernel(const CfPtrs* __restrict__ tab,
                                                         const int* __restrict__ blocks, int n)
{
    const int t = blockIdx.x * NT + threadIdx.x;
    if (t >= n * 3 * NEDGES) return;
    const int e = t % NEDGES;
    const int bd = t / NEDGES;
    const int dir = bd % 3;
    const int b = blocks[bd / 3];
    const int nk = dir == 2 ? RB : RB + 1;
    const int nj = dir == 1 ? RB : RB + 1;
    const int k = e % nk - NG, j = (e / nk) % nj - NG, i = e / (nk * nj) - NG;
    const int im = dir == 0 ? BS - 1 : BS, jm = dir == 1 ? BS - 1 : BS, km = dir == 2 ? BS - 1 : BS;
    const bool act = i >= 0 && i <= im && j >= 0 && j <= jm && k >= 0 && k <= km;
    tab[b].fid[dir][e] = act ? FID_SAME : FID_NONE;
}

// prj_mhd_write_emf_edge (src/prj_mhd.c:831-851) of one restricted value
__global__ void __launch_bounds__(NT) cf_emf_kernel(const CfPtrs* __restrict__ tab,
                                                    const CfEmf* __restrict__ recs, int n,
                                                    int* status)
{
    const int t = blockIdx.x * NT + threadIdx.x;
    if (t >= n) return;
    const CfEmf r = recs[t];
    const double* E = tab[r.src].emf[r.dir];
    double value = 0.0;
    value += E[r.a];
    value += E[r.b];
    value *= 0.5;
    if (!isfinite(value)) {
        if (status != nullptr && atomicCAS(status, 0, 1) == 0) {
            status[1] = r.dst;
            status[2] = r.dir;
            status[3] = r.dst_idx;
        }
        return;
    }
    if (!r.store) return;
    int* fid = tab[r.dst].fid[r.dir];
    if (fid != nullptr && FID_FINER < fid[r.dst_idx]) return;
    tab[r.dst].emf[r.dir][r.dst_idx] = value;
    if (fid != nullptr && FID_FINER > fid[r.dst_idx]) fid[r.dst_idx] = FID_FINER;
}

}  // namespace

void cf_build_plan(int nblocks, const fx_cf_block* blocks, int nlinks, const fx_cf_link* links,
                   int nvar, CfHostPlan& p)
{
    p.flux_err = build_flux(nblocks, blocks, nlinks, links, nvar, p);
    if (!p.flux_err.empty()) {
        p.flux.clear();
        p.flux_off.assign(1, 0);
    }
    p.emf_err = build_emf(nblocks, blocks, nlinks, links, p);
    if (!p.emf_err.empty()) {
        p.emf.clear();
        p.emf_off.assign(1, 0);
        p.fid_blocks.clear();
    }
}

hipError_t launch_cf_flux(const CfPtrs* tab, const CfFlux* recs, int n, int nvar, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(cf_flux_kernel, dim3((n + NT - 1) / NT), dim3(NT), 0, s, tab, recs, n, nvar);
    return hipGetLastError();
}

hipError_t launch_cf_fid_init(const CfPtrs* tab, const int* blocks, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int total = n * 3 * NEDGES;
    hipLaunchKernelGGL(cf_fid_init_kernel, dim3((total + NT - 1) / NT), dim3(NT), 0, s, tab,
                       blocks, n);
    return hipGetLastError();
}

hipError_t launch_cf_emf(const CfPtrs* tab, const CfEmf* recs, int n, int* status, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(cf_emf_kernel, dim3((n + NT - 1) / NT), dim3(NT), 0, s, tab, recs, n,
                       status);
    return hipGetLastError();
}

}  // namespace fx
