/*
 * CPU restatement of the coarse-fine corrections on a refined mesh:
 * the local phase of prj_riemann_flux_send (src/prj_riemann.c:1701-1768)
 * and prj_mhd_emf_send (src/prj_mhd.c:779-1029).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP coarse-fine kernels.
 * Only tests/ may load it; the product library never links it.
 *
 * Mesh description, as the reference holds it (src/prj_types.h:77-213):
 *   geo[b*9 + 0..8]   xmin[3], xmax[3], dx[3] of block b
 *   info[b*4 + 0..3]  id, active, level, rank
 *   slots[(b*56 + n)*16 + 0..15]  slot n of block b: id, rank, rel_level,
 *                     type, send_loc_start[3], send_loc_end[3],
 *                     recv_loc_start[3], recv_loc_end[3]
 * Block arrays: flux[b*3 + axis] (nvar x 8000, VIDX = BIDX(nvar, ...)),
 * emf[b*3 + dir] and fid[b*3 + dir] (8820 each, EDGE_IDX); NULL where a block
 * has no storage.
 */
#include <math.h>
#include <stddef.h>

#define CF_BS 16
#define CF_NG 2
#define CF_RB (CF_BS + 2 * CF_NG)
#define CF_NEDGES ((CF_RB + 1) * (CF_RB + 1) * CF_RB)
#define CF_NSLOT 56
#define CF_SLOTW 16

enum { FID_NONE = 0, FID_SAME = 2, FID_FINER = 3 };   /* src/prj_defs.h:594-599 */
enum { NB_FACE = 1 };                                  /* src/prj_types.h:64-69 */

/* BIDX(nc, v, i, j, k) with AoSoA width 4 (src/prj_defs.h:419-434) */
static size_t cf_vidx(int nc, int v, int i, int j, int k)
{
    const int kk = k + CF_NG;
    return ((((size_t)(i + CF_NG) * CF_RB + (size_t)(j + CF_NG)) * (CF_RB / 4) + (size_t)(kk / 4)) *
                (size_t)nc + (size_t)v) * 4 + (size_t)(kk % 4);
}

/* EDGE_IDX (src/prj_defs.h:400-405) */
static size_t cf_eidx(int dir, int i, int j, int k)
{
    const size_t I = (size_t)(i + CF_NG), J = (size_t)(j + CF_NG), K = (size_t)(k + CF_NG);
    if (dir == 0) return (I * (CF_RB + 1) + J) * (CF_RB + 1) + K;
    if (dir == 1) return (I * CF_RB + J) * (CF_RB + 1) + K;
    return (I * (CF_RB + 1) + J) * CF_RB + K;
}

static int cf_overlap_open(double amin, double amax, double bmin, double bmax)
{
    const double lo = amin > bmin ? amin : bmin;
    const double hi = amax < bmax ? amax : bmax;
    return hi - lo > 1.0e-12;
}

/* prj_riemann_face_axis (src/prj_riemann.c:1676-1699): the axis along which
 * block a's box meets box b face-on (exactly one touching axis, open overlap
 * on the others), side 1 if b lies above a.  -1 otherwise. */
static int cf_face_axis(const double *ga, const double *gb, int *side)
{
    int axis = -1, s = -1, d;
    for (d = 0; d < 3; ++d) {
        const double tol = 1.0e-12 * ga[6 + d];
        if (fabs(ga[3 + d] - gb[d]) < tol) {
            if (axis >= 0) return -1;
            axis = d;
            s = 1;
        } else if (fabs(gb[3 + d] - ga[d]) < tol) {
            if (axis >= 0) return -1;
            axis = d;
            s = 0;
        }
    }
    if (axis < 0) return -1;
    for (d = 0; d < 3; ++d) {
        if (d != axis && !cf_overlap_open(ga[d], ga[3 + d], gb[d], gb[3 + d])) return -1;
    }
    *side = s;
    return axis;
}

static int cf_local(const int *info, int b, int my_rank)
{
    return info[4 * b] >= 0 && info[4 * b + 1] == 1 && info[4 * b + 3] == my_rank;
}

/* Local phase of prj_riemann_flux_send: every face slot of a local block whose
 * neighbour is coarser and on this rank overwrites the neighbour's face flux
 * with the mean of the four fine faces behind each coarse face, summed
 * s0 + s1 + s2 + s3 with s1/s2 stepping tan0/tan1 (src/prj_riemann.c:1730-1765). */
void fxo_cf_flux_send(int nblocks, const double *geo, const int *info, const int *slots,
    int nvar, double *const *flux, int my_rank)
{
    int b, n;
    for (b = 0; b < nblocks; ++b) {
        if (!cf_local(info, b, my_rank)) continue;
        for (n = 0; n < CF_NSLOT; ++n) {
            const int *s = slots + ((size_t)b * CF_NSLOT + n) * CF_SLOTW;
            const int nid = s[0];
            const int *rs = s + 10, *re = s + 13, *ss = s + 4;
            int side = 0, axis, t0, t1, a, c;
            if (s[3] != NB_FACE || s[2] >= 0 || s[1] != my_rank) continue;
            if (nid < 0 || nid >= nblocks) continue;
            axis = cf_face_axis(geo + 9 * (size_t)b, geo + 9 * (size_t)nid, &side);
            if (axis < 0) continue;
            t0 = (axis + 1) % 3;
            t1 = (axis + 2) % 3;
            for (a = 0; a < re[t0] - rs[t0]; ++a) {
                for (c = 0; c < re[t1] - rs[t1]; ++c) {
                    int rc[3], q[4][3], v, m;
                    rc[axis] = side == 1 ? 0 : CF_BS;
                    rc[t0] = rs[t0] + a;
                    rc[t1] = rs[t1] + c;
                    for (m = 0; m < 4; ++m) {
                        q[m][axis] = side == 1 ? CF_BS : 0;
                        q[m][t0] = ss[t0] + 2 * a + (m & 1);
                        q[m][t1] = ss[t1] + 2 * c + (m >> 1);
                    }
                    for (v = 0; v < nvar; ++v) {
                        const double *F = flux[3 * (size_t)b + axis];
                        const double sum = F[cf_vidx(nvar, v, q[0][0], q[0][1], q[0][2])] +
                                           F[cf_vidx(nvar, v, q[1][0], q[1][1], q[1][2])] +
                                           F[cf_vidx(nvar, v, q[2][0], q[2][1], q[2][2])] +
                                           F[cf_vidx(nvar, v, q[3][0], q[3][1], q[3][2])];
                        flux[3 * (size_t)nid + axis][cf_vidx(nvar, v, rc[0], rc[1], rc[2])] = 0.25 * sum;
                    }
                }
            }
        }
    }
}

/* one coarse-edge write of prj_mhd_write_emf_edge (src/prj_mhd.c:831-851) */
static int cf_emf_write(double *emf, int *fid, size_t idx, double value)
{
    if (!isfinite(value)) return 1;
    if (fid != NULL) {
        if (FID_FINER < fid[idx]) return 0;
        emf[idx] = value;
        if (FID_FINER > fid[idx]) fid[idx] = FID_FINER;
    } else {
        emf[idx] = value;
    }
    return 0;
}

/* prj_mhd_restrict_emf_to_coarse (src/prj_mhd.c:853-965): the coarse edges
 * along the contact of fine block f and coarse block c take the mean of the
 * two fine edges they cover, accumulated as (0 + e0) + e1, then halved. */
static int cf_restrict(const double *gf, const double *gc, double *const *emf_f,
    double *const *emf_c, int *const *fid_c)
{
    int touch = 0, d, dir, ax[3], snd[3], rcv[3];
    for (d = 0; d < 3; ++d) {
        const double tol = 1.0e-12 * gf[6 + d];
        ax[d] = 0;
        if (fabs(gf[3 + d] - gc[d]) < tol) {
            ax[d] = 1;
            ++touch;
        } else if (fabs(gc[3 + d] - gf[d]) < tol) {
            ax[d] = -1;
            ++touch;
        }
        snd[d] = -100;
        rcv[d] = -100;
        if (ax[d] == 1) {
            snd[d] = CF_BS;
            rcv[d] = 0;
        } else if (ax[d] == -1) {
            snd[d] = 0;
            rcv[d] = CF_BS;
        }
    }
    if (touch == 0 || touch > 2) return 0;
    for (dir = 0; dir < 3; ++dir) {
        const int t0 = (dir + 1) % 3, t1 = (dir + 2) % 3;
        int off, i, j, jt = -1, joff = 0, jn;
        if (ax[dir] != 0) continue;
        off = (gf[dir] + gf[3 + dir]) < (gc[dir] + gc[3 + dir]) ? 0 : CF_BS / 2;
        if (ax[t0] != 0 && ax[t1] != 0) {
            jn = 1;
        } else {
            jt = ax[t0] != 0 ? t1 : t0;
            joff = (gf[jt] + gf[3 + jt]) < (gc[jt] + gc[3 + jt]) ? 0 : CF_BS / 2;
            jn = CF_BS / 2 + 1;
        }
        for (i = 0; i < CF_BS; i += 2) {
            for (j = 0; j < jn; ++j) {
                double value = 0.0;
                if (jt >= 0) {
                    snd[jt] = 2 * j;
                    rcv[jt] = j + joff;
                }
                snd[dir] = i;
                value += emf_f[dir][cf_eidx(dir, snd[0], snd[1], snd[2])];
                snd[dir] = i + 1;
                value += emf_f[dir][cf_eidx(dir, snd[0], snd[1], snd[2])];
                value *= 0.5;
                rcv[dir] = i / 2 + off;
                if (cf_emf_write(emf_c[dir], fid_c ? fid_c[dir] : NULL,
                        cf_eidx(dir, rcv[0], rcv[1], rcv[2]), value))
                    return 1;
            }
        }
    }
    return 0;
}

/* prj_mhd_emf_send (src/prj_mhd.c:967-1029): edge fidelity of every local
 * block reset (NONE, SAME on active edges), then each local block restricts
 * onto every coarser local neighbour in slot order.  Returns 1 at the first
 * non-finite restricted value (the reference aborts there), else 0. */
int fxo_cf_emf_send(int nblocks, const double *geo, const int *info, const int *slots,
    double *const *emf, int *const *fid, int my_rank)
{
    int b, n;
    for (b = 0; b < nblocks; ++b) {
        int dir, i, j, k;
        if (!cf_local(info, b, my_rank) || fid == NULL || fid[3 * (size_t)b] == NULL) continue;
        for (dir = 0; dir < 3; ++dir) {
            int *F = fid[3 * (size_t)b + dir];
            for (n = 0; n < CF_NEDGES; ++n) F[n] = FID_NONE;
            for (i = 0; i <= (dir == 0 ? CF_BS - 1 : CF_BS); ++i)
                for (j = 0; j <= (dir == 1 ? CF_BS - 1 : CF_BS); ++j)
                    for (k = 0; k <= (dir == 2 ? CF_BS - 1 : CF_BS); ++k)
                        F[cf_eidx(dir, i, j, k)] = FID_SAME;
        }
    }
    for (b = 0; b < nblocks; ++b) {
        if (!cf_local(info, b, my_rank)) continue;
        for (n = 0; n < CF_NSLOT; ++n) {
            const int *s = slots + ((size_t)b * CF_NSLOT + n) * CF_SLOTW;
            const int nid = s[0];
            if (nid < 0 || nid >= nblocks || s[2] >= 0) continue;
            if (!cf_local(info, nid, my_rank) || info[4 * nid + 3] != info[4 * b + 3]) continue;
            if (cf_restrict(geo + 9 * (size_t)b, geo + 9 * (size_t)nid, emf + 3 * (size_t)b,
                    emf + 3 * (size_t)nid, fid ? fid + 3 * (size_t)nid : NULL))
                return 1;
        }
    }
    return 0;
}
