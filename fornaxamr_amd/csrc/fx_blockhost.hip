// Per-block entry points on the reference's own block layouts
// (include/fornax_gpu.h): the device-array forms launch the kernels of
// fx_block.hip directly; the host-array forms -- one prj_flux_update,
// prj_mhd_bf2bc or prj_flux_div call on a block in host memory -- stage the
// block through device memory the library keeps per device (allocated on the
// first call there, then reused) and run on a private stream, waiting for
// that stream only.  Host-array calls are serialised by a lock, as the
// reference's block functions are not reentrant either.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fornax_gpu.h"
#include "fx_api.hpp"
#include "fx_kernels.hpp"

using namespace fx;
using namespace fx_api;

namespace {

constexpr size_t RB_CELLS = 8000;    // 20^3 (PRJ_BLOCK_NCELLS at 16^3 + 2 ghosts)
constexpr size_t RB_FACES = 8400;    // 21 * 20 * 20 (PRJ_BLOCK_NFACES)

// Device staging of one block, in doubles: W compact as [var][8000] (at most 9
// variables), eosvar, flux[3] (up to 9 variables), v_riemann[3], the face
// fields, Bv1/Bv2[3].  A call on n blocks lays out n such regions, then the n
// descriptors, then the status word.
constexpr size_t ST_W = 0;
constexpr size_t ST_E = ST_W + 9 * RB_CELLS;
constexpr size_t ST_F = ST_E + 3 * RB_CELLS;
constexpr size_t ST_V = ST_F + 3 * 9 * RB_CELLS;
constexpr size_t ST_BF = ST_V + 3 * 3 * RB_CELLS;
constexpr size_t ST_BV = ST_BF + 3 * RB_FACES;
constexpr size_t ST_DOUBLES = ST_BV + 6 * RB_CELLS;          // 457,200 doubles
// one block's region, rounded up to 256 bytes: 3,657,728 bytes
constexpr size_t ST_BYTES = (ST_DOUBLES * sizeof(double) + 255) / 256 * 256;
// after the regions of the n blocks of a call: n descriptors of ST_DESC_BYTES,
// then the status word
constexpr size_t ST_DESC_BYTES = 256;
constexpr size_t ST_STATUS_BYTES = 256;
static_assert(sizeof(fx_block_io) <= ST_DESC_BYTES && sizeof(fx_bf2bc_io) <= ST_DESC_BYTES,
              "descriptor slot");
// staging of a call on n blocks
constexpr size_t stage_bytes(int n)
{
    return (size_t)n * (ST_BYTES + ST_DESC_BYTES) + ST_STATUS_BYTES;
}
// most blocks of one fx_flux_update_block_noalloc call
constexpr int ST_MAX_BLOCKS = FX_NOALLOC_MAX_BLOCKS;

// Stages prj_flux_update on n blocks of host arrays through `base` on stream
// s: block b's arrays go to the region of ST_BYTES bytes at base + b *
// ST_BYTES, its descriptor to desc[b]; the status word follows the n
// descriptors.  Copies the blocks in, launches once, copies the outputs back
// and waits for s.  The output arrays are copied in as well, so that the
// entries prj_flux_update does not write come back unchanged.
int staged_flux_update(const fx_block_io* hbs, int n, int mhd, char* base, hipStream_t s)
{
    if (hbs == nullptr || n <= 0 || n > ST_MAX_BLOCKS) return fail(FX_EINVAL, "bad block descriptor");
    const int nh = mhd ? 9 : 6;
    for (int b = 0; b < n; ++b) {
        const fx_block_io* hb = hbs + b;
        if (hb->W == nullptr || hb->eosvar == nullptr || hb->var_stride < (long long)RB_CELLS)
            return fail(FX_EINVAL, "bad block descriptor");
        for (int d = 0; d < 3; ++d) {
            if (hb->flux[d] == nullptr || hb->v_riemann[d] == nullptr)
                return fail(FX_EINVAL, "missing flux or v_riemann array");
            if (mhd && (hb->bf[d] == nullptr || hb->Bv1[d] == nullptr || hb->Bv2[d] == nullptr))
                return fail(FX_EINVAL, "missing MHD face array");
        }
    }
    const size_t cb = RB_CELLS * sizeof(double);
    const size_t f_len = (size_t)nh * RB_CELLS, v_len = 3 * RB_CELLS;
    fx_block_io db[ST_MAX_BLOCKS];
    for (int b = 0; b < n; ++b) {
        const fx_block_io* hb = hbs + b;
        double* D = reinterpret_cast<double*>(base + (size_t)b * ST_BYTES);
        db[b] = fx_block_io{};
        db[b].W = D + ST_W;
        db[b].var_stride = (long long)RB_CELLS;
        db[b].eosvar = D + ST_E;
        // W: nh variables of 8000 cells at the caller's variable stride, compacted
        FX_HIP(hipMemcpy2DAsync(D + ST_W, cb, hb->W, (size_t)hb->var_stride * sizeof(double), cb,
                                nh, hipMemcpyHostToDevice, s));
        FX_HIP(hipMemcpyAsync(D + ST_E, hb->eosvar, 3 * cb, hipMemcpyHostToDevice, s));
        for (int d = 0; d < 3; ++d) {
            db[b].flux[d] = D + ST_F + d * f_len;
            db[b].v_riemann[d] = D + ST_V + d * v_len;
            FX_HIP(hipMemcpyAsync(db[b].flux[d], hb->flux[d], f_len * sizeof(double),
                                  hipMemcpyHostToDevice, s));
            FX_HIP(hipMemcpyAsync(db[b].v_riemann[d], hb->v_riemann[d], v_len * sizeof(double),
                                  hipMemcpyHostToDevice, s));
            if (mhd) {
                double* bfd = D + ST_BF + d * RB_FACES;
                FX_HIP(hipMemcpyAsync(bfd, hb->bf[d], RB_FACES * sizeof(double),
                                      hipMemcpyHostToDevice, s));
                db[b].bf[d] = bfd;
                db[b].Bv1[d] = D + ST_BV + (2 * d) * RB_CELLS;
                db[b].Bv2[d] = D + ST_BV + (2 * d + 1) * RB_CELLS;
                FX_HIP(hipMemcpyAsync(db[b].Bv1[d], hb->Bv1[d], cb, hipMemcpyHostToDevice, s));
                FX_HIP(hipMemcpyAsync(db[b].Bv2[d], hb->Bv2[d], cb, hipMemcpyHostToDevice, s));
            }
        }
    }
    fx_block_io* dd = reinterpret_cast<fx_block_io*>(base + (size_t)n * ST_BYTES);
    int* dstat = reinterpret_cast<int*>(base + (size_t)n * ST_BYTES + (size_t)n * ST_DESC_BYTES);
    const int zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    FX_HIP(hipMemcpyAsync(dd, db, (size_t)n * sizeof(fx_block_io), hipMemcpyHostToDevice, s));
    FX_HIP(hipMemcpyAsync(dstat, zero, sizeof(zero), hipMemcpyHostToDevice, s));
    FX_HIP(launch_block_flux(dd, n, mhd ? 1 : 0, dstat, s));
    int st[8];
    FX_HIP(hipMemcpyAsync(st, dstat, sizeof(st), hipMemcpyDeviceToHost, s));
    for (int b = 0; b < n; ++b) {
        const fx_block_io* hb = hbs + b;
        for (int d = 0; d < 3; ++d) {
            FX_HIP(hipMemcpyAsync(hb->flux[d], db[b].flux[d], f_len * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
            FX_HIP(hipMemcpyAsync(hb->v_riemann[d], db[b].v_riemann[d], v_len * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
            if (mhd) {
                FX_HIP(hipMemcpyAsync(hb->Bv1[d], db[b].Bv1[d], cb, hipMemcpyDeviceToHost, s));
                FX_HIP(hipMemcpyAsync(hb->Bv2[d], db[b].Bv2[d], cb, hipMemcpyDeviceToHost, s));
            }
        }
    }
    FX_HIP(hipStreamSynchronize(s));
    if (st[0] != 0) {
        char buf[256];
        snprintf(buf, sizeof(buf), "%s (dir %d, face %d %d %d)", phys_message(st[0]), st[1], st[2],
                 st[3], st[4]);
        return fail(FX_EPHYS, buf);
    }
    return FX_OK;
}

// prj_mhd_bf2bc(_all) on host arrays through `base`: the face fields (and the
// done flags) in, the three field variables of W in and out.
int staged_bf2bc(const fx_bf2bc_io* hb, char* base, hipStream_t s)
{
    if (hb == nullptr || hb->W == nullptr || hb->var_stride < (long long)RB_CELLS ||
        hb->bf[0] == nullptr || hb->bf[1] == nullptr || hb->bf[2] == nullptr)
        return fail(FX_EINVAL, "bad bf2bc descriptor");
    double* D = reinterpret_cast<double*>(base);
    const size_t cb = RB_CELLS * sizeof(double);
    fx_bf2bc_io db{};
    db.W = D + ST_W;
    db.var_stride = (long long)RB_CELLS;
    const size_t src_pitch = (size_t)hb->var_stride * sizeof(double);
    const double* srcB = hb->W + (size_t)6 * (size_t)hb->var_stride;   // B1..B3 are prims 6..8
    double* dstB = D + ST_W + (size_t)6 * RB_CELLS;
    FX_HIP(hipMemcpy2DAsync(dstB, cb, srcB, src_pitch, cb, 3, hipMemcpyHostToDevice, s));
    for (int d = 0; d < 3; ++d) {
        double* bfd = D + ST_BF + d * RB_FACES;
        FX_HIP(hipMemcpyAsync(bfd, hb->bf[d], RB_FACES * sizeof(double), hipMemcpyHostToDevice, s));
        db.bf[d] = bfd;
    }
    if (hb->done != nullptr) {
        int* dn = reinterpret_cast<int*>(D + ST_E);   // 8000 ints in the eosvar slot
        FX_HIP(hipMemcpyAsync(dn, hb->done, RB_CELLS * sizeof(int), hipMemcpyHostToDevice, s));
        db.done = dn;
    }
    fx_bf2bc_io* dd = reinterpret_cast<fx_bf2bc_io*>(base + ST_BYTES);
    int* dstat = reinterpret_cast<int*>(base + ST_BYTES + ST_DESC_BYTES);
    const int zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    FX_HIP(hipMemcpyAsync(dd, &db, sizeof(db), hipMemcpyHostToDevice, s));
    FX_HIP(hipMemcpyAsync(dstat, zero, sizeof(zero), hipMemcpyHostToDevice, s));
    FX_HIP(launch_block_bf2bc(dd, 1, dstat, s));
    int st[8];
    FX_HIP(hipMemcpyAsync(st, dstat, sizeof(st), hipMemcpyDeviceToHost, s));
    FX_HIP(hipMemcpy2DAsync(const_cast<double*>(srcB), src_pitch, dstB, cb, cb, 3,
                            hipMemcpyDeviceToHost, s));
    FX_HIP(hipStreamSynchronize(s));
    if (st[0] != 0) {
        char buf[256];
        snprintf(buf, sizeof(buf),
                 "prj_mhd_bf2bc: non-finite cell-centered magnetic field (cell %d %d %d)", st[2],
                 st[3], st[4]);
        return fail(FX_EPHYS, buf);
    }
    return FX_OK;
}

// Device staging kept per device for the host-array entry points.
struct HostStage {
    char* base = nullptr;
    hipStream_t stream = nullptr;
};
std::mutex g_stage_mu;
std::vector<HostStage> g_stage;

// the current device's staging (allocated on first use); g_stage_mu held
int current_stage(HostStage*& out)
{
    int dev = 0;
    FX_HIP(hipGetDevice(&dev));
    if ((int)g_stage.size() <= dev) g_stage.resize(dev + 1);
    HostStage& h = g_stage[dev];
    if (h.base == nullptr) {
        hipError_t e = hipMalloc(&h.base, stage_bytes(1));
        if (e != hipSuccess) {
            h.base = nullptr;
            return fail(e == hipErrorOutOfMemory ? FX_ENOMEM : FX_EHIP,
                        std::string("per-block staging: ") + hipGetErrorString(e));
        }
        e = hipStreamCreateWithFlags(&h.stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            (void)hipFree(h.base);
            h.base = nullptr;
            h.stream = nullptr;
            return hip_fail(e, "hipStreamCreateWithFlags");
        }
    }
    out = &h;
    return FX_OK;
}

}  // namespace

// Staging of fx_flux_update_block_noalloc: a fixed device array, so that the
// entry point allocates nothing (29,264,128 bytes: FX_NOALLOC_MAX_BLOCKS
// blocks).  Each GPU holds its own copy at its own address; host code uses the
// address hipGetSymbolAddress gives for the current device, never the array's
// name, which on the host is only a stand-in.
static __device__ __align__(256) char scratch[stage_bytes(ST_MAX_BLOCKS)];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
char* g_scratch_ptr[NOALLOC_MAX_DEVICES];   // per device; null until its first call
std::mutex g_noalloc_mu;

// the current device's address of `scratch` (looked up on the first call on
// that device, with the device current); g_noalloc_mu held
int noalloc_scratch(char*& out)
{
    int dev = 0;
    FX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(FX_EINVAL, "fx_flux_update_block_noalloc: device index beyond 63");
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        FX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) return fail(FX_EHIP, "hipGetSymbolAddress(scratch) gave no address");
        g_scratch_ptr[dev] = static_cast<char*>(p);
    }
    out = g_scratch_ptr[dev];
    return FX_OK;
}

}  // namespace

extern "C" {

int fx_flux_update_blocks(int n, const fx_block_io* blocks, int mhd, int* status, void* stream)
{
    if (n < 0 || (n > 0 && blocks == nullptr)) return fail(FX_EINVAL, "bad block array");
    if (mhd && status == nullptr) return fail(FX_EINVAL, "MHD needs a device status word");
    FX_HIP(launch_block_flux(blocks, n, mhd ? 1 : 0, status, (hipStream_t)stream));
    return FX_OK;
}

int fx_flux_update_block(const fx_block_io* hb, int mhd)
{
    std::lock_guard<std::mutex> lock(g_stage_mu);
    HostStage* h = nullptr;
    if (int rc = current_stage(h)) return rc;
    return staged_flux_update(hb, 1, mhd, h->base, h->stream);
}

int fx_flux_update_block_noalloc(int n, const fx_block_io* blocks, int mhd, void* stream)
{
    if (n < 1 || n > FX_NOALLOC_MAX_BLOCKS)
        return fail(FX_EINVAL, "fx_flux_update_block_noalloc: n = " + std::to_string(n) +
                                   " blocks, outside 1.." + std::to_string(FX_NOALLOC_MAX_BLOCKS) +
                                   " (the fixed device staging holds " +
                                   std::to_string(FX_NOALLOC_MAX_BLOCKS) + " blocks)");
    if (blocks == nullptr) return fail(FX_EINVAL, "bad block array");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    char* base = nullptr;
    if (int rc = noalloc_scratch(base)) return rc;
    return staged_flux_update(blocks, n, mhd, base, (hipStream_t)stream);
}

int fx_bf2bc_blocks(int n, const fx_bf2bc_io* blocks, int* status, void* stream)
{
    if (n < 0 || (n > 0 && blocks == nullptr)) return fail(FX_EINVAL, "bad block array");
    FX_HIP(launch_block_bf2bc(blocks, n, status, (hipStream_t)stream));
    return FX_OK;
}

int fx_bf2bc_block(const fx_bf2bc_io* hb)
{
    std::lock_guard<std::mutex> lock(g_stage_mu);
    HostStage* h = nullptr;
    if (int rc = current_stage(h)) return rc;
    return staged_bf2bc(hb, h->base, h->stream);
}

int fx_flux_div_block(double* const* flux, const double* area, double vol, int nvar,
                      double* fluxdiv)
{
    if (flux == nullptr || area == nullptr || fluxdiv == nullptr || (nvar != 6 && nvar != 9))
        return fail(FX_EINVAL, "bad flux_div arguments");
    for (int d = 0; d < 3; ++d)
        if (flux[d] == nullptr) return fail(FX_EINVAL, "missing flux array");
    std::lock_guard<std::mutex> lock(g_stage_mu);
    HostStage* h = nullptr;
    if (int rc = current_stage(h)) return rc;
    double* D = reinterpret_cast<double*>(h->base);
    const size_t f_len = (size_t)nvar * RB_CELLS;
    double* f[3];
    for (int d = 0; d < 3; ++d) {
        f[d] = D + ST_F + d * f_len;
        FX_HIP(hipMemcpyAsync(f[d], flux[d], f_len * sizeof(double), hipMemcpyHostToDevice,
                              h->stream));
    }
    double* out = D + ST_W;    // nvar x 4096 <= 9 x 8000
    FX_HIP(launch_block_flux_div(f[0], f[1], f[2], area, vol, nvar, out, h->stream));
    FX_HIP(hipMemcpyAsync(fluxdiv, out, (size_t)nvar * 4096 * sizeof(double),
                          hipMemcpyDeviceToHost, h->stream));
    FX_HIP(hipStreamSynchronize(h->stream));
    return FX_OK;
}

}  // extern "C"
