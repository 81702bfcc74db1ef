// C ABI over the gfx950 kernels (include/fornax_gpu.h).
//
// A fx_mesh owns the device-resident state of one uniform mesh: two stage
// slots of the cell primitives (the reference's W_mhd slots 0/1), for MHD
// three face-field buffers (slots 0/1 plus a ping-pong target for the
// stage-2 CT update), the MHD scratch, the device dt and an error word.
// All work is stream-ordered on the mesh's stream; nothing synchronises
// except the calls documented to do so.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fornax_gpu.h"
#include "fx_kernels.hpp"

using namespace fx;

struct fx_mesh {
    Geom g;
    int nh;
    double* W[2] = {nullptr, nullptr};
    double* bf[3] = {nullptr, nullptr, nullptr};   // [slot 0, slot 1, spare]
    double* U1 = nullptr;
    double* fdat = nullptr;
    double* dev = nullptr;   // [0] dt, [1] dt_prev, [2] reduction scratch
    int* err = nullptr;      // [code, dir, x, y, z, stage]
    hipStream_t stream = nullptr;
    size_t cell_bytes = 0;   // bytes of one padded variable
    // optional event timing of the stage kernels (excluding the ghost fill)
    bool timing = false;
    struct Span { hipEvent_t a, b; int stage; };
    std::vector<Span> spans;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char* where)
{
    return fail(FX_EHIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define FX_HIP(call)                                    \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

const char* phys_message(int code)
{
    switch (code) {
    case 11: return "prj_riemann_hlld: non-positive density";
    case 12: return "prj_riemann_hlld: non-positive gas pressure";
    case 13: return "prj_riemann_hlld: non-positive adiabatic index";
    case 14: return "prj_riemann_hlld: negative fast-speed discriminant";
    case 15: return "prj_riemann_hlld: invalid fast speed";
    case 16: return "prj_riemann_hlld: invalid fast-wave ordering";
    case 17: return "prj_riemann_hlld: degenerate contact-speed denominator";
    case 18: return "prj_riemann_hlld: outer wave coincides with contact wave";
    case 19: return "prj_riemann_hlld: invalid star density";
    case 20: return "prj_riemann_hlld: invalid double-star density weights";
    case 32: return "prj_timeint_mhd_cell_emf_all: non-finite cell emf";
    case 33: return "prj_mhd_emf_upwind: non-finite input or result";
    case 34: return "prj_timeint_mhd_ct_value: non-finite updated magnetic field";
    default: return "device error";
    }
}

int check_mesh(fx_mesh* m)
{
    if (m == nullptr) return fail(FX_EINVAL, "null mesh");
    return FX_OK;
}

int read_error_word(fx_mesh* m)
{
    int h[6] = {0, 0, 0, 0, 0, 0};
    FX_HIP(hipMemcpyAsync(h, m->err, sizeof(h), hipMemcpyDeviceToHost, m->stream));
    FX_HIP(hipStreamSynchronize(m->stream));
    if (h[0] != 0) {
        char buf[256];
        snprintf(buf, sizeof(buf), "%s (stage %d, dir %d, at %d %d %d)", phys_message(h[0]), h[5],
                 h[1], h[2], h[3], h[4]);
        // clear for the next run
        int z[6] = {0, 0, 0, 0, 0, 0};
        FX_HIP(hipMemcpy(m->err, z, sizeof(z), hipMemcpyHostToDevice));
        return fail(FX_EPHYS, buf);
    }
    return FX_OK;
}

// host global [nh][X][Y][Z] <-> device padded, per (var, x) plane
int copy_cells(fx_mesh* m, double* dev, const double* host_in, double* host_out, int nh)
{
    const Geom& g = m->g;
    const size_t row = (size_t)g.nz * sizeof(double);
    for (int v = 0; v < nh; ++v) {
        for (int x = 0; x < g.nx; ++x) {
            double* d = dev + (int64_t)v * g.vstride + cell_off(g, x, 0, 0);
            size_t hoff = ((size_t)v * g.nx + x) * (size_t)g.ny * g.nz;
            if (host_in) {
                FX_HIP(hipMemcpy2DAsync(d, g.sy * sizeof(double), host_in + hoff, row, row, g.ny,
                                        hipMemcpyHostToDevice, m->stream));
            } else {
                FX_HIP(hipMemcpy2DAsync(host_out + hoff, row, d, g.sy * sizeof(double), row, g.ny,
                                        hipMemcpyDeviceToHost, m->stream));
            }
        }
    }
    FX_HIP(hipStreamSynchronize(m->stream));
    return FX_OK;
}

int copy_faces(fx_mesh* m, double* dev, double* const* host_in, double* const* host_out)
{
    const Geom& g = m->g;
    for (int d = 0; d < 3; ++d) {
        int fx_ = g.nx + (d == 0), fy = g.ny + (d == 1), fz = g.nz + (d == 2);
        const size_t row = (size_t)fz * sizeof(double);
        for (int x = 0; x < fx_; ++x) {
            double* dp = dev + (int64_t)d * g.fvstride + cell_off(g, x, 0, 0);
            size_t hoff = (size_t)x * fy * fz;
            if (host_in) {
                FX_HIP(hipMemcpy2DAsync(dp, g.sy * sizeof(double), host_in[d] + hoff, row, row, fy,
                                        hipMemcpyHostToDevice, m->stream));
            } else {
                FX_HIP(hipMemcpy2DAsync(host_out[d] + hoff, row, dp, g.sy * sizeof(double), row, fy,
                                        hipMemcpyDeviceToHost, m->stream));
            }
        }
    }
    FX_HIP(hipStreamSynchronize(m->stream));
    return FX_OK;
}

int run_stage(fx_mesh* m, int stage, const double* dt_ptr, double dt)
{
    const Geom& g = m->g;
    double* Win = stage == 1 ? m->W[0] : m->W[1];
    if (!g.mhd) {
        FX_HIP(launch_boundary_fill(g, Win, nullptr, m->stream));
        HydroStageArgs a;
        a.g = g;
        a.Win = Win;
        a.W0 = m->W[0];
        a.Wout = stage == 1 ? m->W[1] : m->W[0];
        a.dt_ptr = dt_ptr;
        a.dt = dt;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (m->timing) {
            FX_HIP(hipEventCreate(&e0));
            FX_HIP(hipEventCreate(&e1));
            FX_HIP(hipEventRecord(e0, m->stream));
        }
        FX_HIP(launch_hydro_stage(a, stage, m->stream));
        if (m->timing) {
            FX_HIP(hipEventRecord(e1, m->stream));
            m->spans.push_back({e0, e1, stage});
        }
        return FX_OK;
    }
    // MHD: stage 1 reads face slot 0 and writes slot 1; stage 2 reads slots
    // 0 and 1 and writes the spare buffer, which then becomes slot 0.
    double* bf_in = stage == 1 ? m->bf[0] : m->bf[1];
    FX_HIP(launch_boundary_fill(g, Win, bf_in, m->stream));
    MhdStageArgs a;
    a.g = g;
    a.Win = Win;
    a.W0 = m->W[0];
    a.Wout = stage == 1 ? m->W[1] : m->W[0];
    a.bf_in = bf_in;
    a.bf0 = m->bf[0];
    a.bf_out = stage == 1 ? m->bf[1] : m->bf[2];
    a.U1 = m->U1;
    a.fdat = m->fdat;
    a.err = m->err;
    a.dt_ptr = dt_ptr;
    a.dt = dt;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (m->timing) {
        FX_HIP(hipEventCreate(&e0));
        FX_HIP(hipEventCreate(&e1));
        FX_HIP(hipEventRecord(e0, m->stream));
    }
    FX_HIP(launch_mhd_stage(a, stage, m->stream));
    if (m->timing) {
        FX_HIP(hipEventRecord(e1, m->stream));
        m->spans.push_back({e0, e1, stage});
    }
    if (stage == 2) {
        double* t = m->bf[0];
        m->bf[0] = m->bf[2];
        m->bf[2] = t;
    }
    return FX_OK;
}

}  // namespace

extern "C" {

const char* fx_last_error(void) { return g_err.c_str(); }

const char* fx_version(void)
{
    return "fornax_gpu gfx950 fp64 (ffp-contract=off)";
}

int fx_mesh_create(const fx_mesh_desc* desc, fx_mesh** out)
{
    if (desc == nullptr || out == nullptr) return fail(FX_EINVAL, "null argument");
    *out = nullptr;
    for (int d = 0; d < 3; ++d) {
        if (desc->n[d] <= 0 || desc->n[d] % BS != 0)
            return fail(FX_EINVAL, "grid extent must be a positive multiple of 16");
        bool lo = desc->bc[2 * d] == 3, hi = desc->bc[2 * d + 1] == 3;
        if (lo != hi) return fail(FX_EINVAL, "periodic boundaries must be paired");
        if (lo && desc->n[d] < 2 * BS)
            return fail(FX_EINVAL, "a periodic axis needs at least two blocks");
        for (int s = 0; s < 2; ++s) {
            int b = desc->bc[2 * d + s];
            if (b != 0 && b != 1 && b != 3) return fail(FX_EINVAL, "unsupported boundary code");
            if (desc->mhd && b != 0)
                return fail(FX_EINVAL,
                            "only outflow magnetic-field boundary conditions are implemented");
        }
    }
    if (!(desc->gamma > 1.0)) return fail(FX_EINVAL, "gamma must exceed 1");
    fx_mesh* m = new fx_mesh();
    m->g = make_geom(desc->n[0], desc->n[1], desc->n[2], desc->xlim, desc->bc, desc->gamma,
                     desc->mhd ? 1 : 0, desc->eta);
    m->nh = desc->mhd ? 9 : 6;
    m->cell_bytes = (size_t)m->g.vstride * sizeof(double);
    const size_t slot = m->cell_bytes * m->nh;
    hipError_t e = hipSuccess;
    for (int s = 0; s < 2 && e == hipSuccess; ++s) {
        e = hipMalloc(&m->W[s], slot);
        if (e == hipSuccess) e = hipMemset(m->W[s], 0, slot);
    }
    if (e == hipSuccess && desc->mhd) {
        for (int s = 0; s < 3 && e == hipSuccess; ++s) {
            e = hipMalloc(&m->bf[s], 3 * m->cell_bytes);
            if (e == hipSuccess) e = hipMemset(m->bf[s], 0, 3 * m->cell_bytes);
        }
        if (e == hipSuccess) e = hipMalloc(&m->U1, 6 * m->cell_bytes);
        if (e == hipSuccess) e = hipMalloc(&m->fdat, 9 * m->cell_bytes);
    }
    if (e == hipSuccess) e = hipMalloc(&m->dev, 4 * sizeof(double));
    if (e == hipSuccess) e = hipMemset(m->dev, 0, 4 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&m->err, 8 * sizeof(int));
    if (e == hipSuccess) e = hipMemset(m->err, 0, 8 * sizeof(int));
    if (e != hipSuccess) {
        fx_mesh_destroy(m);
        return fail(e == hipErrorOutOfMemory ? FX_ENOMEM : FX_EHIP,
                    std::string("fx_mesh_create: ") + hipGetErrorString(e));
    }
    *out = m;
    return FX_OK;
}

static void drop_spans(fx_mesh* m)
{
    for (auto& sp : m->spans) {
        (void)hipEventDestroy(sp.a);
        (void)hipEventDestroy(sp.b);
    }
    m->spans.clear();
}

int fx_mesh_destroy(fx_mesh* m)
{
    if (m == nullptr) return FX_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    else (void)hipDeviceSynchronize();
    drop_spans(m);
    for (double* p : m->W) (void)hipFree(p);
    for (double* p : m->bf) (void)hipFree(p);
    (void)hipFree(m->U1);
    (void)hipFree(m->fdat);
    (void)hipFree(m->dev);
    (void)hipFree(m->err);
    delete m;
    return FX_OK;
}

int fx_mesh_set_stream(fx_mesh* m, void* s)
{
    if (int rc = check_mesh(m)) return rc;
    m->stream = (hipStream_t)s;
    return FX_OK;
}

int fx_mesh_set_prim(fx_mesh* m, const double* W)
{
    if (int rc = check_mesh(m)) return rc;
    if (W == nullptr) return fail(FX_EINVAL, "null W");
    for (int s = 0; s < 2; ++s)
        if (int rc = copy_cells(m, m->W[s], W, nullptr, m->nh)) return rc;
    double zero[2] = {0.0, 0.0};
    FX_HIP(hipMemcpy(m->dev, zero, sizeof(zero), hipMemcpyHostToDevice));
    return FX_OK;
}

int fx_mesh_get_prim(fx_mesh* m, int slot, double* W)
{
    if (int rc = check_mesh(m)) return rc;
    if (W == nullptr || (slot != 0 && slot != 1)) return fail(FX_EINVAL, "bad slot or null W");
    return copy_cells(m, m->W[slot], nullptr, W, m->nh);
}

int fx_mesh_set_bf(fx_mesh* m, const double* bx, const double* by, const double* bz)
{
    if (int rc = check_mesh(m)) return rc;
    if (!m->g.mhd) return fail(FX_EINVAL, "mesh has no face fields");
    double* in[3] = {(double*)bx, (double*)by, (double*)bz};
    for (int s = 0; s < 2; ++s)
        if (int rc = copy_faces(m, m->bf[s], in, nullptr)) return rc;
    // cell B of every slot from the faces (prj_mhd_bf2bc_all)
    for (int s = 0; s < 2; ++s) FX_HIP(launch_bf2bc(m->g, m->W[s], m->bf[s], m->stream));
    FX_HIP(hipStreamSynchronize(m->stream));
    return FX_OK;
}

int fx_mesh_get_bf(fx_mesh* m, int slot, double* bx, double* by, double* bz)
{
    if (int rc = check_mesh(m)) return rc;
    if (!m->g.mhd) return fail(FX_EINVAL, "mesh has no face fields");
    if (slot != 0 && slot != 1) return fail(FX_EINVAL, "bad slot");
    double* out[3] = {bx, by, bz};
    return copy_faces(m, m->bf[slot], nullptr, out);
}

int fx_calc_dt(fx_mesh* m, double cfl, double* dt)
{
    if (int rc = check_mesh(m)) return rc;
    FX_HIP(launch_calc_dt(m->g, m->W[0], cfl, m->dev + 2, m->dev + 3, nullptr, 1.0, m->stream));
    double h = 0.0;
    FX_HIP(hipMemcpyAsync(&h, m->dev + 3, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    FX_HIP(hipStreamSynchronize(m->stream));
    if (dt) *dt = h;
    return FX_OK;
}

int fx_stage(fx_mesh* m, int stage, double dt)
{
    if (int rc = check_mesh(m)) return rc;
    if (stage != 1 && stage != 2) return fail(FX_EINVAL, "stage must be 1 or 2");
    return run_stage(m, stage, nullptr, dt);
}

int fx_step_async(fx_mesh* m, double cfl, double dt_factor)
{
    if (int rc = check_mesh(m)) return rc;
    FX_HIP(launch_calc_dt(m->g, m->W[0], cfl, m->dev + 2, m->dev + 0, m->dev + 1, dt_factor,
                          m->stream));
    if (int rc = run_stage(m, 1, m->dev, 0.0)) return rc;
    return run_stage(m, 2, m->dev, 0.0);
}

int fx_step(fx_mesh* m, double cfl, double dt_factor, double* dt_used)
{
    if (int rc = fx_step_async(m, cfl, dt_factor)) return rc;
    if (int rc = fx_sync(m)) return rc;
    if (dt_used) return fx_last_dt(m, dt_used);
    return FX_OK;
}

int fx_sync(fx_mesh* m)
{
    if (int rc = check_mesh(m)) return rc;
    FX_HIP(hipStreamSynchronize(m->stream));
    return read_error_word(m);
}

int fx_last_dt(fx_mesh* m, double* dt)
{
    if (int rc = check_mesh(m)) return rc;
    FX_HIP(hipMemcpyAsync(dt, m->dev, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    FX_HIP(hipStreamSynchronize(m->stream));
    return FX_OK;
}

int fx_mesh_set_timing(fx_mesh* m, int on)
{
    if (int rc = check_mesh(m)) return rc;
    FX_HIP(hipStreamSynchronize(m->stream));
    drop_spans(m);
    m->timing = on != 0;
    return FX_OK;
}

int fx_mesh_kernel_time(fx_mesh* m, int stage, double* total_ms, int* launches)
{
    if (int rc = check_mesh(m)) return rc;
    FX_HIP(hipStreamSynchronize(m->stream));
    double tot = 0.0;
    int n = 0;
    for (auto& sp : m->spans) {
        if (stage != 0 && sp.stage != stage) continue;
        float ms = 0.0f;
        FX_HIP(hipEventElapsedTime(&ms, sp.a, sp.b));
        tot += ms;
        ++n;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = n;
    return FX_OK;
}

int fx_mesh_view_get(fx_mesh* m, fx_mesh_view* out)
{
    if (int rc = check_mesh(m)) return rc;
    out->W[0] = m->W[0];
    out->W[1] = m->W[1];
    for (int s = 0; s < 3; ++s) out->bf[s] = m->bf[s];
    out->vstride = m->g.vstride;
    out->sx = m->g.sx;
    out->sy = m->g.sy;
    out->origin = m->g.origin;
    out->n[0] = m->g.nx;
    out->n[1] = m->g.ny;
    out->n[2] = m->g.nz;
    out->nh = m->nh;
    return FX_OK;
}

int fx_riemann_hllc_batch(int n, const double* WL, const double* WR, const double* pL,
                          const double* pR, const double* gL, const double* gR, const double* du,
                          const double* dv, const double* dw, double* flux, double* v_face,
                          void* stream)
{
    if (n < 0) return fail(FX_EINVAL, "negative count");
    FX_HIP(launch_hllc_batch(n, WL, WR, pL, pR, gL, gR, du, dv, dw, flux, v_face,
                             (hipStream_t)stream));
    return FX_OK;
}

int fx_riemann_hlld_batch(int n, const double* WL, const double* WR, const double* pL,
                          const double* pR, const double* gL, const double* gR, const double* bn,
                          double* flux, double* v_face, double* Bv1, double* Bv2, int* status,
                          void* stream)
{
    if (n < 0) return fail(FX_EINVAL, "negative count");
    FX_HIP(launch_hlld_batch(n, WL, WR, pL, pR, gL, gR, bn, flux, v_face, Bv1, Bv2, status,
                             (hipStream_t)stream));
    return FX_OK;
}

int fx_emf_upwind_batch(int n, const double* emf_face, const double* emf_cell,
                        const double* v_norm, double* emf, void* stream)
{
    if (n < 0) return fail(FX_EINVAL, "negative count");
    FX_HIP(launch_emf_batch(n, emf_face, emf_cell, v_norm, emf, (hipStream_t)stream));
    return FX_OK;
}

}  // extern "C"
