// C ABI over the gfx950 kernels (include/fornax_gpu.h).
//
// A fx_mesh owns the device-resident state of one uniform mesh: two stage
// slots of the cell primitives (the reference's W_mhd slots 0/1), for MHD
// three face-field buffers (slots 0/1 plus a ping-pong target for the
// stage-2 CT update), the MHD scratch, the device dt and an error word.
// All work is stream-ordered on the mesh's stream; nothing synchronises
// except the calls documented to do so.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <algorithm>
#include <vector>

#include "../../include/fornax_gpu.h"
#include "fx_api.hpp"
#include "fx_kernels.hpp"

using namespace fx;

struct fx_mesh {
    Geom g;
    int nh;
    double* W[2] = {nullptr, nullptr};
    double* bf[3] = {nullptr, nullptr, nullptr};   // [slot 0, slot 1, spare]
    double* fdat = nullptr;
    double* zring = nullptr;   // MHD: compact copies of the ghost planes z = -1, nz
    double* dev = nullptr;   // [0] dt, [1] dt_prev, [2] (unused), [3] fx_calc_dt,
                             // [4] eSSPRK substep dt; from [8] the CFL partials
    int nblocks = 0;
    // dev + 8 holds cfl_nparts CFL maxima of the current state W[0] (per
    // block for hydro, per CT workgroup for MHD), as recorded by the stage 2
    // that wrote it; anything else that writes W[0] (or the partials) clears
    // cfl_ready
    int cfl_nparts = 0;
    bool cfl_ready = false;
    int* err = nullptr;      // [code, dir, x, y, z, stage]
    int integrator = FX_RK2;
    int riemann = FX_RIEMANN_HLLC;   // hydro meshes
    int ct_split = 1;        // MHD CT workgroups per block
    // MHD meshes whose flux kernel fits the chip in one round (at most 2
    // workgroups per CU): stream and events on which the ghost-layer face
    // kernel runs beside the flux kernel (default priority, non-blocking).
    // On larger meshes the flux kernel has every CU busy and the two kernels
    // run in order (running them together measured 0.4 % slower at 256^3,
    // 2.7 % faster at 128^3).
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    bool side_on = false;    // the ghost-layer face kernel runs on `side`
    double* saved[2] = {nullptr, nullptr};   // eSSPRK(10,4) saved states u^n, u5
    double* saved_bf[2] = {nullptr, nullptr};   // and their face fields (MHD)
    hipStream_t stream = nullptr;
    size_t cell_bytes = 0;   // bytes of one padded variable
    // optional event timing of the stage kernels (excluding the ghost fill)
    bool timing = false;
    struct Span { hipEvent_t a, b; int stage; };
    std::vector<Span> spans;
};

namespace fx_api {

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
    case 31: return "prj_mhd_emf_upwind: non-finite input";
    case 32: return "prj_timeint_mhd_cell_emf_all: non-finite cell emf";
    case 33: return "prj_mhd_emf_upwind: non-finite input or result";
    case 34: return "prj_timeint_mhd_ct_value: non-finite updated magnetic field";
    case 35: return "prj_mhd_emf_upwind: resistive curl B reads outside the face storage";
    case 36: return "prj_mhd_emf_upwind: non-finite resistive curl B";
    case 37: return "prj_mhd_emf_upwind: edge index is out of storage bounds";
    case 38: return "prj_mhd_emf_upwind: null face field for resistive term";
    case 39: return "prj_mhd_emf_upwind: non-finite result";
    default: return "device error";
    }
}

}  // namespace fx_api

using namespace fx_api;

namespace {

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

// the second stream and the fork/join events of the MHD ghost-layer face kernel
hipError_t create_side_stream(fx_mesh* m)
{
    hipError_t e = hipStreamCreateWithFlags(&m->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->join, hipEventDisableTiming);
    return e;
}

int run_stage(fx_mesh* m, int stage, const double* dt_ptr, double dt)
{
    const Geom& g = m->g;
    double* Win = stage == 1 ? m->W[0] : m->W[1];
    if (!g.mhd) {
        // the hydro stage reads cells outside the domain at their sources:
        // no ghost fill
        HydroStageArgs a;
        a.g = g;
        a.Win = Win;
        a.W0 = m->W[0];
        a.Wout = stage == 1 ? m->W[1] : m->W[0];
        a.dt_ptr = dt_ptr;
        a.dt = dt;
        a.cfl_part = stage == 2 ? m->dev + 8 : nullptr;
        a.riemann = m->riemann;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (m->timing) {
            FX_HIP(hipEventCreate(&e0));
            FX_HIP(hipEventCreate(&e1));
            FX_HIP(hipEventRecord(e0, m->stream));
        }
        m->cfl_ready = false;
        FX_HIP(launch_hydro_stage(a, stage, m->stream));
        m->cfl_ready = stage == 2;
        if (m->timing) {
            FX_HIP(hipEventRecord(e1, m->stream));
            m->spans.push_back({e0, e1, stage});
        }
        return FX_OK;
    }
    // MHD: stage 1 reads face slot 0 and writes slot 1; stage 2 reads slots
    // 0 and 1 and writes the spare buffer, which then becomes slot 0.
    m->cfl_ready = false;
    double* bf_in = stage == 1 ? m->bf[0] : m->bf[1];
    FX_HIP(launch_boundary_fill(g, Win, bf_in, m->zring, m->stream));
    MhdStageArgs a;
    a.g = g;
    a.Win = Win;
    a.W0 = m->W[0];
    a.Wout = stage == 1 ? m->W[1] : m->W[0];
    a.bf_in = bf_in;
    a.bf0 = m->bf[0];
    a.bf_out = stage == 1 ? m->bf[1] : m->bf[2];
    a.fdat = m->fdat;
    a.zring = m->zring;
    a.err = m->err;
    a.dt_ptr = dt_ptr;
    a.dt = dt;
    a.ct_split = m->ct_split;
    a.cfl_part = stage == 2 ? m->dev + 8 : nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (m->timing) {
        FX_HIP(hipEventCreate(&e0));
        FX_HIP(hipEventCreate(&e1));
        FX_HIP(hipEventRecord(e0, m->stream));
    }
    FX_HIP(launch_mhd_stage(a, stage, m->stream, m->side_on ? m->side : nullptr, m->fork,
                            m->join));
    m->cfl_ready = stage == 2;
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

const char* fx_status_message(int status) { return status == 0 ? "ok" : phys_message(status); }

const char* fx_version(void)
{
    return "fornax_gpu gfx950 fp64 (ffp-contract=off)";
}

// Cell and shell indices are decomposed in 32-bit arithmetic on the device.
constexpr long long FX_MAX_CELLS = (1LL << 31) - (1LL << 26);

static bool too_many_cells(const int* n)
{
    return (long long)n[0] * n[1] * n[2] > FX_MAX_CELLS;
}

int fx_mesh_create(const fx_mesh_desc* desc, fx_mesh** out)
{
    if (desc == nullptr || out == nullptr) return fail(FX_EINVAL, "null argument");
    *out = nullptr;
    if (too_many_cells(desc->n)) return fail(FX_EINVAL, "mesh exceeds 2^31 - 2^26 cells");
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
        if (e == hipSuccess) e = hipMalloc(&m->fdat, 9 * m->cell_bytes);
        if (e == hipSuccess) e = hipMalloc(&m->zring, zring_doubles(m->g) * sizeof(double));
    }
    if (e == hipSuccess && desc->mhd) {
        int dev_id = 0, ncu = 0;
        e = hipGetDevice(&dev_id);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev_id);
        if (e == hipSuccess) m->ct_split = mhd_ct_split(m->g, ncu);
        const int64_t nb = (int64_t)m->g.nbx * m->g.nby * m->g.nbz;
        if (e == hipSuccess && nb <= 2 * (int64_t)ncu) {
            e = create_side_stream(m);
            m->side_on = e == hipSuccess;
        }
    }
    // dev[0..7] scalars, then the CFL partials: one per workgroup of the
    // reduction, or the maxima a stage 2 records
    m->nblocks = m->g.nbx * m->g.nby * m->g.nbz;
    m->cfl_nparts = m->nblocks * (desc->mhd ? m->ct_split : 1);
    // room for the partials of the largest CT split fx_mesh_set_mhd_schedule allows
    const size_t npart = (size_t)std::max(CALC_DT_WGS, m->nblocks * (desc->mhd ? 4 : 1));
    if (e == hipSuccess) e = hipMalloc(&m->dev, (8 + npart) * sizeof(double));
    if (e == hipSuccess) e = hipMemset(m->dev, 0, 8 * sizeof(double));
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
    if (m->side) (void)hipStreamSynchronize(m->side);
    drop_spans(m);
    for (double* p : m->W) (void)hipFree(p);
    for (double* p : m->bf) (void)hipFree(p);
    (void)hipFree(m->fdat);
    (void)hipFree(m->zring);
    (void)hipFree(m->dev);
    for (int s = 0; s < 2; ++s) {
        (void)hipFree(m->saved[s]);
        (void)hipFree(m->saved_bf[s]);
    }
    (void)hipFree(m->err);
    if (m->fork) (void)hipEventDestroy(m->fork);
    if (m->join) (void)hipEventDestroy(m->join);
    if (m->side) (void)hipStreamDestroy(m->side);
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
    m->cfl_ready = false;
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
    m->cfl_ready = false;
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
    // the integrator's scale of the CFL minimum (src/prj_timeint.c:2248-2253)
    const double scale = m->integrator == FX_ESSPRK10_4 ? 6.0 : 1.0;
    if (m->cfl_ready)
        FX_HIP(launch_dt_final(m->dev + 8, m->cfl_nparts, cfl, m->dev + 3, nullptr, 1.0, scale,
                               m->stream));
    else
        FX_HIP(launch_calc_dt(m->g, m->W[0], cfl, m->dev + 8, m->dev + 3, nullptr, 1.0, m->stream,
                              scale));
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

int fx_mesh_set_integrator(fx_mesh* m, int integrator)
{
    if (int rc = check_mesh(m)) return rc;
    if (integrator != FX_RK2 && integrator != FX_ESSPRK10_4)
        return fail(FX_EINVAL, "unknown integrator");
    if (integrator == FX_ESSPRK10_4) {
        FX_HIP(hipStreamSynchronize(m->stream));
        for (int s = 0; s < 2; ++s) {
            hipError_t e = hipSuccess;
            if (m->saved[s] == nullptr) e = hipMalloc(&m->saved[s], m->cell_bytes * m->nh);
            if (e == hipSuccess && m->g.mhd && m->saved_bf[s] == nullptr)
                e = hipMalloc(&m->saved_bf[s], 3 * m->cell_bytes);
            if (e != hipSuccess)
                return fail(e == hipErrorOutOfMemory ? FX_ENOMEM : FX_EHIP,
                            std::string("fx_mesh_set_integrator: ") + hipGetErrorString(e));
        }
    }
    m->integrator = integrator;
    return FX_OK;
}

int fx_mesh_set_mhd_schedule(fx_mesh* m, int ct_split, int side_stream)
{
    if (int rc = check_mesh(m)) return rc;
    if (!m->g.mhd) return fail(FX_EINVAL, "fx_mesh_set_mhd_schedule: not an MHD mesh");
    if (ct_split != 0 && ct_split != 1 && ct_split != 2 && ct_split != 4)
        return fail(FX_EINVAL, "ct_split must be 0 (automatic), 1, 2 or 4");
    if (side_stream < -1 || side_stream > 1)
        return fail(FX_EINVAL, "side_stream must be -1 (automatic), 0 or 1");
    FX_HIP(hipStreamSynchronize(m->stream));
    if (m->side) FX_HIP(hipStreamSynchronize(m->side));
    int dev_id = 0, ncu = 0;
    FX_HIP(hipGetDevice(&dev_id));
    FX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev_id));
    m->ct_split = ct_split != 0 ? ct_split : mhd_ct_split(m->g, ncu);
    m->cfl_nparts = m->nblocks * m->ct_split;
    m->cfl_ready = false;   // recorded per CT workgroup of the old split
    bool on = side_stream == 1 || (side_stream == -1 && m->nblocks <= 2 * ncu);
    if (on && m->side == nullptr) FX_HIP(create_side_stream(m));
    m->side_on = on;
    return FX_OK;
}

int fx_mesh_set_riemann(fx_mesh* m, int solver)
{
    if (int rc = check_mesh(m)) return rc;
    if (solver != FX_RIEMANN_HLLC && solver != FX_RIEMANN_HLLE)
        return fail(FX_EINVAL, "unknown Riemann solver");
    if (m->g.mhd && solver != FX_RIEMANN_HLLC)
        return fail(FX_EINVAL, "MHD meshes take HLLD fluxes only");
    FX_HIP(hipStreamSynchronize(m->stream));
    m->riemann = solver;
    m->cfl_ready = false;
    return FX_OK;
}

namespace {

// One forward-Euler substep of eSSPRK (src/prj_timeint.c:2470-2578): the
// stage-1 form with the device substep dt, slot 0 -> slot 1 (cells, and the
// face fields of MHD: the reference's in-place stage-0 CT update of slot 0
// forms the same values), then the slots swap so that the state stays in
// slot 0.
int essprk_substep(fx_mesh* m)
{
    if (int rc = run_stage(m, 1, m->dev + 4, 0.0)) return rc;
    std::swap(m->W[0], m->W[1]);
    if (m->g.mhd) std::swap(m->bf[0], m->bf[1]);
    return FX_OK;
}

// slot 0 (cells, and faces on MHD meshes) into saved state s
// (prj_timeint_eSSPRK_save_state, src/prj_timeint.c:2310-2345)
int essprk_save(fx_mesh* m, int s)
{
    FX_HIP(hipMemcpyAsync(m->saved[s], m->W[0], m->cell_bytes * m->nh, hipMemcpyDeviceToDevice,
                          m->stream));
    if (m->g.mhd)
        FX_HIP(hipMemcpyAsync(m->saved_bf[s], m->bf[0], 3 * m->cell_bytes,
                              hipMemcpyDeviceToDevice, m->stream));
    return FX_OK;
}

// slot 0 <- w * saved state s + (1 - w) * slot 0
// (prj_timeint_eSSPRK_blend_with_saved, src/prj_timeint.c:2382-2417)
int essprk_blend(fx_mesh* m, int s, double w)
{
    if (m->g.mhd)
        FX_HIP(launch_blend_mhd(m->g, m->W[0], m->saved[s], m->bf[0], m->saved_bf[s], w,
                                m->stream));
    else
        FX_HIP(launch_blend(m->g, m->W[0], m->saved[s], w, m->stream));
    return FX_OK;
}

// Ketcheson eSSPRK(10,4), low-storage form (src/prj_timeint.c:2661-2730)
int essprk10_4_step(fx_mesh* m)
{
    FX_HIP(launch_div_scalar(m->dev + 0, m->dev + 4, 6.0, m->stream));     // dt / 6
    if (int rc = essprk_save(m, 0)) return rc;
    for (int n = 0; n < 5; ++n)
        if (int rc = essprk_substep(m)) return rc;
    if (int rc = essprk_save(m, 1)) return rc;
    if (int rc = essprk_blend(m, 0, 3.0 / 5.0)) return rc;
    for (int n = 0; n < 5; ++n)
        if (int rc = essprk_substep(m)) return rc;
    if (int rc = essprk_blend(m, 1, 3.0 / 8.0)) return rc;
    return essprk_blend(m, 0, 1.0 / 25.0);
}

}  // namespace

int fx_step_async(fx_mesh* m, double cfl, double dt_factor)
{
    if (int rc = check_mesh(m)) return rc;
    const bool ssp = m->integrator == FX_ESSPRK10_4;
    // the CFL maximum of W[0] comes from the stage 2 that wrote it when that
    // was recorded, else from a reduction over the cells
    const double scale = ssp ? 6.0 : 1.0;
    if (m->cfl_ready)
        FX_HIP(launch_dt_final(m->dev + 8, m->cfl_nparts, cfl, m->dev + 0, m->dev + 1, dt_factor,
                               scale, m->stream));
    else
        FX_HIP(launch_calc_dt(m->g, m->W[0], cfl, m->dev + 8, m->dev + 0, m->dev + 1, dt_factor,
                              m->stream, scale));
    if (ssp) {
        m->cfl_ready = false;
        return essprk10_4_step(m);
    }
    if (int rc = run_stage(m, 1, m->dev, 0.0)) return rc;
    return run_stage(m, 2, m->dev, 0.0);
}

int fx_step_dt(fx_mesh* m, double dt)
{
    if (int rc = check_mesh(m)) return rc;
    if (!(dt > 0.0) || !std::isfinite(dt))
        return fail(FX_EINVAL, "fx_step_dt: dt must be finite and > 0");
    // dt and the growth limit's previous dt of a later fx_step
    const double h[2] = {dt, dt};
    FX_HIP(hipMemcpyAsync(m->dev, h, sizeof(h), hipMemcpyHostToDevice, m->stream));
    if (m->integrator == FX_ESSPRK10_4) {
        m->cfl_ready = false;
        if (int rc = essprk10_4_step(m)) return rc;
    } else {
        if (int rc = run_stage(m, 1, m->dev, 0.0)) return rc;
        if (int rc = run_stage(m, 2, m->dev, 0.0)) return rc;
    }
    return fx_sync(m);
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
    m->cfl_ready = false;   // the caller may write the state through the view
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

int fx_reconstruct_mc_slope_batch(int n, const double* qm, const double* q0, const double* qp,
                                  double* slope, void* stream)
{
    if (n < 0 || (n > 0 && (qm == nullptr || q0 == nullptr || qp == nullptr || slope == nullptr)))
        return fail(FX_EINVAL, "bad mc slope arguments");
    FX_HIP(launch_mc_slope_batch(n, qm, q0, qp, slope, (hipStream_t)stream));
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

int fx_emf_upwind_block_batch(int n, const int* edge, const double* emf_face,
                              const double* emf_cell, const double* v_norm, const double* eta,
                              const double* const* bf, const double* dx, double* const* block_emf,
                              double* emf, int* status, void* stream)
{
    if (n < 0) return fail(FX_EINVAL, "negative count");
    if (n == 0) return FX_OK;
    if (edge == nullptr || emf_face == nullptr || emf_cell == nullptr || v_norm == nullptr ||
        eta == nullptr || dx == nullptr || emf == nullptr || status == nullptr)
        return fail(FX_EINVAL, "fx_emf_upwind_block_batch: null array");
    EmfBlockArgs a{};
    a.n = n;
    a.edge = edge;
    a.ef = emf_face;
    a.ec = emf_cell;
    a.vn = v_norm;
    a.eta = eta;
    for (int d = 0; d < 3; ++d) {
        a.bf[d] = bf != nullptr ? bf[d] : nullptr;
        a.block_emf[d] = block_emf != nullptr ? block_emf[d] : nullptr;
        a.dx[d] = dx[d];
    }
    a.emf = emf;
    a.status = status;
    FX_HIP(launch_emf_block_batch(a, (hipStream_t)stream));
    return FX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Ragged ensembles of independent hydro meshes.
struct fx_ensemble {
    int n = 0;
    std::vector<Geom> geoms;
    std::vector<EnsInst> host_inst;
    std::vector<EnsBlock> order[2];      // instance-major and XCD-interleaved tables
    EnsInst* d_inst = nullptr;
    EnsBlock* d_blocks[2] = {nullptr, nullptr};
    int* d_queues = nullptr;
    double* dev = nullptr;               // 4 doubles per instance
    double* partial = nullptr;           // per-block CFL maxima of the per-instance schedule
    std::vector<int> part_off;           // first partial of each instance
    int nblocks = 0;
    // the CFL maxima of the current states are recorded (dev[2] of every
    // instance in the batched schedule, `partial` in the per-instance one)
    bool cfl_ready = false;
    hipStream_t stream = nullptr;        // owned: graph capture needs a non-null stream
    // schedule: mode 0 = one launch set per instance, 1 = batched
    int mode = 1;
    int persistent_wgs = 0;              // 0 = one workgroup per block
    int xcd = 1;                         // remap / per-XCD queues
    int table = 0;                       // 0 instance-major, 1 interleaved
    // hipGraph of one step (batched mode)
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double graph_cfl = -1.0, graph_factor = -1.0;
};

namespace {

void ens_drop_graph(fx_ensemble* e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

int ens_nblocks(const Geom& g) { return g.nbx * g.nby * g.nbz; }

// CFL maxima of the current states, read from the cells, where the step
// expects them (a stage 2 records them for the step after it)
int ens_enqueue_cfl(fx_ensemble* e)
{
    hipStream_t s = e->stream;
    if (e->mode == 0) {
        for (int i = 0; i < e->n; ++i) {
            const EnsInst& I = e->host_inst[i];
            FX_HIP(launch_dt_reduce(I.g, I.W[0], e->partial + e->part_off[i], ens_nblocks(I.g), s));
        }
        return FX_OK;
    }
    FX_HIP(launch_ens_dt_reduce(e->d_inst, e->n, e->d_blocks[e->table], e->nblocks, s));
    return FX_OK;
}

// One RK2 step of every instance; the CFL maxima of the states are recorded.
int ens_enqueue_step(fx_ensemble* e, double cfl, double factor)
{
    hipStream_t s = e->stream;
    if (e->mode == 0) {
        for (int i = 0; i < e->n; ++i) {
            const EnsInst& I = e->host_inst[i];
            double* part = e->partial + e->part_off[i];
            FX_HIP(launch_dt_final(part, ens_nblocks(I.g), cfl, I.dev + 0, I.dev + 1, factor, 1.0, s));
            for (int st = 1; st <= 2; ++st) {
                double* Win = st == 1 ? I.W[0] : I.W[1];
                HydroStageArgs a;
                a.g = I.g;
                a.Win = Win;
                a.W0 = I.W[0];
                a.Wout = st == 1 ? I.W[1] : I.W[0];
                a.dt_ptr = I.dev;
                a.dt = 0.0;
                a.cfl_part = st == 2 ? part : nullptr;
                FX_HIP(launch_hydro_stage(a, st, s));
            }
        }
        return FX_OK;
    }
    const EnsBlock* tbl = e->d_blocks[e->table];
    FX_HIP(launch_ens_dt_finalize(e->d_inst, e->n, cfl, factor, s));
    for (int st = 1; st <= 2; ++st) {
        if (st == 2) FX_HIP(launch_ens_dt_clear(e->d_inst, e->n, s));
        FX_HIP(launch_ens_stage(e->d_inst, tbl, e->nblocks, st, e->persistent_wgs, e->xcd, e->xcd,
                                e->d_queues, s));
    }
    return FX_OK;
}

}  // namespace

extern "C" {

int fx_ensemble_destroy(fx_ensemble* e)
{
    if (e == nullptr) return FX_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    ens_drop_graph(e);
    for (auto& I : e->host_inst) {
        (void)hipFree(I.W[0]);
        (void)hipFree(I.W[1]);
    }
    (void)hipFree(e->d_inst);
    (void)hipFree(e->d_blocks[0]);
    (void)hipFree(e->d_blocks[1]);
    (void)hipFree(e->d_queues);
    (void)hipFree(e->dev);
    (void)hipFree(e->partial);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return FX_OK;
}

int fx_ensemble_create(int n, const fx_mesh_desc* descs, fx_ensemble** out)
{
    if (n <= 0 || descs == nullptr || out == nullptr) return fail(FX_EINVAL, "bad ensemble arguments");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        const fx_mesh_desc& d = descs[i];
        if (d.mhd) return fail(FX_EINVAL, "ensembles are hydro only");
        for (int a = 0; a < 3; ++a) {
            if (d.n[a] <= 0 || d.n[a] % BS != 0)
                return fail(FX_EINVAL, "grid extent must be a positive multiple of 16");
            bool lo = d.bc[2 * a] == 3, hi = d.bc[2 * a + 1] == 3;
            if (lo != hi) return fail(FX_EINVAL, "periodic boundaries must be paired");
            if (lo && d.n[a] < 2 * BS) return fail(FX_EINVAL, "a periodic axis needs at least two blocks");
        }
        if (!(d.gamma > 1.0)) return fail(FX_EINVAL, "gamma must exceed 1");
        if (too_many_cells(d.n)) return fail(FX_EINVAL, "instance exceeds 2^31 - 2^26 cells");
    }
    fx_ensemble* e = new fx_ensemble();
    e->n = n;
    hipError_t err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipMalloc(&e->dev, sizeof(double) * 4 * n);
    if (err == hipSuccess) err = hipMemset(e->dev, 0, sizeof(double) * 4 * n);
    for (int i = 0; i < n && err == hipSuccess; ++i) {
        const fx_mesh_desc& d = descs[i];
        Geom g = make_geom(d.n[0], d.n[1], d.n[2], d.xlim, d.bc, d.gamma, 0, 0.0);
        EnsInst I{};
        I.g = g;
        const size_t bytes = (size_t)g.vstride * 6 * sizeof(double);
        err = hipMalloc(&I.W[0], bytes);
        if (err == hipSuccess) err = hipMemset(I.W[0], 0, bytes);
        if (err == hipSuccess) err = hipMalloc(&I.W[1], bytes);
        if (err == hipSuccess) err = hipMemset(I.W[1], 0, bytes);
        I.dev = e->dev + 4 * i;
        e->geoms.push_back(g);
        e->host_inst.push_back(I);
        for (int bx = 0; bx < g.nbx; ++bx)
            for (int by = 0; by < g.nby; ++by)
                for (int bz = 0; bz < g.nbz; ++bz) e->order[0].push_back({i, bx, by, bz});
    }
    e->nblocks = (int)e->order[0].size();
    for (int i = 0, acc = 0; i < n; ++i) {
        e->part_off.push_back(acc);
        acc += ens_nblocks(e->geoms[i]);
    }
    if (err == hipSuccess) err = hipMalloc(&e->partial, (size_t)std::max(1, e->nblocks) * sizeof(double));
    // interleaved table: instances dealt round robin block by block, so that
    // every part of the table mixes all instances
    {
        std::vector<size_t> next(n, 0), start(n, 0);
        size_t acc = 0;
        for (int i = 0; i < n; ++i) {
            start[i] = acc;
            acc += (size_t)e->geoms[i].nbx * e->geoms[i].nby * e->geoms[i].nbz;
        }
        bool left = true;
        while (left) {
            left = false;
            for (int i = 0; i < n; ++i) {
                size_t cnt = (size_t)e->geoms[i].nbx * e->geoms[i].nby * e->geoms[i].nbz;
                if (next[i] < cnt) {
                    e->order[1].push_back(e->order[0][start[i] + next[i]]);
                    ++next[i];
                    left = true;
                }
            }
        }
    }
    if (err == hipSuccess) err = hipMalloc(&e->d_inst, sizeof(EnsInst) * n);
    if (err == hipSuccess)
        err = hipMemcpy(e->d_inst, e->host_inst.data(), sizeof(EnsInst) * n, hipMemcpyHostToDevice);
    for (int t = 0; t < 2 && err == hipSuccess; ++t) {
        err = hipMalloc(&e->d_blocks[t], sizeof(EnsBlock) * e->nblocks);
        if (err == hipSuccess)
            err = hipMemcpy(e->d_blocks[t], e->order[t].data(), sizeof(EnsBlock) * e->nblocks,
                            hipMemcpyHostToDevice);
    }
    if (err == hipSuccess) err = hipMalloc(&e->d_queues, 64 * sizeof(int));
    if (err == hipSuccess) err = hipMemset(e->d_queues, 0, 64 * sizeof(int));
    if (err != hipSuccess) {
        fx_ensemble_destroy(e);
        return fail(err == hipErrorOutOfMemory ? FX_ENOMEM : FX_EHIP,
                    std::string("fx_ensemble_create: ") + hipGetErrorString(err));
    }
    *out = e;
    return FX_OK;
}

static int ens_check(fx_ensemble* e, int i)
{
    if (e == nullptr) return fail(FX_EINVAL, "null ensemble");
    if (i < 0 || i >= e->n) return fail(FX_EINVAL, "instance index out of range");
    return FX_OK;
}

int fx_ensemble_set_prim(fx_ensemble* e, int i, const double* W)
{
    if (int rc = ens_check(e, i)) return rc;
    e->cfl_ready = false;
    fx_mesh tmp;
    tmp.g = e->geoms[i];
    tmp.stream = e->stream;
    for (int s = 0; s < 2; ++s)
        if (int rc = copy_cells(&tmp, e->host_inst[i].W[s], W, nullptr, 6)) return rc;
    double zero[4] = {0.0, 0.0, 0.0, 0.0};
    FX_HIP(hipMemcpyAsync(e->host_inst[i].dev, zero, sizeof(zero), hipMemcpyHostToDevice, e->stream));
    FX_HIP(hipStreamSynchronize(e->stream));
    return FX_OK;
}

int fx_ensemble_get_prim(fx_ensemble* e, int i, int slot, double* W)
{
    if (int rc = ens_check(e, i)) return rc;
    if (slot != 0 && slot != 1) return fail(FX_EINVAL, "bad slot");
    fx_mesh tmp;
    tmp.g = e->geoms[i];
    tmp.stream = e->stream;
    return copy_cells(&tmp, e->host_inst[i].W[slot], nullptr, W, 6);
}

int fx_ensemble_set_schedule(fx_ensemble* e, int mode, int persistent_wgs, int xcd, int table)
{
    if (e == nullptr) return fail(FX_EINVAL, "null ensemble");
    if (mode < 0 || mode > 1 || persistent_wgs < 0 || table < 0 || table > 1)
        return fail(FX_EINVAL, "bad schedule");
    FX_HIP(hipStreamSynchronize(e->stream));
    e->mode = mode;
    e->persistent_wgs = persistent_wgs;
    e->xcd = xcd ? 1 : 0;
    e->table = table;
    e->cfl_ready = false;   // the schedules record the CFL maxima in different places
    ens_drop_graph(e);
    return FX_OK;
}

int fx_ensemble_set_graph(fx_ensemble* e, int on)
{
    if (e == nullptr) return fail(FX_EINVAL, "null ensemble");
    FX_HIP(hipStreamSynchronize(e->stream));
    e->use_graph = on != 0;
    ens_drop_graph(e);
    return FX_OK;
}

static int ens_step(fx_ensemble* e, double cfl, double dt_factor)
{
    if (!e->cfl_ready) {
        if (int rc = ens_enqueue_cfl(e)) return rc;
        e->cfl_ready = true;
    }
    if (!e->use_graph) return ens_enqueue_step(e, cfl, dt_factor);
    if (e->exec == nullptr || e->graph_cfl != cfl || e->graph_factor != dt_factor) {
        ens_drop_graph(e);
        FX_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int rc = ens_enqueue_step(e, cfl, dt_factor);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(e->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        FX_HIP(ce);
        e->graph = g;
        FX_HIP(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
        e->graph_cfl = cfl;
        e->graph_factor = dt_factor;
    }
    FX_HIP(hipGraphLaunch(e->exec, e->stream));
    return FX_OK;
}

int fx_ensemble_step_async(fx_ensemble* e, double cfl, double dt_factor)
{
    if (e == nullptr) return fail(FX_EINVAL, "null ensemble");
    const int rc = ens_step(e, cfl, dt_factor);
    // a step that failed part-way may have cleared the maxima it was to raise
    if (rc) e->cfl_ready = false;
    return rc;
}

int fx_ensemble_sync(fx_ensemble* e)
{
    if (e == nullptr) return fail(FX_EINVAL, "null ensemble");
    FX_HIP(hipStreamSynchronize(e->stream));
    return FX_OK;
}

int fx_ensemble_last_dt(fx_ensemble* e, double* dt)
{
    if (e == nullptr || dt == nullptr) return fail(FX_EINVAL, "null argument");
    std::vector<double> h(4 * (size_t)e->n);
    FX_HIP(hipMemcpyAsync(h.data(), e->dev, sizeof(double) * h.size(), hipMemcpyDeviceToHost, e->stream));
    FX_HIP(hipStreamSynchronize(e->stream));
    for (int i = 0; i < e->n; ++i) dt[i] = h[4 * (size_t)i];
    return FX_OK;
}

int fx_ensemble_info(fx_ensemble* e, int* nblocks, long long* ncells)
{
    if (e == nullptr) return fail(FX_EINVAL, "null ensemble");
    if (nblocks) *nblocks = e->nblocks;
    if (ncells) *ncells = (long long)e->nblocks * BS * BS * BS;
    return FX_OK;
}

// Time `reps` steps of every candidate schedule on the current state (the
// state advances; callers tune on a copy or re-set it) and keep the fastest.
// out[0..3] = mode, persistent_wgs, xcd, table; *best_ms = ms per step.
int fx_ensemble_autotune(fx_ensemble* e, double cfl, double dt_factor, int reps, int* out,
                         double* best_ms)
{
    if (e == nullptr || reps <= 0) return fail(FX_EINVAL, "bad autotune arguments");
    int dev_id = 0;
    FX_HIP(hipGetDevice(&dev_id));
    int ncu = 0;
    FX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev_id));
    const int cand[][4] = {
        {1, 0, 1, 0}, {1, 0, 0, 0}, {1, 0, 1, 1},
        {1, ncu, 1, 0}, {1, 2 * ncu, 1, 0}, {1, 2 * ncu, 0, 1},
    };
    const bool graph = e->use_graph;
    double best = 1e300;
    int bi = 0;
    struct Events {
        hipEvent_t a = nullptr, b = nullptr;
        ~Events()
        {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev;
    FX_HIP(hipEventCreate(&ev.a));
    FX_HIP(hipEventCreate(&ev.b));
    hipEvent_t a = ev.a, b = ev.b;
    for (int c = 0; c < (int)(sizeof(cand) / sizeof(cand[0])); ++c) {
        if (int rc = fx_ensemble_set_schedule(e, cand[c][0], cand[c][1], cand[c][2], cand[c][3])) return rc;
        e->use_graph = graph;
        if (int rc = fx_ensemble_step_async(e, cfl, dt_factor)) return rc;   // warm-up / capture
        FX_HIP(hipEventRecord(a, e->stream));
        for (int r = 0; r < reps; ++r)
            if (int rc = fx_ensemble_step_async(e, cfl, dt_factor)) return rc;
        FX_HIP(hipEventRecord(b, e->stream));
        FX_HIP(hipEventSynchronize(b));
        float ms = 0.0f;
        FX_HIP(hipEventElapsedTime(&ms, a, b));
        if (ms < best) {
            best = ms;
            bi = c;
        }
    }
    if (int rc = fx_ensemble_set_schedule(e, cand[bi][0], cand[bi][1], cand[bi][2], cand[bi][3])) return rc;
    e->use_graph = graph;
    if (out) for (int q = 0; q < 4; ++q) out[q] = cand[bi][q];
    if (best_ms) *best_ms = best / reps;
    return FX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Differentiable scheme: forward with checkpoints, reverse sweep with recompute.
struct fx_adj {
    AdjGeom g;
    int every = 8;                        // checkpoint interval (steps)
    size_t state_bytes = 0;               // 6 * N doubles
    double* S = nullptr;                  // running state; W_T after a forward
    double* W1 = nullptr;                 // stage-1 output of the plain forward
    double* F = nullptr;                  // 18 N
    double* G = nullptr;                  // 36 N
    double* Ub[2] = {nullptr, nullptr};   // 6 N each
    double* seg = nullptr;                // 2 * every states of one recomputed segment
    double* ckpt = nullptr;               // checkpoints of the last forward
    int ckpt_cap = 0;
    int nsteps = -1;                      // of the last forward (-1: none)
    double dt = 0.0;
};

namespace {

inline double* adj_state(double* base, const fx_adj* a, int i)
{
    return base + (size_t)i * 6 * (size_t)a->g.N;
}

}  // namespace

extern "C" {

int fx_adj_destroy(fx_adj* a)
{
    if (a == nullptr) return FX_OK;
    (void)hipDeviceSynchronize();
    (void)hipFree(a->S);
    (void)hipFree(a->W1);
    (void)hipFree(a->F);
    (void)hipFree(a->G);
    (void)hipFree(a->Ub[0]);
    (void)hipFree(a->Ub[1]);
    (void)hipFree(a->seg);
    (void)hipFree(a->ckpt);
    delete a;
    return FX_OK;
}

int fx_adj_create(const int* n, const double* xlim, double gamma, double va_eps2,
                  int checkpoint_every, fx_adj** out)
{
    if (n == nullptr || xlim == nullptr || out == nullptr) return fail(FX_EINVAL, "null argument");
    *out = nullptr;
    for (int d = 0; d < 3; ++d)
        if (n[d] < 4) return fail(FX_EINVAL, "the periodic box needs at least 4 cells per axis");
    if (!(gamma > 1.0)) return fail(FX_EINVAL, "gamma must exceed 1");
    if (!(va_eps2 > 0.0)) return fail(FX_EINVAL, "van Albada eps^2 must be positive");
    if (checkpoint_every < 1) return fail(FX_EINVAL, "checkpoint interval must be >= 1");
    fx_adj* a = new fx_adj();
    AdjGeom& g = a->g;
    for (int d = 0; d < 3; ++d) {
        g.n[d] = n[d];
        g.dx[d] = (xlim[2 * d + 1] - xlim[2 * d]) / (double)n[d];
    }
    g.N = (int64_t)n[0] * n[1] * n[2];
    g.vol = g.dx[0] * g.dx[1] * g.dx[2];
    g.area[0] = g.dx[1] * g.dx[2];
    g.area[1] = g.dx[0] * g.dx[2];
    g.area[2] = g.dx[0] * g.dx[1];
    g.gamma = gamma;
    g.eps2 = va_eps2;
    a->every = checkpoint_every;
    a->state_bytes = sizeof(double) * 6 * (size_t)g.N;
    hipError_t e = hipMalloc(&a->S, a->state_bytes);
    if (e == hipSuccess) e = hipMalloc(&a->W1, a->state_bytes);
    if (e == hipSuccess) e = hipMalloc(&a->F, 3 * a->state_bytes);
    if (e == hipSuccess) e = hipMalloc(&a->G, 6 * a->state_bytes);
    if (e == hipSuccess) e = hipMalloc(&a->Ub[0], a->state_bytes);
    if (e == hipSuccess) e = hipMalloc(&a->Ub[1], a->state_bytes);
    if (e == hipSuccess) e = hipMalloc(&a->seg, 2 * (size_t)checkpoint_every * a->state_bytes);
    if (e != hipSuccess) {
        fx_adj_destroy(a);
        return fail(e == hipErrorOutOfMemory ? FX_ENOMEM : FX_EHIP,
                    std::string("fx_adj_create: ") + hipGetErrorString(e));
    }
    *out = a;
    return FX_OK;
}

int fx_adj_forward(fx_adj* a, const double* W0, double* WT, int nsteps, double dt, void* stream)
{
    if (a == nullptr || W0 == nullptr || WT == nullptr) return fail(FX_EINVAL, "null argument");
    if (nsteps < 0) return fail(FX_EINVAL, "negative step count");
    hipStream_t s = (hipStream_t)stream;
    const int nck = nsteps / a->every + 1;
    if (nck > a->ckpt_cap) {
        FX_HIP(hipStreamSynchronize(s));
        (void)hipFree(a->ckpt);
        a->ckpt = nullptr;
        a->ckpt_cap = 0;
        FX_HIP(hipMalloc(&a->ckpt, (size_t)nck * a->state_bytes));
        a->ckpt_cap = nck;
    }
    a->nsteps = -1;
    FX_HIP(hipMemcpyAsync(a->S, W0, a->state_bytes, hipMemcpyDeviceToDevice, s));
    for (int k = 0; k < nsteps; ++k) {
        if (k % a->every == 0)
            FX_HIP(hipMemcpyAsync(adj_state(a->ckpt, a, k / a->every), a->S, a->state_bytes,
                                  hipMemcpyDeviceToDevice, s));
        FX_HIP(launch_adj_stage(a->g, 1, dt, a->S, a->S, a->F, a->W1, s));
        FX_HIP(launch_adj_stage(a->g, 2, dt, a->W1, a->S, a->F, a->S, s));
    }
    FX_HIP(hipMemcpyAsync(WT, a->S, a->state_bytes, hipMemcpyDeviceToDevice, s));
    a->nsteps = nsteps;
    a->dt = dt;
    return FX_OK;
}

int fx_adj_backward(fx_adj* a, const double* gWT, double* gW0, void* stream)
{
    if (a == nullptr || gWT == nullptr || gW0 == nullptr) return fail(FX_EINVAL, "null argument");
    if (a->nsteps < 0) return fail(FX_EINVAL, "fx_adj_backward needs a preceding fx_adj_forward");
    hipStream_t s = (hipStream_t)stream;
    const AdjGeom& g = a->g;
    const int n = a->nsteps, K = a->every;
    double* Ub2 = a->Ub[0];
    double* Ub1 = a->Ub[1];
    FX_HIP(launch_adj_w2u(g, a->S, gWT, Ub2, s));
    for (int sg = (n - 1) / K; n > 0 && sg >= 0; --sg) {
        const int k0 = sg * K, k1 = std::min(n, k0 + K), m = k1 - k0;
        // recompute the segment's stage inputs: seg[2k] = W^n, seg[2k+1] = W^(1)
        FX_HIP(hipMemcpyAsync(adj_state(a->seg, a, 0), adj_state(a->ckpt, a, sg), a->state_bytes,
                              hipMemcpyDeviceToDevice, s));
        for (int k = 0; k < m; ++k) {
            FX_HIP(launch_adj_stage(g, 1, a->dt, adj_state(a->seg, a, 2 * k),
                                    adj_state(a->seg, a, 2 * k), a->F,
                                    adj_state(a->seg, a, 2 * k + 1), s));
            if (k + 1 < m)
                FX_HIP(launch_adj_stage(g, 2, a->dt, adj_state(a->seg, a, 2 * k + 1),
                                        adj_state(a->seg, a, 2 * k), a->F,
                                        adj_state(a->seg, a, 2 * k + 2), s));
        }
        for (int k = m - 1; k >= 0; --k)
            FX_HIP(launch_adj_step_reverse(g, a->dt, adj_state(a->seg, a, 2 * k),
                                           adj_state(a->seg, a, 2 * k + 1), Ub2, Ub1, a->G, s));
    }
    // initial state: W0 is checkpoint 0 (or S itself when nsteps == 0)
    const double* W0 = n > 0 ? adj_state(a->ckpt, a, 0) : a->S;
    FX_HIP(launch_adj_u2w(g, W0, Ub2, gW0, s));
    return FX_OK;
}

}  // extern "C"


struct fx_cf_plan {
    int nblocks = 0;
    int nvar = 0;
    void* mem = nullptr;            // one device allocation: table, records, block list
    CfPtrs* tab = nullptr;
    CfFlux* flux = nullptr;
    CfEmf* emf = nullptr;
    int* fid_blocks = nullptr;
    int nfid = 0;
    std::vector<int> flux_off, emf_off;
    std::string flux_err, emf_err;
};

extern "C" {

int fx_cf_plan_create(int nblocks, const fx_cf_block* blocks, int nlinks, const fx_cf_link* links,
                      int nvar, fx_cf_plan** out)
{
    if (out == nullptr) return fail(FX_EINVAL, "null plan pointer");
    *out = nullptr;
    if (nblocks <= 0 || blocks == nullptr) return fail(FX_EINVAL, "empty block table");
    if (nlinks < 0 || (nlinks > 0 && links == nullptr)) return fail(FX_EINVAL, "bad link array");
    if (nvar < 1 || nvar > 16) return fail(FX_EINVAL, "nvar must be 1..16");
    CfHostPlan hp;
    cf_build_plan(nblocks, blocks, nlinks, links, nvar, hp);
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t b_tab = align(sizeof(CfPtrs) * (size_t)nblocks);
    const size_t b_flux = align(sizeof(CfFlux) * hp.flux.size());
    const size_t b_emf = align(sizeof(CfEmf) * hp.emf.size());
    const size_t b_fid = align(sizeof(int) * hp.fid_blocks.size());
    fx_cf_plan* p = new fx_cf_plan();
    p->nblocks = nblocks;
    p->nvar = nvar;
    hipError_t e = hipMalloc(&p->mem, b_tab + b_flux + b_emf + b_fid);
    if (e != hipSuccess) {
        delete p;
        return fail(e == hipErrorOutOfMemory ? FX_ENOMEM : FX_EHIP,
                    std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    char* base = (char*)p->mem;
    p->tab = (CfPtrs*)base;
    p->flux = (CfFlux*)(base + b_tab);
    p->emf = (CfEmf*)(base + b_tab + b_flux);
    p->fid_blocks = (int*)(base + b_tab + b_flux + b_emf);
    std::vector<CfPtrs> tab(nblocks);
    for (int b = 0; b < nblocks; ++b)
        for (int d = 0; d < 3; ++d) {
            tab[b].flux[d] = blocks[b].flux[d];
            tab[b].emf[d] = blocks[b].emf[d];
            tab[b].fid[d] = blocks[b].edge_fidelity[d];
        }
    struct Copy { void* dst; const void* src; size_t n; };
    const Copy copies[] = {{p->tab, tab.data(), sizeof(CfPtrs) * tab.size()},
                           {p->flux, hp.flux.data(), sizeof(CfFlux) * hp.flux.size()},
                           {p->emf, hp.emf.data(), sizeof(CfEmf) * hp.emf.size()},
                           {p->fid_blocks, hp.fid_blocks.data(), sizeof(int) * hp.fid_blocks.size()}};
    for (const Copy& c : copies) {
        if (c.n == 0) continue;
        e = hipMemcpy(c.dst, c.src, c.n, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(p->mem);
            delete p;
            return hip_fail(e, "hipMemcpy");
        }
    }
    p->nfid = (int)hp.fid_blocks.size();
    p->flux_off = hp.flux_off;
    p->emf_off = hp.emf_off;
    p->flux_err = hp.flux_err;
    p->emf_err = hp.emf_err;
    *out = p;
    return FX_OK;
}

int fx_cf_plan_info(const fx_cf_plan* p, int* info)
{
    if (p == nullptr || info == nullptr) return fail(FX_EINVAL, "null argument");
    info[0] = p->flux_off.empty() ? 0 : p->flux_off.back();
    info[1] = p->flux_off.empty() ? 0 : (int)p->flux_off.size() - 1;
    info[2] = p->emf_off.empty() ? 0 : p->emf_off.back();
    info[3] = p->emf_off.empty() ? 0 : (int)p->emf_off.size() - 1;
    return FX_OK;
}

int fx_cf_plan_destroy(fx_cf_plan* p)
{
    if (p == nullptr) return FX_OK;
    const hipError_t e = p->mem ? hipFree(p->mem) : hipSuccess;
    delete p;
    if (e != hipSuccess) return hip_fail(e, "hipFree");
    return FX_OK;
}

int fx_cf_flux_send(fx_cf_plan* p, void* stream)
{
    if (p == nullptr) return fail(FX_EINVAL, "null plan");
    if (!p->flux_err.empty()) return fail(FX_EINVAL, "flux correction unavailable: " + p->flux_err);
    for (size_t l = 0; l + 1 < p->flux_off.size(); ++l)
        FX_HIP(launch_cf_flux(p->tab, p->flux + p->flux_off[l], p->flux_off[l + 1] - p->flux_off[l],
                              p->nvar, (hipStream_t)stream));
    return FX_OK;
}

int fx_cf_emf_send(fx_cf_plan* p, int* status, void* stream)
{
    if (p == nullptr) return fail(FX_EINVAL, "null plan");
    if (!p->emf_err.empty()) return fail(FX_EINVAL, "EMF restriction unavailable: " + p->emf_err);
    FX_HIP(launch_cf_fid_init(p->tab, p->fid_blocks, p->nfid, (hipStream_t)stream));
    for (size_t l = 0; l + 1 < p->emf_off.size(); ++l)
        FX_HIP(launch_cf_emf(p->tab, p->emf + p->emf_off[l], p->emf_off[l + 1] - p->emf_off[l],
                             status, (hipStream_t)stream));
    return FX_OK;
}

}  // extern "C"
