// sph_context.hip -- resident single-scene API (sph_hip_create ... destroy).
//
// The scene stays in HBM between steps; only uploads and downloads cross
// PCIe.  A step is five kernels (scan, scatter, reorder, density,
// force_integrate); by default it is captured once into a HIP graph and
// replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "sph_runtime.h"

using namespace sph;

struct sph_ctx {
    hipStream_t stream = nullptr;
    Workspace w;
    DevBuf aos, xf;
    StepParams P;
    sph_scene_desc_v1 desc;
    uint32_t n = 0;
    bool uploaded = false;
    bool use_graph = true;
    bool keep_force = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    void drop_graph()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

namespace {

constexpr uint32_t kFlagKeepForce = 1u;

int step_once(sph_ctx *c, StepTimer *timer)
{
    float4 *force = c->keep_force ? (float4 *)c->w.force.ptr : nullptr;
    enqueue_step(c->stream, c->w, c->n, c->P, force, /*next_keys=*/true, timer);
    return SPH_OK;
}

int capture_graph(sph_ctx *c)
{
    c->drop_graph();
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess)
        return hip_fail(e, "hipStreamBeginCapture");
    step_once(c, nullptr);
    e = hipStreamEndCapture(c->stream, &c->graph);
    if (e != hipSuccess)
        return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        c->drop_graph();
        return hip_fail(e, "hipGraphInstantiate");
    }
    return SPH_OK;
}

}  // namespace

extern "C" {

int sph_hip_create(const sph_scene_desc_v1 *desc, sph_ctx **out)
{
    int rc;
    if (!desc || !out)
        return fail(SPH_E_INVALID, "create: NULL argument");
    *out = nullptr;
    if (desc->n == 0)
        return fail(SPH_E_INVALID, "create: n must be > 0");
    if (desc->n > (1u << 30))
        return fail(SPH_E_LIMIT, "create: %u particles exceed the 2^30 limit", desc->n);
    if ((desc->flags & ~kFlagKeepForce) != 0u)
        return fail(SPH_E_INVALID, "create: unknown flags 0x%x", desc->flags);
    if ((rc = check_settings(&desc->settings)))
        return rc;
    if (!std::isfinite(desc->dt))
        return fail(SPH_E_INVALID, "create: dt must be finite");
    if ((rc = ensure_device()))
        return rc;
    sph_ctx *c = new (std::nothrow) sph_ctx();
    if (!c)
        return fail(SPH_E_NOMEM, "create: host allocation failed");
    c->desc = *desc;
    c->n = desc->n;
    c->keep_force = (desc->flags & kFlagKeepForce) != 0u;
    c->P = make_params(desc->settings, desc->box, desc->dt);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreate");
    }
    if ((rc = c->w.reserve_particles(c->n)) ||
        (c->keep_force && (rc = c->w.force.reserve(sizeof(float4) * (size_t)c->n)))) {
        sph_hip_destroy(c);
        return rc;
    }
    *out = c;
    return SPH_OK;
}

int sph_hip_destroy(sph_ctx *c)
{
    if (!c)
        return SPH_OK;
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    c->drop_graph();
    c->w.release_all();
    c->aos.release();
    c->xf.release();
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
    return SPH_OK;
}

int sph_hip_upload(sph_ctx *c, const float *pos, const float *vel)
{
    int rc;
    if (!c || !pos)
        return fail(SPH_E_INVALID, "upload: NULL argument");
    const uint32_t n = c->n;
    // Grid bounds: the finite walls of the box, widened to the particles;
    // an open side gets head room of one particle extent (at least 16 h).
    double pmin[3] = {INFINITY, INFINITY, INFINITY}, pmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            double v = pos[3 * (size_t)i + k];
            if (std::isfinite(v)) {
                pmin[k] = std::min(pmin[k], v);
                pmax[k] = std::max(pmax[k], v);
            }
        }
    double lo[3], hi[3];
    const double h = c->P.h;
    for (int k = 0; k < 3; ++k) {
        if (!(pmin[k] <= pmax[k])) {
            pmin[k] = 0.0;
            pmax[k] = 0.0;
        }
        double ext = std::max(pmax[k] - pmin[k], 16.0 * h);
        double blo = c->desc.box.lo[k], bhi = c->desc.box.hi[k];
        lo[k] = std::isfinite(blo) ? std::min(blo, pmin[k]) : pmin[k] - ext;
        hi[k] = std::isfinite(bhi) ? std::max(bhi, pmax[k]) : pmax[k] + ext;
    }
    const double scale = c->desc.cell_scale > 1.0f ? (double)c->desc.cell_scale : 1.0;
    uint64_t max_cells = std::max<uint64_t>(16ull * n, 1ull << 20);
    make_grid(c->P, lo, hi, scale, max_cells);
    (void)hipStreamSynchronize(c->stream);
    c->drop_graph();
    if ((rc = c->w.reserve_cells((uint64_t)c->P.nx * c->P.ny * c->P.nz)))
        return rc;

    std::vector<float4> p4(n), v4(n);
    for (uint32_t i = 0; i < n; ++i) {
        p4[i] = make_float4(pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2], 0.0f);
        float vx = vel ? vel[3 * (size_t)i] : 0.0f, vy = vel ? vel[3 * (size_t)i + 1] : 0.0f;
        float vz = vel ? vel[3 * (size_t)i + 2] : 0.0f;
        uint32_t id = i;
        float idf;
        std::memcpy(&idf, &id, 4);
        v4[i] = make_float4(vx, vy, vz, idf);
    }
    ParticleArrays a = c->w.particles();
    CellArrays cells = c->w.cells();
    hipError_t e = hipMemcpyAsync(a.posA, p4.data(), sizeof(float4) * (size_t)n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(a.velA, v4.data(), sizeof(float4) * (size_t)n, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess)
        return hip_fail(e, "upload");
    launch_fill_u32(c->stream, cells.counts, 0u, cells.ncells_padded);
    launch_fill_u32(c->stream, (uint32_t *)a.rhoB, 0u, n);
    launch_fill_u32(c->stream, (uint32_t *)a.posR, 0u, 4ull * n);
    launch_fill_u32(c->stream, (uint32_t *)a.velR, 0u, 4ull * n);
    launch_fill_u32(c->stream, (uint32_t *)a.posB, 0u, 4ull * n);
    if (c->keep_force)
        launch_fill_u32(c->stream, (uint32_t *)c->w.force.ptr, 0u, 4ull * n);
    c->w.counts_dirty = false;
    launch_init_keys(c->stream, a, cells, n, c->P);
    e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess)
        e = hipGetLastError();
    if (e != hipSuccess)
        return hip_fail(e, "upload keys");
    c->uploaded = true;
    return SPH_OK;
}

int sph_hip_step(sph_ctx *c, int nsteps)
{
    int rc;
    if (!c || nsteps < 0)
        return fail(SPH_E_INVALID, "step: bad arguments");
    if (!c->uploaded)
        return fail(SPH_E_STATE, "step: no particles uploaded");
    if (c->use_graph) {
        if (!c->exec && (rc = capture_graph(c)))
            return rc;
        for (int k = 0; k < nsteps; ++k) {
            hipError_t e = hipGraphLaunch(c->exec, c->stream);
            if (e != hipSuccess)
                return hip_fail(e, "hipGraphLaunch");
        }
    } else {
        for (int k = 0; k < nsteps; ++k)
            step_once(c, nullptr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return hip_fail(e, "step launch");
    }
    return SPH_OK;
}

int sph_hip_sync(sph_ctx *c)
{
    if (!c)
        return fail(SPH_E_INVALID, "sync: NULL context");
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess)
        return hip_fail(e, "hipStreamSynchronize");
    return SPH_OK;
}

int sph_hip_download(sph_ctx *c, float *pos, float *vel, float *rho, float *pressure, uint32_t *ids)
{
    if (!c)
        return fail(SPH_E_INVALID, "download: NULL context");
    if (!c->uploaded)
        return fail(SPH_E_STATE, "download: no particles uploaded");
    const uint32_t n = c->n;
    ParticleArrays a = c->w.particles();
    std::vector<float4> p4, v4;
    std::vector<float> r;
    std::vector<float4> pi;
    hipError_t e = hipSuccess;
    if (pos) {
        p4.resize(n);
        e = hipMemcpyAsync(p4.data(), a.posA, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess && (vel || ids)) {
        v4.resize(n);
        e = hipMemcpyAsync(v4.data(), a.velA, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess && rho) {
        r.resize(n);
        e = hipMemcpyAsync(r.data(), a.rhoB, 4ull * n, hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess && pressure) {
        pi.resize(n);
        e = hipMemcpyAsync(pi.data(), a.velR, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess)
        return hip_fail(e, "download");
    for (uint32_t i = 0; i < n; ++i) {
        if (pos) {
            pos[3 * (size_t)i] = p4[i].x;
            pos[3 * (size_t)i + 1] = p4[i].y;
            pos[3 * (size_t)i + 2] = p4[i].z;
        }
        if (vel) {
            vel[3 * (size_t)i] = v4[i].x;
            vel[3 * (size_t)i + 1] = v4[i].y;
            vel[3 * (size_t)i + 2] = v4[i].z;
        }
        if (ids)
            std::memcpy(&ids[i], &v4[i].w, 4);
        if (rho)
            rho[i] = r[i];
        if (pressure)
            pressure[i] = pi[i].w;
    }
    return SPH_OK;
}

int sph_hip_download_particles(sph_ctx *c, sph_particle_v1 *out, float *transforms)
{
    int rc;
    if (!c || !out)
        return fail(SPH_E_INVALID, "download_particles: NULL argument");
    if (!c->uploaded)
        return fail(SPH_E_STATE, "download_particles: no particles uploaded");
    const uint32_t n = c->n;
    if ((rc = c->aos.reserve(sizeof(sph_particle_v1) * (size_t)n)) ||
        (transforms && (rc = c->xf.reserve(64ull * n))))
        return rc;
    ParticleArrays a = c->w.particles();
    launch_pack_state(c->stream, a, c->keep_force ? (const float4 *)c->w.force.ptr : nullptr, n, c->P,
                      c->aos.ptr, transforms ? (float *)c->xf.ptr : nullptr);
    hipError_t e = hipMemcpyAsync(out, c->aos.ptr, sizeof(sph_particle_v1) * (size_t)n, hipMemcpyDeviceToHost,
                                  c->stream);
    if (e == hipSuccess && transforms)
        e = hipMemcpyAsync(transforms, c->xf.ptr, 64ull * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess)
        return hip_fail(e, "download_particles");
    return SPH_OK;
}

int sph_hip_set_graph(sph_ctx *c, int enable)
{
    if (!c)
        return fail(SPH_E_INVALID, "set_graph: NULL context");
    c->use_graph = enable != 0;
    if (!c->use_graph)
        c->drop_graph();
    return SPH_OK;
}

uint32_t sph_hip_count(const sph_ctx *c) { return c ? c->n : 0u; }

int sph_hip_profile(sph_ctx *c, int nsteps, double *ms, int max_k, int *nk, char *names_buf,
                    size_t names_cap)
{
    if (!c || nsteps <= 0 || !ms || !nk || max_k <= 0)
        return fail(SPH_E_INVALID, "profile: bad arguments");
    if (!c->uploaded)
        return fail(SPH_E_STATE, "profile: no particles uploaded");
    std::vector<double> acc;
    std::vector<const char *> names;
    for (int s = 0; s < nsteps; ++s) {
        StepTimer t;
        step_once(c, &t);
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess)
            return hip_fail(e, "profile step");
        if (acc.empty()) {
            acc.assign(t.events.size() - 1, 0.0);
            names.assign(t.names.begin() + 1, t.names.end());
        }
        for (size_t k = 1; k < t.events.size(); ++k) {
            float dt = 0.0f;
            (void)hipEventElapsedTime(&dt, t.events[k - 1], t.events[k]);
            acc[k - 1] += dt;
        }
    }
    int k_out = (int)std::min<size_t>(acc.size(), (size_t)max_k);
    size_t pos = 0;
    for (int k = 0; k < k_out; ++k) {
        ms[k] = acc[k] / nsteps;
        if (names_buf) {
            size_t len = std::strlen(names[k]);
            if (pos + len + 1 <= names_cap) {
                std::memcpy(names_buf + pos, names[k], len + 1);
                pos += len + 1;
            }
        }
    }
    *nk = k_out;
    return SPH_OK;
}

}  // extern "C"
