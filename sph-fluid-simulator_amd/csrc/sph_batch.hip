// sph_batch.hip -- resident scenes: the single-scene context
// (sph_hip_create ... destroy) and the multi-scene batch (sph_batch_*).
//
// A batch keeps any number of independent scenes in one set of HBM arrays:
// scene s owns slots [p_begin, p_end) and grid cells [cell_base, cell_base +
// nx ny nz), so one launch per pass advances every scene, and the global
// counting sort never mixes scenes.  Each scene's grid depends only on its
// own box and particles, and every per-particle result depends only on its
// own scene, so a scene steps bit-identically alone or in any mix.  A step
// is captured once into a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "sph_batch.h"

namespace sph {

namespace {
constexpr uint32_t kKnownFlags = SPH_SCENE_KEEP_FORCE | SPH_SCENE_LANE_WALK | SPH_SCENE_AUTO_GRID;
}

int validate_desc(const sph_scene_desc_v1 &d)
{
    int rc;
    if (d.n == 0)
        return fail(SPH_E_INVALID, "scene: n must be > 0");
    if (d.n > kMaxParticles)
        return fail(SPH_E_LIMIT, "scene: %u particles exceed the 2^30 limit", d.n);
    if ((d.flags & ~kKnownFlags) != 0u)
        return fail(SPH_E_INVALID, "scene: unknown flags 0x%x", d.flags);
    if ((rc = check_settings(&d.settings)))
        return rc;
    if (!std::isfinite(d.dt))
        return fail(SPH_E_INVALID, "scene: dt must be finite");
    if (grid_reach(d.cell_scale) == 0)
        return fail(SPH_E_INVALID, "scene: cell_scale must be 0, >= 1, or in [1/%d, 1)", kMaxReach);
    return SPH_OK;
}

// Grid of one scene from its box and particles: the finite walls, widened
// to the particles; an open side gets head room of one particle extent (at
// least 16 h).  Depends on nothing but the scene.
void scene_grid(BatchScene &b, const float *pos, uint32_t n)
{
    const sph_scene_desc_v1 &d = b.desc;
    double pmin[3] = {INFINITY, INFINITY, INFINITY}, pmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            double v = pos[3 * (size_t)i + k];
            if (std::isfinite(v)) {
                pmin[k] = std::min(pmin[k], v);
                pmax[k] = std::max(pmax[k], v);
            }
        }
    double *lo = b.lo, *hi = b.hi;
    const double h = b.P.h;
    for (int k = 0; k < 3; ++k) {
        if (!(pmin[k] <= pmax[k])) {
            pmin[k] = 0.0;
            pmax[k] = 0.0;
        }
        double ext = std::max(pmax[k] - pmin[k], 16.0 * h);
        double blo = d.box.lo[k], bhi = d.box.hi[k];
        lo[k] = std::isfinite(blo) ? std::min(blo, pmin[k]) : pmin[k] - ext;
        hi[k] = std::isfinite(bhi) ? std::max(bhi, pmax[k]) : pmax[k] + ext;
    }
    (void)scene_regrid(b, 0.0);  // desc.cell_scale was validated
}

int scene_regrid(BatchScene &b, double scale)
{
    if (!(scale > 0.0))
        scale = b.desc.cell_scale > 0.0f ? (double)b.desc.cell_scale : 1.0;
    const uint64_t max_cells = std::max<uint64_t>(64ull * b.desc.n, 1ull << 16);
    return make_grid(b.P, b.lo, b.hi, scale, max_cells);
}

void Batch::drop_graph()
{
    if (exec)
        (void)hipGraphExecDestroy(exec);
    if (graph)
        (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
}

Batch::~Batch()
{
    if (stream)
        (void)hipStreamSynchronize(stream);
    for (hipEvent_t e : events)
        (void)hipEventDestroy(e);
    drop_graph();
    w.release_all();
    if (stream)
        (void)hipStreamDestroy(stream);
}

int make_scene(const sph_scene_desc_v1 &d, BatchScene &out)
{
    int rc;
    if ((rc = validate_desc(d)))
        return rc;
    out = BatchScene();
    out.desc = d;
    out.P = make_params(d.settings, d.box, d.dt);
    out.P.lane_walk = (d.flags & SPH_SCENE_LANE_WALK) ? 1u : 0u;
    return SPH_OK;
}

int Batch::init(const sph_scene_desc_v1 *descs, uint32_t k)
{
    int rc;
    if (!descs || k == 0)
        return fail(SPH_E_INVALID, "batch: no scenes");
    uint64_t total = 0;
    scenes.clear();
    for (uint32_t s = 0; s < k; ++s) {
        BatchScene b;
        if ((rc = make_scene(descs[s], b)))
            return rc;
        b.P.p_begin = (uint32_t)total;
        total += descs[s].n;
        if (total > kMaxParticles)
            return fail(SPH_E_LIMIT, "batch: %llu particles exceed the 2^30 limit", (unsigned long long)total);
        b.P.p_end = (uint32_t)total;
        keep_force = keep_force || (descs[s].flags & SPH_SCENE_KEEP_FORCE);
        scenes.push_back(b);
    }
    total_n = (uint32_t)total;
    if ((rc = init_empty(keep_force)))
        return rc;
    if ((rc = w.reserve_particles(total_n)) ||
        (keep_force && (rc = w.force.reserve(sizeof(float4) * (size_t)total_n))))
        return rc;
    return SPH_OK;
}

int Batch::init_empty(bool keep_force_)
{
    int rc;
    keep_force = keep_force_;
    if ((rc = ensure_device()))
        return rc;
    hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        stream = nullptr;
        return hip_fail(e, "hipStreamCreate");
    }
    return SPH_OK;
}

namespace {
void swap_buf(DevBuf &x, DevBuf &y)
{
    std::swap(x.ptr, y.ptr);
    std::swap(x.bytes, y.bytes);
}
}  // namespace

int Batch::relayout(std::vector<BatchScene> next, const std::vector<int> &from)
{
    int rc;
    if (next.size() != from.size())
        return fail(SPH_E_INVALID, "relayout: size mismatch");
    uint64_t total = 0;
    for (size_t k = 0; k < next.size(); ++k) {
        if (from[k] >= (int)scenes.size())
            return fail(SPH_E_INVALID, "relayout: bad source scene");
        next[k].P.p_begin = (uint32_t)total;
        total += next[k].desc.n;
        if (total > kMaxParticles)
            return fail(SPH_E_LIMIT, "relayout: %llu particles exceed the 2^30 limit", (unsigned long long)total);
        next[k].P.p_end = (uint32_t)total;
        next[k].uploaded = from[k] >= 0 && scenes[from[k]].uploaded;
    }
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess)
        return hip_fail(e, "relayout");
    drop_graph();
    // Destination of the moved state: the B arrays (scratch between steps)
    // when they are large enough, else fresh buffers with head room.
    const uint64_t cap = w.posA.bytes / sizeof(float4);
    const bool grow = total > cap;
    DevBuf np, nv;
    float4 *dp = (float4 *)w.posB.ptr, *dv = (float4 *)w.velB.ptr;
    if (grow) {
        uint64_t want = std::max<uint64_t>(total, cap + cap / 2);
        if (max_capacity)
            want = std::min<uint64_t>(want, std::max<uint64_t>(total, max_capacity));
        want = std::min<uint64_t>(want, kMaxParticles);
        if ((rc = np.reserve(sizeof(float4) * want)) || (rc = nv.reserve(sizeof(float4) * want)))
            return rc;
        dp = (float4 *)np.ptr;
        dv = (float4 *)nv.ptr;
    }
    for (size_t k = 0; k < next.size(); ++k) {
        if (from[k] < 0 || !next[k].uploaded)
            continue;
        const BatchScene &o = scenes[from[k]];
        const size_t bytes = sizeof(float4) * (size_t)o.desc.n;
        e = hipMemcpyAsync(dp + next[k].P.p_begin, (float4 *)w.posA.ptr + o.P.p_begin, bytes,
                           hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dv + next[k].P.p_begin, (float4 *)w.velA.ptr + o.P.p_begin, bytes,
                               hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess)
            return hip_fail(e, "relayout copy");
    }
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess)
        return hip_fail(e, "relayout copy");
    if (inject_failure == kFailRelayout) {
        inject_failure = 0;
        return fail(SPH_E_NOMEM, "relayout: injected failure");
    }
    if (grow) {
        // np / nv keep the old state until every other array has grown, so
        // a failure here leaves the batch on its old layout (a failed
        // growth keeps each array as it was)
        swap_buf(w.posA, np);
        swap_buf(w.velA, nv);
        const uint32_t want = (uint32_t)(w.posA.bytes / sizeof(float4));
        if ((rc = w.reserve_particles(want)) ||
            (keep_force && (rc = w.force.reserve(sizeof(float4) * (size_t)want)))) {
            swap_buf(w.posA, np);
            swap_buf(w.velA, nv);
            return rc;
        }
        np.release();
        nv.release();
    } else {
        swap_buf(w.posA, w.posB);
        swap_buf(w.velA, w.velB);
    }
    scenes = std::move(next);
    total_n = (uint32_t)total;
    layout_dirty = true;
    keys_dirty = true;
    stat_ok = false;
    return SPH_OK;
}

int Batch::upload(uint32_t s, const float *pos, const float *vel)
{
    if (s >= scenes.size() || !pos)
        return fail(SPH_E_INVALID, "upload: bad scene or NULL positions");
    if (inject_failure == kFailUpload) {
        inject_failure = 0;
        return fail(SPH_E_NOMEM, "upload: injected failure");
    }
    BatchScene &b = scenes[s];
    const uint32_t n = b.desc.n, p0 = b.P.p_begin;
    scene_grid(b, pos, n);
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
    (void)hipStreamSynchronize(stream);
    drop_graph();
    ParticleArrays a = w.particles();
    hipError_t e = hipMemcpyAsync(a.posA + p0, p4.data(), sizeof(float4) * (size_t)n, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(a.velA + p0, v4.data(), sizeof(float4) * (size_t)n, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess)
        return hip_fail(e, "upload");
    launch_fill_u32(stream, (uint32_t *)(a.rhoB + p0), 0u, n);
    launch_fill_u32(stream, (uint32_t *)(a.posR + p0), 0u, 4ull * n);
    launch_fill_u32(stream, (uint32_t *)(a.velR + p0), 0u, 4ull * n);
    launch_fill_u32(stream, (uint32_t *)(a.posB + p0), 0u, 4ull * n);
    if (keep_force)
        launch_fill_u32(stream, (uint32_t *)w.force.ptr + 4ull * p0, 0u, 4ull * n);
    e = hipStreamSynchronize(stream);
    if (e == hipSuccess)
        e = hipGetLastError();
    if (e != hipSuccess)
        return hip_fail(e, "upload");
    b.uploaded = true;
    b.steps = 0;
    b.scale = 0.0;
    b.last_ratio = 0.0;
    b.evaluated = ~0ull;
    layout_dirty = true;
    keys_dirty = true;
    stat_ok = false;
    // Once every scene holds particles, the first step's set-up (cells,
    // scene table, keys, the automatic grid's first choice) runs here
    // rather than inside the first step.
    for (const BatchScene &o : scenes)
        if (!o.uploaded)
            return SPH_OK;
    int rc;
    if (!hold_grid && (rc = auto_grid()))
        return rc;
    if ((rc = prepare()))
        return rc;
    if (use_graph && (rc = capture()))  // the step's graph too, outside the first step
        return rc;
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess)
        return hip_fail(e, "upload");
    return SPH_OK;
}

int Batch::prepare()
{
    int rc;
    for (const BatchScene &b : scenes)
        if (!b.uploaded)
            return fail(SPH_E_STATE, "batch: every scene needs an upload before stepping");
    if (layout_dirty) {
        uint64_t base = 0;
        std::vector<StepParams> table;
        for (BatchScene &b : scenes) {
            b.P.cell_base = (uint32_t)base;
            base += (uint64_t)b.P.nx * b.P.ny * b.P.nz;
            if (base > (1ull << 31))
                return fail(SPH_E_LIMIT, "batch: %llu grid cells exceed the 2^31 limit", (unsigned long long)base);
            table.push_back(b.P);
        }
        drop_graph();
        stat_ok = false;
        if ((rc = w.reserve_cells(base)) || (rc = w.set_scenes(table, stream)))
            return rc;
        layout_dirty = false;
        keys_dirty = true;
    }
    if (keys_dirty) {
        CellArrays c = w.cells();
        launch_fill_u32(stream, c.counts, 0u, c.ncells_padded);
        launch_init_keys(stream, w.particles(), c, w.scene_set());
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return hip_fail(e, "keys");
        keys_dirty = false;
    }
    return SPH_OK;
}

void Batch::enqueue(StepTimer *timer)
{
    enqueue_step(stream, w, keep_force ? (float4 *)w.force.ptr : nullptr, /*next_keys=*/true, timer);
}

int Batch::step(int nsteps)
{
    int rc;
    if (nsteps < 0)
        return fail(SPH_E_INVALID, "step: negative step count");
    if (scenes.empty() || nsteps == 0)
        return SPH_OK;
    bool any_auto = false;
    for (const BatchScene &b : scenes)
        any_auto = any_auto || (b.desc.flags & SPH_SCENE_AUTO_GRID) != 0u;
    if (hold_grid || !any_auto) {
        if ((rc = step_plain(nsteps)))
            return rc;
        if (!hold_grid)
            for (BatchScene &b : scenes)
                b.steps += (uint64_t)nsteps;
        return SPH_OK;
    }
    // Evaluates the automatic scenes due (step count a multiple of
    // kRegridEvery, the upload included) and runs up to the next point.
    uint64_t left = (uint64_t)nsteps;
    while (left > 0) {
        if ((rc = auto_grid()))
            return rc;
        uint64_t k = left;
        for (const BatchScene &b : scenes)
            if (b.desc.flags & SPH_SCENE_AUTO_GRID)
                k = std::min<uint64_t>(k, kRegridEvery - b.steps % kRegridEvery);
        if ((rc = step_plain((int)k)))
            return rc;
        for (BatchScene &b : scenes)
            b.steps += k;
        left -= k;
    }
    return SPH_OK;
}

int Batch::grid_stat(std::vector<double> &ratio, bool fresh)
{
    int rc;
    ratio.assign(scenes.size(), 0.0);
    if (fresh) {
        // the cell order of the current state (the next step redoes it and
        // recomputes the keys first: the scatter consumes the cell counts)
        if ((rc = prepare()))
            return rc;
        const ParticleArrays a = w.particles();
        const CellArrays c = w.cells();
        const SceneSet S = w.scene_set();
        launch_scan(stream, c, false);
        launch_scatter(stream, a, c, S);
        launch_sort_big_cells(stream, a, c);
        launch_reorder(stream, a, c, S);
        keys_dirty = true;
        stat_ok = true;
    }
    if (!stat_ok)
        return SPH_OK;
    const size_t ns = scenes.size();
    if ((rc = gstat.reserve(16ull * ns)))
        return rc;
    launch_grid_stat(stream, w.particles(), w.cells(), w.scene_set(), (unsigned long long *)gstat.ptr, (uint32_t)ns);
    std::vector<unsigned long long> st(2 * ns);
    hipError_t e = hipMemcpyAsync(st.data(), gstat.ptr, 16ull * ns, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(stream);
    if (e == hipSuccess)
        e = hipGetLastError();
    if (e != hipSuccess)
        return hip_fail(e, "grid statistic");
    for (size_t k = 0; k < ns; ++k)
        ratio[k] = st[2 * k + 1] ? (double)st[2 * k] / (double)st[2 * k + 1] : (st[2 * k] ? INFINITY : 0.0);
    return SPH_OK;
}

int Batch::grid_info(uint32_t s, float *cell_scale, double *cand_per_nbr)
{
    if (s >= scenes.size())
        return fail(SPH_E_INVALID, "grid_info: bad scene");
    const BatchScene &b = scenes[s];
    if (cell_scale)
        *cell_scale = (float)(b.scale > 0.0 ? b.scale : (b.desc.cell_scale > 0.0f ? (double)b.desc.cell_scale : 1.0));
    if (cand_per_nbr) {
        std::vector<double> ratio;
        int rc = grid_stat(ratio, false);
        if (rc)
            return rc;
        *cand_per_nbr = ratio[s];
    }
    return SPH_OK;
}

int Batch::unordered_cells(uint64_t *out)
{
    if (!out)
        return fail(SPH_E_INVALID, "unordered_cells: NULL output");
    *out = 0;
    if (!w.unordered.ptr)
        return SPH_OK;
    unsigned long long v = 0;
    hipError_t e = hipMemcpyAsync(&v, w.unordered.ptr, sizeof(v), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(stream);
    if (e != hipSuccess)
        return hip_fail(e, "reading the unordered-cell counter");
    *out = (uint64_t)v;
    return SPH_OK;
}

int Batch::auto_grid()
{
    int rc;
    auto due = [](const BatchScene &b) {
        const double cur = b.scale > 0.0 ? b.scale : (b.desc.cell_scale > 0.0f ? (double)b.desc.cell_scale : 1.0);
        const bool clumped = b.steps > 0 && cur >= 1.0 && b.last_ratio < kClumpedRatio;
        return (b.desc.flags & SPH_SCENE_AUTO_GRID) != 0u && b.steps != b.evaluated &&
               b.steps % (clumped ? kRegridClumped : kRegridEvery) == 0;
    };
    bool any = false;
    for (const BatchScene &b : scenes)
        any = any || due(b);
    if (!any)
        return SPH_OK;
    std::vector<double> ratio;
    if ((rc = grid_stat(ratio, /*fresh=*/true)))
        return rc;
    for (size_t k = 0; k < scenes.size(); ++k) {
        BatchScene &b = scenes[k];
        if (!due(b))
            continue;
        b.last_ratio = ratio[k];
        b.evaluated = b.steps;
        const double cur = b.scale > 0.0 ? b.scale : (b.desc.cell_scale > 0.0f ? (double)b.desc.cell_scale : 1.0);
        double want = cur;
        if (cur >= 1.0 && ratio[k] > kToFineRatio && b.desc.n >= kAutoGridMinParticles)
            want = 0.5;
        else if (cur < 1.0 && ratio[k] < kToCoarseRatio)
            want = 1.0;
        if (want != cur) {
            if ((rc = scene_regrid(b, want)))
                return rc;
            b.scale = want;
            layout_dirty = true;  // new cells, scene table and keys before the next step
        }
    }
    return SPH_OK;
}

int Batch::capture()
{
    if (exec)
        return SPH_OK;
    hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess)
        return hip_fail(e, "hipStreamBeginCapture");
    enqueue(nullptr);
    e = hipStreamEndCapture(stream, &graph);
    if (e == hipSuccess)
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph();
        return hip_fail(e, "graph capture");
    }
    return SPH_OK;
}

int Batch::step_plain(int nsteps)
{
    int rc;
    if (nsteps < 0)
        return fail(SPH_E_INVALID, "step: negative step count");
    if (scenes.empty() || nsteps == 0)
        return SPH_OK;
    if ((rc = prepare()))
        return rc;
    if (use_graph) {
        if ((rc = capture()))
            return rc;
        for (int k = 0; k < nsteps; ++k) {
            hipError_t e = hipGraphLaunch(exec, stream);
            if (e != hipSuccess)
                return hip_fail(e, "hipGraphLaunch");
        }
    } else {
        for (int k = 0; k < nsteps; ++k)
            enqueue(nullptr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return hip_fail(e, "step launch");
    }
    stat_ok = true;  // the B order and cell ranges are the last step's
    return SPH_OK;
}

int Batch::step_timed(int nsteps, float *ms)
{
    int rc;
    if (nsteps < 0)
        return fail(SPH_E_INVALID, "step: negative step count");
    if (scenes.empty() || nsteps == 0)
        return SPH_OK;
    if ((rc = prepare()))
        return rc;
    while (events.size() < (size_t)nsteps + 1) {
        hipEvent_t e;
        hipError_t err = hipEventCreate(&e);
        if (err != hipSuccess)
            return hip_fail(err, "hipEventCreate");
        events.push_back(e);
    }
    hipError_t e = hipEventRecord(events[0], stream);
    for (int s = 0; s < nsteps && e == hipSuccess; ++s) {
        if ((rc = step(1)))
            return rc;
        e = hipEventRecord(events[s + 1], stream);
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(stream);
    if (e != hipSuccess)
        return hip_fail(e, "timed step");
    for (int s = 0; ms && s < nsteps; ++s) {
        ms[s] = 0.0f;
        (void)hipEventElapsedTime(&ms[s], events[s], events[s + 1]);
    }
    return SPH_OK;
}

int Batch::autotune(int trial_steps, sph_tune_v1 *out, bool grid)
{
    int rc;
    if (trial_steps <= 0 || !out)
        return fail(SPH_E_INVALID, "autotune: bad arguments");
    if (scenes.empty())
        return fail(SPH_E_STATE, "autotune: no scenes");
    if ((rc = prepare()))
        return rc;
    std::memset(out, 0, sizeof *out);
    // The state every trial starts from (and is restored to at the end).
    DevBuf sp, sv;
    const size_t bytes = sizeof(float4) * (size_t)total_n;
    if ((rc = sp.reserve(bytes)) || (rc = sv.reserve(bytes)))
        return rc;
    hipError_t e = hipMemcpyAsync(sp.ptr, w.posA.ptr, bytes, hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(sv.ptr, w.velA.ptr, bytes, hipMemcpyDeviceToDevice, stream);
    if (e != hipSuccess)
        return hip_fail(e, "autotune snapshot");
    auto restore = [&]() -> int {
        hipError_t err = hipMemcpyAsync(w.posA.ptr, sp.ptr, bytes, hipMemcpyDeviceToDevice, stream);
        if (err == hipSuccess)
            err = hipMemcpyAsync(w.velA.ptr, sv.ptr, bytes, hipMemcpyDeviceToDevice, stream);
        if (err != hipSuccess)
            return hip_fail(err, "autotune restore");
        keys_dirty = true;
        stat_ok = false;
        return SPH_OK;
    };
    // Trial steps neither count toward the scenes' step counts nor re-grid.
    struct Hold {
        bool &flag;
        ~Hold() { flag = false; }
    } hold{hold_grid};
    hold_grid = true;
    // The walk (per lane or union) is not a knob: it may change split sums.
    auto apply = [&](int block, double scale) -> int {
        w.stencil_block = block;
        if (grid)
            for (BatchScene &b : scenes) {
                b.desc.cell_scale = (float)scale;  // kept by later uploads
                b.scale = 0.0;
                int r = scene_regrid(b, 0.0);
                if (r)
                    return r;
            }
        layout_dirty = true;  // re-uploads the scene table (and re-keys on a new grid)
        return SPH_OK;
    };
    const int blocks[3] = {64, 128, 256};
    const double scales[2] = {1.0, 0.5};
    const int nscales = grid ? 2 : 1;
    std::vector<float> ms((size_t)trial_steps);
    double best = -1.0, best_scale = 0.0;
    int k = 0;
    for (int si = 0; si < nscales; ++si)
        for (int bi = 0; bi < 3; ++bi, ++k) {
            const double scale = grid ? scales[si] : 0.0;
            if ((rc = apply(blocks[bi], scale)) || (rc = restore()) || (rc = step_timed(1, nullptr)) ||
                (rc = restore()) || (rc = step_timed(trial_steps, ms.data())))
                return rc;
            double sum = 0.0;
            for (float v : ms)
                sum += v;
            const double mean = sum / trial_steps;
            out->trial_ms[k] = mean;
            if (blocks[bi] == kStencilBlock && si == 0)
                out->default_ms_per_step = mean;
            if (best < 0.0 || mean < best) {
                best = mean;
                best_scale = scale;
                out->stencil_block = blocks[bi];
            }
        }
    out->ms_per_step = best;
    out->trials = k;
    out->lane_walk = 0;
    out->cell_scale = (float)best_scale;
    if ((rc = apply(out->stencil_block, best_scale)))
        return rc;
    if ((rc = restore()))
        return rc;
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess)
        return hip_fail(e, "autotune");
    return SPH_OK;
}

int Batch::sync()
{
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess)
        return hip_fail(e, "hipStreamSynchronize");
    return SPH_OK;
}

int Batch::download(uint32_t s, float *pos, float *vel, float *rho, float *pressure, uint32_t *ids)
{
    if (s >= scenes.size())
        return fail(SPH_E_INVALID, "download: bad scene");
    if (!scenes[s].uploaded)
        return fail(SPH_E_STATE, "download: scene not uploaded");
    const uint32_t n = scenes[s].desc.n, p0 = scenes[s].P.p_begin;
    ParticleArrays a = w.particles();
    std::vector<float4> p4, v4, r4;
    std::vector<float> r;
    hipError_t e = hipSuccess;
    if (pos) {
        p4.resize(n);
        e = hipMemcpyAsync(p4.data(), a.posA + p0, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost, stream);
    }
    if (e == hipSuccess && (vel || ids)) {
        v4.resize(n);
        e = hipMemcpyAsync(v4.data(), a.velA + p0, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost, stream);
    }
    if (e == hipSuccess && rho) {
        r.resize(n);
        e = hipMemcpyAsync(r.data(), a.rhoB + p0, 4ull * n, hipMemcpyDeviceToHost, stream);
    }
    if (e == hipSuccess && pressure) {
        r4.resize(n);
        e = hipMemcpyAsync(r4.data(), a.velR + p0, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost, stream);
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(stream);
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
            pressure[i] = r4[i].w;
    }
    return SPH_OK;
}

int Batch::download_particles(uint32_t s, sph_particle_v1 *out, float *transforms)
{
    int rc;
    if (s >= scenes.size() || !out)
        return fail(SPH_E_INVALID, "download_particles: bad scene or NULL output");
    if (!scenes[s].uploaded)
        return fail(SPH_E_STATE, "download_particles: scene not uploaded");
    const uint32_t n = scenes[s].desc.n, p0 = scenes[s].P.p_begin;
    if ((rc = aos.reserve(sizeof(sph_particle_v1) * (size_t)n)) || (transforms && (rc = xf.reserve(64ull * n))))
        return rc;
    launch_pack_state(stream, w.particles(), keep_force ? (const float4 *)w.force.ptr : nullptr, p0, n,
                      scenes[s].P.xf_scale, aos.ptr, transforms ? (float *)xf.ptr : nullptr);
    hipError_t e = hipMemcpyAsync(out, aos.ptr, sizeof(sph_particle_v1) * (size_t)n, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess && transforms)
        e = hipMemcpyAsync(transforms, xf.ptr, 64ull * n, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(stream);
    if (e != hipSuccess)
        return hip_fail(e, "download_particles");
    return SPH_OK;
}

int Batch::profile(int nsteps, double *ms, int max_k, int *nk, char *names_buf, size_t names_cap)
{
    int rc;
    if (nsteps <= 0 || !ms || !nk || max_k <= 0)
        return fail(SPH_E_INVALID, "profile: bad arguments");
    if ((rc = prepare()))
        return rc;
    std::vector<double> acc;
    std::vector<const char *> names;
    for (int s = 0; s < nsteps; ++s) {
        // the same schedule of grid choices as step() (outside the timed kernels)
        if (!hold_grid && ((rc = auto_grid()) || (rc = prepare())))
            return rc;
        StepTimer t;
        enqueue(&t);
        stat_ok = true;
        if (!hold_grid)
            for (BatchScene &b : scenes)
                ++b.steps;
        hipError_t e = hipStreamSynchronize(stream);
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

int Batch::walk_dump(uint32_t s, uint32_t *count, uint32_t *walk, uint32_t *list_count, uint32_t *list)
{
    int rc;
    if (s >= scenes.size() || !count)
        return fail(SPH_E_INVALID, "walk_dump: bad scene or NULL count");
    if ((rc = prepare()))
        return rc;
    const uint32_t N = total_n;
    // The B-side arrays hold the last step's order and densities, which the
    // downloads read: the dump's own go to temporaries.
    DevBuf tpos, tvel, tcell, tperm, trho, tposR, tvelR, tcount;
    if ((rc = tpos.reserve(16ull * N)) || (rc = tvel.reserve(16ull * N)) || (rc = tcell.reserve(4ull * N)) ||
        (rc = tperm.reserve(4ull * N)) || (rc = trho.reserve(4ull * N)) || (rc = tposR.reserve(16ull * N)) ||
        (rc = tvelR.reserve(16ull * N)) || (rc = tcount.reserve(4ull * N)))
        return rc;
    ParticleArrays a = w.particles();
    a.posB = (float4 *)tpos.ptr;
    a.velB = (float4 *)tvel.ptr;
    a.cellcB = (uint32_t *)tcell.ptr;
    a.perm = (uint32_t *)tperm.ptr;
    a.rhoB = (float *)trho.ptr;
    a.posR = (float4 *)tposR.ptr;
    a.velR = (float4 *)tvelR.ptr;
    const CellArrays c = w.cells();
    const SceneSet S = w.scene_set();
    const PartPlan pp = w.part_plan();
    const NeighborLists nl = w.lists();
    launch_scan(stream, c, false);
    keys_dirty = true;  // the scan consumed the cell counts: the next step recomputes the keys
    stat_ok = false;    // and the cell ranges now belong to this dump
    launch_scatter(stream, a, c, S);
    launch_sort_big_cells(stream, a, c);
    launch_reorder(stream, a, c, S);
    launch_plan_parts(stream, a, c, S, pp);
    NeighborDump d{(uint32_t *)tcount.ptr, nullptr, nullptr, true};
    launch_density(stream, a, c, S, &d, &pp, kStencilBlock, &nl);

    const BatchScene &b = scenes[s];
    const uint32_t n = b.desc.n, p0 = b.P.p_begin;
    uint32_t t0 = 0;  // first wave task of scene s (tasks follow the scenes in order)
    for (uint32_t k = 0; k < s; ++k)
        t0 += (scenes[k].desc.n + 63u) / 64u;
    const uint32_t nt = (n + 63u) / 64u;
    std::vector<uint32_t> cnt(n), lc(n), off(nt + 1), ids(n);
    std::vector<float4> vb(n);
    std::vector<int4> wk(2ull * nt);
    std::vector<uint32_t> lst(list ? (size_t)kListMax * n : 0);
    hipError_t e = hipMemcpyAsync(cnt.data(), (uint32_t *)tcount.ptr + p0, 4ull * n, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(lc.data(), nl.count + p0, 4ull * n, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(vb.data(), a.velB + p0, 16ull * n, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(off.data(), pp.off + t0, 4ull * (nt + 1), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(wk.data(), pp.walk + 2ull * t0, sizeof(int4) * 2ull * nt, hipMemcpyDeviceToHost, stream);
    for (uint32_t k = 0; list && k < kListMax && e == hipSuccess; ++k)  // entry k of every slot
        e = hipMemcpyAsync(lst.data() + list_at(k, 0, n), nl.list + list_at(k, p0, nl.stride), 4ull * n,
                           hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(stream);
    if (e == hipSuccess)
        e = hipGetLastError();
    if (e != hipSuccess)
        return hip_fail(e, "walk_dump");
    for (uint32_t q = 0; q < n; ++q)
        std::memcpy(&ids[q], &vb[q].w, 4);
    for (uint32_t q = 0; q < n; ++q) {
        const uint32_t id = ids[q], t = q / 64u;
        if (id >= n)
            return fail(SPH_E_HIP, "walk_dump: slot %u holds id %u of %u", q, id, n);
        count[id] = cnt[q];
        if (walk)
            walk[id] = (wk[2 * t + 1].w ? 1u : 0u) | (off[t + 1] != off[t] ? 2u : 0u) | (lc[q] != kNoList ? 4u : 0u);
        if (list_count)
            list_count[id] = lc[q];
        if (list && lc[q] != kNoList)
            for (uint32_t k = 0; k < lc[q]; ++k)
                list[(size_t)id * kListMax + k] = ids[lst[list_at(k, q, n)] - p0];
    }
    return SPH_OK;
}

}  // namespace sph

using namespace sph;

struct sph_ctx {
    Batch b;
};

struct sph_batch {
    Batch b;
};

namespace {

template <class T>
int create_object(const sph_scene_desc_v1 *descs, uint32_t k, T **out)
{
    if (!out)
        return fail(SPH_E_INVALID, "create: NULL output");
    *out = nullptr;
    T *o = new (std::nothrow) T();
    if (!o)
        return fail(SPH_E_NOMEM, "create: host allocation failed");
    int rc = o->b.init(descs, k);
    if (rc) {
        delete o;
        return rc;
    }
    *out = o;
    return SPH_OK;
}

}  // namespace

extern "C" {

// ---- single scene

int sph_hip_create(const sph_scene_desc_v1 *desc, sph_ctx **out)
{
    if (!desc)
        return fail(SPH_E_INVALID, "create: NULL argument");
    return create_object(desc, 1u, out);
}

int sph_hip_destroy(sph_ctx *c)
{
    delete c;
    return SPH_OK;
}

int sph_hip_upload(sph_ctx *c, const float *pos, const float *vel)
{
    if (!c)
        return fail(SPH_E_INVALID, "upload: NULL context");
    return c->b.upload(0, pos, vel);
}

int sph_hip_step(sph_ctx *c, int nsteps)
{
    if (!c)
        return fail(SPH_E_INVALID, "step: NULL context");
    return c->b.step(nsteps);
}

int sph_hip_sync(sph_ctx *c)
{
    if (!c)
        return fail(SPH_E_INVALID, "sync: NULL context");
    return c->b.sync();
}

int sph_hip_download(sph_ctx *c, float *pos, float *vel, float *rho, float *pressure, uint32_t *ids)
{
    if (!c)
        return fail(SPH_E_INVALID, "download: NULL context");
    return c->b.download(0, pos, vel, rho, pressure, ids);
}

int sph_hip_download_particles(sph_ctx *c, sph_particle_v1 *out, float *transforms)
{
    if (!c)
        return fail(SPH_E_INVALID, "download_particles: NULL context");
    return c->b.download_particles(0, out, transforms);
}

int sph_hip_set_graph(sph_ctx *c, int enable)
{
    if (!c)
        return fail(SPH_E_INVALID, "set_graph: NULL context");
    c->b.use_graph = enable != 0;
    if (!c->b.use_graph)
        c->b.drop_graph();
    return SPH_OK;
}

uint32_t sph_hip_count(const sph_ctx *c) { return c ? c->b.total_n : 0u; }

int sph_hip_profile(sph_ctx *c, int nsteps, double *ms, int max_k, int *nk, char *names_buf,
                    size_t names_cap)
{
    if (!c)
        return fail(SPH_E_INVALID, "profile: NULL context");
    return c->b.profile(nsteps, ms, max_k, nk, names_buf, names_cap);
}

// ---- batch of scenes

int sph_hip_walk_dump(sph_ctx *c, uint32_t *count, uint32_t *walk, uint32_t *list_count, uint32_t *list)
{
    if (!c)
        return fail(SPH_E_INVALID, "walk_dump: NULL context");
    return c->b.walk_dump(0, count, walk, list_count, list);
}

int sph_batch_download_particles(sph_batch *b, uint32_t scene, sph_particle_v1 *out, float *transforms)
{
    if (!b)
        return fail(SPH_E_INVALID, "download_particles: NULL batch");
    return b->b.download_particles(scene, out, transforms);
}

int sph_hip_grid_info(sph_ctx *c, float *cell_scale, double *cand_per_nbr)
{
    if (!c)
        return fail(SPH_E_INVALID, "grid_info: NULL context");
    return c->b.grid_info(0, cell_scale, cand_per_nbr);
}

int sph_batch_grid_info(sph_batch *b, uint32_t scene, float *cell_scale, double *cand_per_nbr)
{
    if (!b)
        return fail(SPH_E_INVALID, "grid_info: NULL batch");
    return b->b.grid_info(scene, cell_scale, cand_per_nbr);
}

int sph_hip_unordered_cells(sph_ctx *c, uint64_t *out)
{
    if (!c)
        return fail(SPH_E_INVALID, "unordered_cells: NULL context");
    return c->b.unordered_cells(out);
}

int sph_batch_unordered_cells(sph_batch *b, uint64_t *out)
{
    if (!b)
        return fail(SPH_E_INVALID, "unordered_cells: NULL batch");
    return b->b.unordered_cells(out);
}

int sph_batch_walk_dump(sph_batch *b, uint32_t scene, uint32_t *count, uint32_t *walk, uint32_t *list_count,
                        uint32_t *list)
{
    if (!b)
        return fail(SPH_E_INVALID, "walk_dump: NULL batch");
    return b->b.walk_dump(scene, count, walk, list_count, list);
}

int sph_batch_create(const sph_scene_desc_v1 *descs, uint32_t nscenes, sph_batch **out)
{
    return create_object(descs, nscenes, out);
}

int sph_batch_destroy(sph_batch *b)
{
    delete b;
    return SPH_OK;
}

int sph_batch_upload(sph_batch *b, uint32_t scene, const float *pos, const float *vel)
{
    if (!b)
        return fail(SPH_E_INVALID, "batch upload: NULL batch");
    return b->b.upload(scene, pos, vel);
}

int sph_batch_step(sph_batch *b, int nsteps)
{
    if (!b)
        return fail(SPH_E_INVALID, "batch step: NULL batch");
    return b->b.step(nsteps);
}

int sph_batch_sync(sph_batch *b)
{
    if (!b)
        return fail(SPH_E_INVALID, "batch sync: NULL batch");
    return b->b.sync();
}

int sph_batch_download(sph_batch *b, uint32_t scene, float *pos, float *vel, float *rho,
                       float *pressure, uint32_t *ids)
{
    if (!b)
        return fail(SPH_E_INVALID, "batch download: NULL batch");
    return b->b.download(scene, pos, vel, rho, pressure, ids);
}

int sph_batch_set_graph(sph_batch *b, int enable)
{
    if (!b)
        return fail(SPH_E_INVALID, "batch set_graph: NULL batch");
    b->b.use_graph = enable != 0;
    if (!b->b.use_graph)
        b->b.drop_graph();
    return SPH_OK;
}

uint32_t sph_batch_scene_count(const sph_batch *b) { return b ? (uint32_t)b->b.scenes.size() : 0u; }

int sph_batch_profile(sph_batch *b, int nsteps, double *ms, int max_k, int *nk, char *names_buf,
                      size_t names_cap)
{
    if (!b)
        return fail(SPH_E_INVALID, "batch profile: NULL batch");
    return b->b.profile(nsteps, ms, max_k, nk, names_buf, names_cap);
}

int sph_batch_step_timed(sph_batch *b, int nsteps, float *ms)
{
    if (!b)
        return fail(SPH_E_INVALID, "batch step_timed: NULL batch");
    return b->b.step_timed(nsteps, ms);
}

int sph_batch_autotune(sph_batch *b, int trial_steps, sph_tune_v1 *out)
{
    if (!b)
        return fail(SPH_E_INVALID, "batch autotune: NULL batch");
    return b->b.autotune(trial_steps, out);
}

int sph_hip_autotune(sph_ctx *c, int trial_steps, sph_tune_v1 *out)
{
    if (!c)
        return fail(SPH_E_INVALID, "autotune: NULL context");
    return c->b.autotune(trial_steps, out);
}

int sph_batch_autotune_grid(sph_batch *b, int trial_steps, sph_tune_v1 *out)
{
    if (!b)
        return fail(SPH_E_INVALID, "batch autotune: NULL batch");
    return b->b.autotune(trial_steps, out, true);
}

int sph_hip_autotune_grid(sph_ctx *c, int trial_steps, sph_tune_v1 *out)
{
    if (!c)
        return fail(SPH_E_INVALID, "autotune: NULL context");
    return c->b.autotune(trial_steps, out, true);
}

}  // extern "C"
