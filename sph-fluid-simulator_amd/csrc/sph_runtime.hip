// sph_runtime.hip -- host runtime and C ABI of libsph_hip.so.
//
// Owns device memory, streams and HIP graphs; builds the per-scene
// constants (StepParams) from the reference settings; implements the
// stateless drop-in (updateParticles replacement) and the resident
// single-scene context.  No CPU fallback: every compute entry point fails
// with SPH_E_NODEVICE when no HIP device is usable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/sph_hip.h"
#include "sph_kernels.h"
#include "sph_runtime.h"

static_assert(sizeof(sph_particle_v1) == 60, "Particle layout (src/Particle.h:4-10) is 60 bytes");
static_assert(sizeof(sph_settings_v1) == 116, "SPHSettings layout (src/SPHSystem.h:13-22) is 116 bytes");

namespace sph {

// ------------------------------------------------------------ errors

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    if (e == hipErrorOutOfMemory)
        return fail(SPH_E_NOMEM, "%s: %s", what, hipGetErrorString(e));
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
        return fail(SPH_E_NODEVICE, "%s: %s", what, hipGetErrorString(e));
    return fail(SPH_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

const char *last_error() { return g_last_error.c_str(); }

int ensure_device()
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return fail(SPH_E_NODEVICE, "no HIP device available (%s)",
                    e == hipSuccess ? "device count 0" : hipGetErrorString(e));
    return SPH_OK;
}

// ---------------------------------------------------------- settings

// SPHSettings::SPHSettings, src/SPHSystem.cpp:8-26, with the reference's
// promotions spelled out: pow(h, n) is the double pow and the products with
// it are double, rounded to float on assignment.
void settings_init(sph_settings_v1 *s, float mass, float restDensity, float gasConst,
                   float viscosity, float h, float g, float tension)
{
    const float PI = 3.14159265f;  // src/SPHSystem.cpp:6
    std::memset(s, 0, sizeof *s);
    s->mass = mass;
    s->restDensity = restDensity;
    s->gasConstant = gasConst;
    s->viscosity = viscosity;
    s->h = h;
    s->g = g;
    s->tension = tension;
    s->poly6 = (float)(315.0 / ((double)(64.0f * PI) * std::pow((double)h, 9.0)));
    s->spikyGrad = (float)(-45.0 / ((double)PI * std::pow((double)h, 6.0)));
    s->spikyLap = (float)(45.0 / ((double)PI * std::pow((double)h, 6.0)));
    s->h2 = h * h;
    s->selfDens = (float)((double)(mass * s->poly6) * std::pow((double)h, 6.0));
    s->massPoly6Product = mass * s->poly6;
    const float sc = h / 2.f;
    s->sphereScale[0] = sc;
    s->sphereScale[5] = sc;
    s->sphereScale[10] = sc;
    s->sphereScale[15] = 1.0f;
}

void box_reference_cpu(sph_box_v1 *b)
{
    b->lo[0] = -8.0f;
    b->lo[1] = 0.0f;
    b->lo[2] = -8.0f;
    b->hi[0] = 8.0f;
    b->hi[1] = INFINITY;
    b->hi[2] = 8.0f;
    b->restitution = 0.5f;
    b->offset = 0.0001f;
}

void box_reference_gpu(sph_box_v1 *b)
{
    b->lo[0] = -10.0f;
    b->lo[1] = 0.0f;
    b->lo[2] = -10.0f;
    b->hi[0] = 10.0f;
    b->hi[1] = INFINITY;
    b->hi[2] = 10.0f;
    b->restitution = 1.0f;
    b->offset = (float)0.00001;
}

int check_settings(const sph_settings_v1 *s)
{
    if (!s)
        return fail(SPH_E_INVALID, "settings is NULL");
    if (!(s->h > 0.0f) || !std::isfinite(s->h))
        return fail(SPH_E_INVALID, "settings.h must be positive and finite (got %g)", (double)s->h);
    return SPH_OK;
}

int grid_reach(double cell_scale)
{
    if (cell_scale == 0.0 || cell_scale >= 1.0)
        return std::isfinite(cell_scale) ? 1 : 0;
    if (!(cell_scale * kMaxReach >= 1.0))
        return 0;
    int R = (int)std::ceil(1.0 / cell_scale);
    if (cell_scale * R < 1.0)  // 1 / scale rounded down
        ++R;
    return R <= kMaxReach ? R : 0;
}

StepParams make_params(const sph_settings_v1 &s, const sph_box_v1 &box, float dt)
{
    StepParams P;
    std::memset(&P, 0, sizeof P);
    P.reach = 1;
    P.h = s.h;
    P.h2 = s.h2;
    P.mass = s.mass;
    P.mp6 = s.mass * s.poly6;  // local product of src/sph.cpp:30
    P.self_dens = s.selfDens;
    P.gas = s.gasConstant;
    P.rho0 = s.restDensity;
    P.visc = s.viscosity;
    P.spiky_grad = s.spikyGrad;
    P.spiky_lap = s.spikyLap;
    P.g = s.g;
    P.dt = dt;
    P.neg_m_spg = -(s.mass * s.spikyGrad);
    P.mu_m_lap = (s.viscosity * s.mass) * s.spikyLap;
    for (int k = 0; k < 3; ++k) {
        P.wall_lo[k] = box.lo[k] + s.h;
        P.wall_hi[k] = box.hi[k] - s.h;
    }
    P.restitution = box.restitution;
    P.offset = box.offset;
    P.xf_scale = s.h / 2.f;
    return P;
}

// Grid over [lo, hi] per axis with cell edge H >= h (1 + 2^-10) * scale,
// one padding cell on each side, per-axis dims within the packing limits
// and at most max_cells cells in total (H grows until both hold; larger
// cells only add candidates).  Positions outside clamp to the border cells.
// A scale below 1 (at least 1 / kMaxReach) makes cells finer than h: the
// stencil then reaches R = grid_reach(scale) cells, and H * R >= h (1 + 2^-10).
int make_grid(StepParams &P, const double lo[3], const double hi[3], double cell_scale,
              uint64_t max_cells)
{
    const int R = grid_reach(cell_scale);
    if (R == 0)
        return fail(SPH_E_INVALID, "cell scale %g is not 0, >= 1, or in [1/%d, 1)", cell_scale, kMaxReach);
    if (cell_scale == 0.0)
        cell_scale = 1.0;
    P.reach = R;
    double H = (double)P.h * (1.0 + kCellMargin) * cell_scale;
    const int lim[3] = {kGridMaxX, kGridMaxY, kGridMaxZ};
    int dims[3];
    for (int iter = 0; iter < 200; ++iter) {
        uint64_t total = 1;
        bool ok = true;
        for (int k = 0; k < 3; ++k) {
            double ext = hi[k] - lo[k];
            if (!(ext >= 0.0) || !std::isfinite(ext))
                ext = 0.0;
            double d = std::floor(ext / H) + 3.0;
            if (d > (double)lim[k]) {
                ok = false;
                d = lim[k];
            }
            dims[k] = (int)d;
            total *= (uint64_t)dims[k];
        }
        if (ok && total <= max_cells)
            break;
        H *= 1.25;
    }
    for (int k = 0; k < 3; ++k) {
        double o = (std::isfinite(lo[k]) ? lo[k] : 0.0) - H;
        P.grid_o[k] = o;
    }
    P.grid_inv = 1.0 / H;
    P.nx = dims[0];
    P.ny = dims[1];
    P.nz = dims[2];
    return SPH_OK;
}

// ------------------------------------------------------- device buffers

int DevBuf::reserve(size_t want)
{
    if (want <= bytes)
        return SPH_OK;
    // the new block first: when it cannot be had, the old one (and the
    // capacity the caller already relies on) stays as it was
    void *fresh = nullptr;
    hipError_t e = hipMalloc(&fresh, want);
    if (e != hipSuccess)
        return hip_fail(e, "hipMalloc");
    release();
    ptr = fresh;
    bytes = want;
    return SPH_OK;
}

void DevBuf::release()
{
    if (ptr)
        (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int Workspace::reserve_particles(uint32_t n)
{
    int rc;
    // cells holding more than kSmallCell particles: at most n / (kSmallCell + 1)
    const uint32_t cap = n / (kSmallCell + 1u) + 1u;
    if ((rc = big.reserve(4ull * (cap + 1u))))
        return rc;
    if (!unordered.ptr) {
        if ((rc = unordered.reserve(8)))
            return rc;
        if (hipMemset(unordered.ptr, 0, 8) != hipSuccess)
            return fail(SPH_E_HIP, "clearing the unordered-cell counter");
    }
    big_cap = (uint32_t)(big.bytes / 4 - 1);
    if ((rc = posA.reserve(sizeof(float4) * (size_t)n)) || (rc = velA.reserve(sizeof(float4) * (size_t)n)) ||
        (rc = posB.reserve(sizeof(float4) * (size_t)n)) || (rc = velB.reserve(sizeof(float4) * (size_t)n)) ||
        (rc = cellcA.reserve(4ull * n)) || (rc = rankA.reserve(4ull * n)) ||
        (rc = cellcB.reserve(4ull * n)) || (rc = perm.reserve(4ull * n)) ||
        (rc = rhoB.reserve(4ull * n)) || (rc = posR.reserve(sizeof(float4) * (size_t)n)) ||
        (rc = velR.reserve(sizeof(float4) * (size_t)n)) || (rc = ncount.reserve(4ull * n)) ||
        (rc = nlist.reserve(4ull * kListMax * n)))
        return rc;
    return SPH_OK;
}

int Workspace::reserve_cells(uint64_t ncells)
{
    uint64_t padded = (ncells + kScanChunk - 1) / kScanChunk * kScanChunk;
    if (padded == 0)
        padded = kScanChunk;
    if (padded > (1ull << 31))
        return fail(SPH_E_LIMIT, "grid of %llu cells exceeds the 2^31 limit", (unsigned long long)ncells);
    int rc;
    if ((rc = counts.reserve(4 * padded)) || (rc = cstart.reserve(4 * (padded + 1))) ||
        (rc = partials.reserve(4 * (padded / kScanChunk))))
        return rc;
    ncells_padded = (uint32_t)padded;
    return SPH_OK;
}

int Workspace::set_scenes(const std::vector<StepParams> &scenes, hipStream_t st)
{
    std::vector<WaveTask> t;
    for (uint32_t k = 0; k < scenes.size(); ++k)
        for (uint32_t f = scenes[k].p_begin; f < scenes[k].p_end; f += 64u)
            t.push_back(WaveTask{k, f});
    int rc;
    if ((rc = params.reserve(sizeof(StepParams) * std::max<size_t>(scenes.size(), 1))) ||
        (rc = tasks.reserve(sizeof(WaveTask) * std::max<size_t>(t.size(), 1))))
        return rc;
    const size_t nw = std::max<size_t>(t.size(), 1);
    size_t nslots = 0;
    for (const StepParams &sp : scenes)
        nslots = std::max<size_t>(nslots, sp.p_end);
    const uint32_t padded = (uint32_t)((nw + kScanChunk - 1) / kScanChunk * kScanChunk);
    size_t cap = 0;  // every scene's wave tasks plus its part room
    for (const StepParams &sp : scenes) {
        const uint64_t waves = ((uint64_t)sp.p_end - sp.p_begin + 63u) / 64u;
        cap += waves + part_room(waves);
    }
    cap = std::max<size_t>(cap, nw);
    if (cap > 0xFFFFFFFFull)
        return fail(SPH_E_LIMIT, "part plan: %zu part tasks", cap);
    const bool fresh_counters = padded > parts_padded || pdone.bytes < 4 * nw ||
                                ptotal.bytes < 8 * std::max<size_t>(scenes.size(), 1);
    if ((rc = pnparts.reserve(4ull * padded)) || (rc = poff.reserve(4ull * (padded + 1))) ||
        (rc = ppartials.reserve(4ull * (padded / kScanChunk))) || (rc = pmeasure.reserve(4ull * nw)) ||
        (rc = ptotal.reserve(8ull * std::max<size_t>(scenes.size(), 1))) || (rc = ptasks.reserve(4ull * cap)) ||
        (rc = pqueue.reserve(16)) ||
        (rc = pbuf.reserve(sizeof(float4) * 64ull * cap)) || (rc = pdone.reserve(4ull * nw)) ||
        (rc = pwalk.reserve(2ull * sizeof(int4) * nw)) || (rc = pown.reserve(4ull * std::max<size_t>(nslots, 1))))
        return rc;
    if (fresh_counters) {  // zero outside use from here on: the step keeps them so
        launch_fill_u32(st, (uint32_t *)pnparts.ptr, 0u, pnparts.bytes / 4);
        launch_fill_u32(st, (uint32_t *)pdone.ptr, 0u, pdone.bytes / 4);
        launch_fill_u32(st, (uint32_t *)ptotal.ptr, 0u, ptotal.bytes / 4);
        parts_padded = (uint32_t)(pnparts.bytes / 4 / kScanChunk * kScanChunk);
    }
    hipError_t e = hipSuccess;
    if (!scenes.empty())
        e = hipMemcpyAsync(params.ptr, scenes.data(), sizeof(StepParams) * scenes.size(),
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess && !t.empty())
        e = hipMemcpyAsync(tasks.ptr, t.data(), sizeof(WaveTask) * t.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);  // the host vectors go out of scope
    if (e != hipSuccess)
        return hip_fail(e, "scene table upload");
    nwaves = (uint32_t)t.size();
    nscenes = (uint32_t)scenes.size();
    part_cap = (uint32_t)cap;
    return SPH_OK;
}

ParticleArrays Workspace::particles() const
{
    ParticleArrays a;
    a.posA = (float4 *)posA.ptr;
    a.velA = (float4 *)velA.ptr;
    a.posB = (float4 *)posB.ptr;
    a.velB = (float4 *)velB.ptr;
    a.cellcA = (uint32_t *)cellcA.ptr;
    a.rankA = (uint32_t *)rankA.ptr;
    a.cellcB = (uint32_t *)cellcB.ptr;
    a.perm = (uint32_t *)perm.ptr;
    a.rhoB = (float *)rhoB.ptr;
    a.posR = (float4 *)posR.ptr;
    a.velR = (float4 *)velR.ptr;
    return a;
}

CellArrays Workspace::cells() const
{
    CellArrays c;
    c.counts = (uint32_t *)counts.ptr;
    c.cstart = (uint32_t *)cstart.ptr;
    c.partials = (uint32_t *)partials.ptr;
    c.big = (uint32_t *)big.ptr;
    c.big_cap = big_cap;
    c.ncells_padded = ncells_padded;
    c.unordered = (unsigned long long *)unordered.ptr;
    return c;
}

SceneSet Workspace::scene_set() const
{
    SceneSet S;
    S.params = (const StepParams *)params.ptr;
    S.tasks = (const WaveTask *)tasks.ptr;
    S.nwaves = nwaves;
    return S;
}

PartPlan Workspace::part_plan() const
{
    PartPlan p;
    p.nparts = (uint32_t *)pnparts.ptr;
    p.off = (uint32_t *)poff.ptr;
    p.scan_partials = (uint32_t *)ppartials.ptr;
    p.padded = (nwaves + kScanChunk - 1) / kScanChunk * kScanChunk;
    p.measure = (uint32_t *)pmeasure.ptr;
    p.total = (unsigned long long *)ptotal.ptr;
    p.nscenes = nscenes;
    p.cap = std::max(part_cap, 1u);
    p.tasks = (uint32_t *)ptasks.ptr;
    p.queue = (uint32_t *)pqueue.ptr;
    p.buf = (float4 *)pbuf.ptr;
    p.done = (uint32_t *)pdone.ptr;
    p.walk = (int4 *)pwalk.ptr;
    p.own = (uint32_t *)pown.ptr;
    return p;
}

NeighborLists Workspace::lists() const
{
    NeighborLists l;
    // the slots both arrays cover (they grow together)
    const size_t slots = std::min<size_t>(ncount.bytes / 4, nlist.bytes / (4ull * kListMax));
    l.list = (uint32_t *)nlist.ptr;
    l.count = (uint32_t *)ncount.ptr;
    l.stride = (uint32_t)slots;
    if (slots == 0)
        l.list = l.count = nullptr;
    return l;
}

void Workspace::release_all()
{
    for (DevBuf *b : {&posA, &velA, &posB, &velB, &cellcA, &rankA, &cellcB, &perm, &rhoB, &posR, &velR,
                      &counts, &cstart, &partials, &big, &force, &params, &tasks, &pnparts, &poff,
                      &ppartials, &pmeasure, &ptotal, &ptasks, &pqueue, &pbuf, &pdone, &pwalk, &pown,
                      &nlist, &ncount, &unordered})
        b->release();
    ncells_padded = 0;
    big_cap = 0;
    nwaves = 0;
    nscenes = 0;
    parts_padded = 0;
}

void enqueue_step(hipStream_t st, Workspace &w, float4 *force_out, bool next_keys, StepTimer *timer)
{
    ParticleArrays a = w.particles();
    CellArrays c = w.cells();
    SceneSet S = w.scene_set();
    if (timer)
        timer->mark(st, nullptr);
    launch_scan(st, c, false);
    if (timer)
        timer->mark(st, "scan");
    launch_scatter(st, a, c, S);
    launch_sort_big_cells(st, a, c);
    if (timer)
        timer->mark(st, "scatter");
    launch_reorder(st, a, c, S);
    const PartPlan pp = w.part_plan();
    launch_plan_parts(st, a, c, S, pp);
    if (timer)
        timer->mark(st, "reorder");
    const NeighborLists nl = w.lists();
    launch_density(st, a, c, S, nullptr, &pp, w.stencil_block, &nl);
    if (timer)
        timer->mark(st, "density");
    launch_force_integrate(st, a, c, S, force_out, next_keys, &pp, w.stencil_block, &nl);
    if (timer)
        timer->mark(st, "force_integrate");
}

// ------------------------------------------------------ stateless path

namespace {

struct Stateless {
    std::mutex mu;
    bool init = false;
    hipStream_t stream = nullptr;
    Workspace w;
    DevBuf aos_in, aos_out, xf, bbox, counts16, cstart16, partials16, perm16;
};

Stateless &stateless()
{
    static Stateless *s = new Stateless();  // never destroyed: no HIP calls at exit
    return *s;
}

int stateless_init(Stateless &S)
{
    if (S.init)
        return SPH_OK;
    int rc = ensure_device();
    if (rc)
        return rc;
    hipError_t e = hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking);
    if (e != hipSuccess)
        return hip_fail(e, "hipStreamCreate");
    S.init = true;
    return SPH_OK;
}

double from_ordered(int32_t o) { return (double)__builtin_bit_cast(float, o >= 0 ? o : (int32_t)(o ^ 0x7FFFFFFF)); }

// Uploads AoS particles, builds the grid from their bounding box, computes
// keys and runs the scan/scatter/reorder/density passes.  Leaves the
// workspace ready for the force pass.
int stateless_prepare(Stateless &S, const sph_particle_v1 *particles, uint32_t n,
                      const sph_settings_v1 *settings, const sph_box_v1 &box, float dt,
                      StepParams &P)
{
    int rc;
    Workspace &w = S.w;
    if ((rc = w.reserve_particles(n)) || (rc = S.aos_in.reserve(sizeof(sph_particle_v1) * (size_t)n)) ||
        (rc = S.bbox.reserve(64)))
        return rc;
    hipStream_t st = S.stream;
    hipError_t e = hipMemcpyAsync(S.aos_in.ptr, particles, sizeof(sph_particle_v1) * (size_t)n,
                                  hipMemcpyHostToDevice, st);
    if (e != hipSuccess)
        return hip_fail(e, "upload particles");
    ParticleArrays a = w.particles();
    launch_unpack_aos(st, S.aos_in.ptr, n, a.posA, a.velA);
    launch_fill_u32(st, (uint32_t *)S.bbox.ptr, 0x7FFFFFFFu, 3);      // min: INT32_MAX
    launch_fill_u32(st, (uint32_t *)S.bbox.ptr + 3, 0x80000000u, 3);  // max: INT32_MIN
    launch_bbox(st, a.posA, n, (int32_t *)S.bbox.ptr);
    int32_t bb[6];
    e = hipMemcpyAsync(bb, S.bbox.ptr, sizeof bb, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, "bbox readback");
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        if (bb[k] == INT32_MAX) {  // no finite coordinate on this axis
            lo[k] = 0.0;
            hi[k] = 0.0;
        } else {
            lo[k] = from_ordered(bb[k]);
            hi[k] = from_ordered(bb[k + 3]);
        }
    }
    P = make_params(*settings, box, dt);
    uint64_t max_cells = std::max<uint64_t>(8ull * n, 1ull << 18);
    make_grid(P, lo, hi, 1.0, max_cells);
    P.cell_base = 0;
    P.p_begin = 0;
    P.p_end = n;
    if ((rc = w.reserve_cells((uint64_t)P.nx * P.ny * P.nz)) || (rc = w.set_scenes({P}, st)))
        return rc;
    CellArrays c = w.cells();
    SceneSet set = w.scene_set();
    launch_fill_u32(st, c.counts, 0u, c.ncells_padded);
    launch_init_keys(st, a, c, set);
    launch_scan(st, c, false);
    launch_scatter(st, a, c, set);
    launch_sort_big_cells(st, a, c);
    launch_reorder(st, a, c, set);
    return SPH_OK;
}

int cells16_reserve(Stateless &S)
{
    int rc;
    if ((rc = S.counts16.reserve(4ull * 65536)) || (rc = S.cstart16.reserve(4ull * 65537)) ||
        (rc = S.partials16.reserve(4ull * 65536 / kScanChunk)))
        return rc;
    return SPH_OK;
}

CellArrays cells16(Stateless &S)
{
    CellArrays c;
    c.counts = (uint32_t *)S.counts16.ptr;
    c.cstart = (uint32_t *)S.cstart16.ptr;
    c.partials = (uint32_t *)S.partials16.ptr;
    c.big = nullptr;
    c.big_cap = 0;
    c.ncells_padded = 65536;
    return c;
}

}  // namespace

int update_particles(sph_particle_v1 *particles, float *transforms, size_t n_,
                     const sph_settings_v1 *settings, float dt, const sph_box_v1 *box_in)
{
    int rc;
    if (n_ == 0)
        return SPH_OK;
    if (!particles)
        return fail(SPH_E_INVALID, "particles is NULL");
    if ((rc = check_settings(settings)))
        return rc;
    if (n_ > (1u << 30))
        return fail(SPH_E_LIMIT, "%zu particles exceed the 2^30 limit", n_);
    const uint32_t n = (uint32_t)n_;
    sph_box_v1 box;
    if (box_in)
        box = *box_in;
    else
        box_reference_cpu(&box);

    Stateless &S = stateless();
    std::lock_guard<std::mutex> lock(S.mu);
    if ((rc = stateless_init(S)))
        return rc;
    StepParams P;
    if ((rc = stateless_prepare(S, particles, n, settings, box, dt, P)))
        return rc;
    Workspace &w = S.w;
    if ((rc = w.force.reserve(sizeof(float4) * (size_t)n)) || (rc = cells16_reserve(S)) ||
        (rc = S.perm16.reserve(4ull * n)) || (rc = S.aos_out.reserve(sizeof(sph_particle_v1) * (size_t)n)) ||
        (transforms && (rc = S.xf.reserve(64ull * n))))
        return rc;
    hipStream_t st = S.stream;
    ParticleArrays a = w.particles();
    CellArrays c = w.cells();
    const SceneSet set = w.scene_set();
    const PartPlan pp = w.part_plan();
    launch_plan_parts(st, a, c, set, pp);
    const NeighborLists nl = w.lists();
    launch_density(st, a, c, set, nullptr, &pp, kStencilBlock, &nl);
    launch_force_integrate(st, a, c, set, (float4 *)w.force.ptr, /*next_keys=*/false, &pp, kStencilBlock, &nl);
    CellArrays c16 = cells16(S);
    launch_fill_u32(st, c16.counts, 0u, 65536);
    launch_h16_rank(st, a, c16, n);
    launch_scan(st, c16, true);
    launch_h16_scatter(st, a, c16, n, (uint32_t *)S.perm16.ptr);
    launch_pack_out(st, a, c16, (const uint32_t *)S.perm16.ptr, (const float4 *)w.force.ptr,
                    S.aos_in.ptr, n, P.xf_scale, S.aos_out.ptr, transforms ? (float *)S.xf.ptr : nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return hip_fail(e, "kernel launch");
    e = hipMemcpyAsync(particles, S.aos_out.ptr, sizeof(sph_particle_v1) * (size_t)n,
                       hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && transforms)
        e = hipMemcpyAsync(transforms, S.xf.ptr, 64ull * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, "download particles");
    return SPH_OK;
}

int neighbors(const sph_particle_v1 *particles, size_t n_, const sph_settings_v1 *settings,
              uint64_t *offsets, uint32_t *idx, size_t cap, size_t *total)
{
    int rc;
    if (total)
        *total = 0;
    if (!offsets)
        return fail(SPH_E_INVALID, "offsets is NULL");
    offsets[0] = 0;
    if (n_ == 0)
        return SPH_OK;
    if (!particles)
        return fail(SPH_E_INVALID, "particles is NULL");
    if ((rc = check_settings(settings)))
        return rc;
    if (n_ > (1u << 30))
        return fail(SPH_E_LIMIT, "%zu particles exceed the 2^30 limit", n_);
    const uint32_t n = (uint32_t)n_;
    sph_box_v1 box;
    box_reference_cpu(&box);
    Stateless &S = stateless();
    std::lock_guard<std::mutex> lock(S.mu);
    if ((rc = stateless_init(S)))
        return rc;
    StepParams P;
    if ((rc = stateless_prepare(S, particles, n, settings, box, 0.0f, P)))
        return rc;
    Workspace &w = S.w;
    hipStream_t st = S.stream;
    ParticleArrays a = w.particles();
    CellArrays c = w.cells();
    // count pass (per sorted slot), then rows in sorted-slot order
    DevBuf dcount, doff, dout;
    if ((rc = dcount.reserve(4ull * n)) || (rc = doff.reserve(8ull * (n + 1))))
        return rc;
    SceneSet set = w.scene_set();
    NeighborDump d{(uint32_t *)dcount.ptr, nullptr, nullptr};
    launch_density(st, a, c, set, &d);
    std::vector<uint32_t> cnt(n), ids(n);
    std::vector<float4> velB(n);
    hipError_t e = hipMemcpyAsync(cnt.data(), dcount.ptr, 4ull * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(velB.data(), a.velB, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        dcount.release();
        doff.release();
        return hip_fail(e, "neighbour count");
    }
    std::vector<uint64_t> soff(n + 1);
    soff[0] = 0;
    for (uint32_t s = 0; s < n; ++s)
        soff[s + 1] = soff[s] + cnt[s];
    const uint64_t tot = soff[n];
    // per-input-index offsets
    for (uint32_t s = 0; s < n; ++s)
        ids[s] = __builtin_bit_cast(uint32_t, velB[s].w);
    std::vector<uint64_t> rowlen(n + 1, 0);
    for (uint32_t s = 0; s < n; ++s)
        rowlen[ids[s] + 1] = cnt[s];
    for (uint32_t i = 0; i < n; ++i)
        offsets[i + 1] = offsets[i] + rowlen[i + 1];
    if (total)
        *total = (size_t)tot;
    if (!idx || cap < tot || tot == 0) {
        dcount.release();
        doff.release();
        return SPH_OK;
    }
    if ((rc = dout.reserve(4ull * tot))) {
        dcount.release();
        doff.release();
        return rc;
    }
    e = hipMemcpyAsync(doff.ptr, soff.data(), 8ull * (n + 1), hipMemcpyHostToDevice, st);
    if (e != hipSuccess)
        return hip_fail(e, "neighbour offsets");
    NeighborDump d2{(uint32_t *)dcount.ptr, (const uint64_t *)doff.ptr, (uint32_t *)dout.ptr};
    launch_density(st, a, c, set, &d2);
    std::vector<uint32_t> rows(tot);
    e = hipMemcpyAsync(rows.data(), dout.ptr, 4ull * tot, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    dcount.release();
    doff.release();
    dout.release();
    if (e != hipSuccess)
        return hip_fail(e, "neighbour rows");
    for (uint32_t s = 0; s < n; ++s) {
        uint32_t i = ids[s];
        uint32_t *dst = idx + offsets[i];
        std::copy(rows.begin() + (ptrdiff_t)soff[s], rows.begin() + (ptrdiff_t)soff[s + 1], dst);
        std::sort(dst, dst + cnt[s]);
    }
    return SPH_OK;
}

// --------------------------------------------------------- copy bench

int copy_bench(size_t bytes, int iters, double *ms)
{
    int rc;
    if (!ms || iters <= 0 || bytes < 16)
        return fail(SPH_E_INVALID, "copy_bench: bad arguments");
    if ((rc = ensure_device()))
        return rc;
    size_t n16 = bytes / 16;
    DevBuf a, b;
    if ((rc = a.reserve(n16 * 16)) || (rc = b.reserve(n16 * 16)))
        return rc;
    hipStream_t st;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess)
        return hip_fail(e, "hipStreamCreate");
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    launch_copy(st, (const uint4 *)a.ptr, (uint4 *)b.ptr, n16);
    (void)hipEventRecord(e0, st);
    for (int i = 0; i < iters; ++i)
        launch_copy(st, (const uint4 *)a.ptr, (uint4 *)b.ptr, n16);
    (void)hipEventRecord(e1, st);
    e = hipEventSynchronize(e1);
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = (double)t / iters;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    if (e != hipSuccess)
        return hip_fail(e, "copy bench");
    return SPH_OK;
}

}  // namespace sph

// ================================================================ C ABI

using namespace sph;

extern "C" {

int sph_abi_version(void) { return SPH_HIP_ABI_VERSION; }

const char *sph_hip_last_error(void) { return last_error(); }

int sph_settings_init(sph_settings_v1 *s, float mass, float restDensity, float gasConst,
                      float viscosity, float h, float g, float tension)
{
    if (!s)
        return fail(SPH_E_INVALID, "settings is NULL");
    settings_init(s, mass, restDensity, gasConst, viscosity, h, g, tension);
    return SPH_OK;
}

int sph_box_reference_cpu(sph_box_v1 *b)
{
    if (!b)
        return fail(SPH_E_INVALID, "box is NULL");
    box_reference_cpu(b);
    return SPH_OK;
}

int sph_box_reference_gpu(sph_box_v1 *b)
{
    if (!b)
        return fail(SPH_E_INVALID, "box is NULL");
    box_reference_gpu(b);
    return SPH_OK;
}

uint32_t sph_get_hash(int32_t cx, int32_t cy, int32_t cz)
{
    return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) %
           262144u;
}

int sph_classify_host(const float pos[3], float h, int32_t out[5])
{
    if (!pos || !out || !(h > 0.0f))
        return fail(SPH_E_INVALID, "classify: bad arguments");
    StepParams P;
    std::memset(&P, 0, sizeof P);
    P.h = h;
    P.nx = P.ny = P.nz = 1;
    P.reach = 1;
    P.grid_inv = 1.0;
    Classified c = classify(pos[0], pos[1], pos[2], P);
    float qx = pos[0] / h, qy = pos[1] / h, qz = pos[2] / h;
    out[0] = (int32_t)c.h16;
    out[1] = (c.cellc & kSlowBit) ? 1 : 0;
    out[2] = trunc_to_int(qx);
    out[3] = trunc_to_int(qy);
    out[4] = trunc_to_int(qz);
    return SPH_OK;
}

int sph_hip_update_particles(sph_particle_v1 *particles, float *transforms, size_t n,
                             const sph_settings_v1 *settings, float dt)
{
    return update_particles(particles, transforms, n, settings, dt, nullptr);
}

int sph_hip_update_particles_box(sph_particle_v1 *particles, float *transforms, size_t n,
                                 const sph_settings_v1 *settings, float dt, const sph_box_v1 *box)
{
    if (!box)
        return fail(SPH_E_INVALID, "box is NULL");
    return update_particles(particles, transforms, n, settings, dt, box);
}

int sph_hip_neighbors(const sph_particle_v1 *particles, size_t n, const sph_settings_v1 *settings,
                      uint64_t *offsets, uint32_t *idx, size_t cap, size_t *total)
{
    return neighbors(particles, n, settings, offsets, idx, cap, total);
}

int sph_hip_copy_bench(size_t bytes, int iters, double *ms) { return copy_bench(bytes, iters, ms); }

}  // extern "C"
