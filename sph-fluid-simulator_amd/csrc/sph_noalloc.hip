// sph_noalloc.hip -- the stateless drop-in step without device allocation.
//
// sph_hip_update_particles_noalloc runs exactly the kernels of
// sph_hip_update_particles_box, but every device buffer it needs (the AoS
// staging copies, the particle arrays, the grid cells, the part plan and
// the 16-bit bucket sort of the output) is carved from one fixed array in
// device memory, `scratch`, instead of being allocated.  It is for callers
// that cannot call hipMalloc (harnesses that forbid it, loops that must not
// allocate).  No counterpart exists in the reference.
//
// The array is a file-scope __device__ variable: every GPU the library runs
// on holds its own copy, kScratchBytes long, until the process ends.  In
// host code its name is only a stand-in for the symbol, never a device
// address; the address is looked up once per device with
// hipGetSymbolAddress (with that device current) and kept in g_scratch.
// Every call on one GPU uses the same array and finds the previous call's
// contents there, so everything that must start at zero is cleared by a
// fill kernel at the start of each call.  Calls are serialised by a
// process-wide lock and return only after their device work has finished,
// so two calls never overlap; device work of other code that wrote into
// this array while a call runs would race with it (nothing in the library
// does).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/sph_hip.h"
#include "sph_kernels.h"
#include "sph_runtime.h"

namespace sph {
namespace {

constexpr size_t kScratchBytes = 64ull << 20;
constexpr uint32_t kMaxParticles = SPH_NOALLOC_MAX_PARTICLES;
constexpr int kMaxDevices = 64;
constexpr size_t kAlign = 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices];  // device address of `scratch`, per device

// Offsets of every buffer of one call inside the scratch array.
struct Layout {
    size_t aos_in, aos_out, xf;
    size_t posA, velA, posB, velB, cellcA, rankA, cellcB, perm, rhoB, posR, velR;
    size_t counts, cstart, partials, big;
    size_t force, perm16, counts16, cstart16, partials16;
    size_t params, tasks;
    size_t nparts, off, ppartials, measure, total, ptasks, queue, buf, done, walk, own;
    size_t nlist, ncount;
    size_t bytes;  // end of the last buffer
    uint32_t ncells_padded, big_cap, nwaves, waves_padded, cap;
};

struct Carver {
    size_t at = 0;
    size_t take(size_t n)
    {
        const size_t o = at;
        at = (at + std::max<size_t>(n, 1) + kAlign - 1) / kAlign * kAlign;
        return o;
    }
};

// Grid cells the drop-in may use for n particles (make_grid's cap).
uint64_t max_cells_for(uint32_t n) { return std::max<uint64_t>(8ull * n, 1ull << 18); }

Layout make_layout(uint32_t n, uint64_t ncells)
{
    Layout L{};
    Carver c;
    L.aos_in = c.take(sizeof(sph_particle_v1) * (size_t)n);
    L.aos_out = c.take(sizeof(sph_particle_v1) * (size_t)n);
    L.xf = c.take(64ull * n);
    L.posA = c.take(16ull * n);
    L.velA = c.take(16ull * n);
    L.posB = c.take(16ull * n);
    L.velB = c.take(16ull * n);
    L.cellcA = c.take(4ull * n);
    L.rankA = c.take(4ull * n);
    L.cellcB = c.take(4ull * n);
    L.perm = c.take(4ull * n);
    L.rhoB = c.take(4ull * n);
    L.posR = c.take(16ull * n);
    L.velR = c.take(16ull * n);
    uint64_t padded = (ncells + kScanChunk - 1) / kScanChunk * kScanChunk;
    if (padded == 0)
        padded = kScanChunk;
    L.ncells_padded = (uint32_t)padded;
    L.counts = c.take(4 * padded);
    L.cstart = c.take(4 * (padded + 1));
    L.partials = c.take(4 * (padded / kScanChunk));
    L.big_cap = n / (kSmallCell + 1u) + 1u;
    L.big = c.take(4ull * (L.big_cap + 1u));
    L.force = c.take(16ull * n);
    L.perm16 = c.take(4ull * n);
    L.counts16 = c.take(4ull * 65536);
    L.cstart16 = c.take(4ull * 65537);
    L.partials16 = c.take(4ull * 65536 / kScanChunk);
    L.params = c.take(sizeof(StepParams));
    L.nwaves = (n + 63u) / 64u;
    L.tasks = c.take(sizeof(WaveTask) * (size_t)L.nwaves);
    L.waves_padded = (L.nwaves + kScanChunk - 1) / kScanChunk * kScanChunk;
    L.cap = (uint32_t)(L.nwaves + part_room(L.nwaves));
    L.nparts = c.take(4ull * L.waves_padded);
    L.off = c.take(4ull * (L.waves_padded + 1));
    L.ppartials = c.take(4ull * (L.waves_padded / kScanChunk));
    L.measure = c.take(4ull * L.nwaves);
    L.total = c.take(8);
    L.ptasks = c.take(4ull * L.cap);
    L.queue = c.take(16);
    L.buf = c.take(sizeof(float4) * 64ull * L.cap);
    L.done = c.take(4ull * L.nwaves);
    L.walk = c.take(2ull * sizeof(int4) * L.nwaves);
    L.own = c.take(4ull * n);
    L.nlist = c.take(4ull * kListMax * n);
    L.ncount = c.take(4ull * n);
    L.bytes = c.at;
    return L;
}

// One wave task per 64 slots of the single scene.
__global__ void k_single_scene_tasks(WaveTask *__restrict__ tasks, uint32_t nwaves)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nwaves)
        tasks[w] = WaveTask{0u, 64u * w};
}

double from_ordered(int32_t o) { return (double)__builtin_bit_cast(float, o >= 0 ? o : (int32_t)(o ^ 0x7FFFFFFF)); }

// Device address of this device's scratch array, looked up on first use.
int scratch_of_current_device(unsigned char **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices)
        return fail(SPH_E_LIMIT, "device %d beyond the %d devices the scratch table holds", dev, kMaxDevices);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch[dev] = (unsigned char *)p;
    }
    *out = g_scratch[dev];
    return SPH_OK;
}

}  // namespace

int update_particles_noalloc(sph_particle_v1 *particles, float *transforms, size_t n_,
                             const sph_settings_v1 *settings, float dt, const sph_box_v1 *box)
{
    int rc;
    if (n_ == 0)
        return SPH_OK;
    if (!particles)
        return fail(SPH_E_INVALID, "particles is NULL");
    if (!box)
        return fail(SPH_E_INVALID, "box is NULL");
    if ((rc = check_settings(settings)))
        return rc;
    if (n_ > kMaxParticles)
        return fail(SPH_E_LIMIT, "%zu particles exceed the %u the no-allocation scratch holds", n_,
                    kMaxParticles);
    const uint32_t n = (uint32_t)n_;
    if ((rc = ensure_device()))
        return rc;
    std::lock_guard<std::mutex> lock(g_mu);
    unsigned char *base = nullptr;
    if ((rc = scratch_of_current_device(&base)))
        return rc;
    const Layout L = make_layout(n, max_cells_for(n));
    if (L.bytes > kScratchBytes)  // kMaxParticles is chosen so that this never happens
        return fail(SPH_E_LIMIT, "no-allocation layout of %u particles needs %zu bytes of %zu", n, L.bytes,
                    kScratchBytes);
    auto at = [&](size_t off) { return (void *)(base + off); };
    const hipStream_t st = nullptr;

    ParticleArrays a;
    a.posA = (float4 *)at(L.posA);
    a.velA = (float4 *)at(L.velA);
    a.posB = (float4 *)at(L.posB);
    a.velB = (float4 *)at(L.velB);
    a.cellcA = (uint32_t *)at(L.cellcA);
    a.rankA = (uint32_t *)at(L.rankA);
    a.cellcB = (uint32_t *)at(L.cellcB);
    a.perm = (uint32_t *)at(L.perm);
    a.rhoB = (float *)at(L.rhoB);
    a.posR = (float4 *)at(L.posR);
    a.velR = (float4 *)at(L.velR);

    hipError_t e = hipMemcpyAsync(at(L.aos_in), particles, sizeof(sph_particle_v1) * (size_t)n,
                                  hipMemcpyHostToDevice, st);
    if (e != hipSuccess)
        return hip_fail(e, "upload particles");
    launch_unpack_aos(st, at(L.aos_in), n, a.posA, a.velA);
    // the bounding box borrows the first words of counts16, which is
    // cleared again before the output sort uses it
    int32_t *bbox = (int32_t *)at(L.counts16);
    launch_fill_u32(st, (uint32_t *)bbox, 0x7FFFFFFFu, 3);
    launch_fill_u32(st, (uint32_t *)bbox + 3, 0x80000000u, 3);
    launch_bbox(st, a.posA, n, bbox);
    int32_t bb[6];
    e = hipMemcpyAsync(bb, bbox, sizeof bb, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, "bbox readback");
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        if (bb[k] == INT32_MAX) {
            lo[k] = 0.0;
            hi[k] = 0.0;
        } else {
            lo[k] = from_ordered(bb[k]);
            hi[k] = from_ordered(bb[k + 3]);
        }
    }
    StepParams P = make_params(*settings, *box, dt);
    make_grid(P, lo, hi, 1.0, max_cells_for(n));
    P.cell_base = 0;
    P.p_begin = 0;
    P.p_end = n;
    const uint64_t ncells = (uint64_t)P.nx * P.ny * P.nz;
    if (ncells > L.ncells_padded)
        return fail(SPH_E_LIMIT, "grid of %llu cells beyond the %u laid out", (unsigned long long)ncells,
                    L.ncells_padded);

    CellArrays c;
    c.counts = (uint32_t *)at(L.counts);
    c.cstart = (uint32_t *)at(L.cstart);
    c.partials = (uint32_t *)at(L.partials);
    c.big = (uint32_t *)at(L.big);
    c.big_cap = L.big_cap;
    // scan only the cells this grid uses (as the allocating path sizes them)
    c.ncells_padded = (uint32_t)((ncells + kScanChunk - 1) / kScanChunk * kScanChunk);

    StepParams *dparams = (StepParams *)at(L.params);
    e = hipMemcpyAsync(dparams, &P, sizeof P, hipMemcpyHostToDevice, st);
    if (e != hipSuccess)
        return hip_fail(e, "scene constants");
    WaveTask *dtasks = (WaveTask *)at(L.tasks);
    hipLaunchKernelGGL(k_single_scene_tasks, dim3(blocks_for(L.nwaves)), dim3(kBlock), 0, st, dtasks, L.nwaves);
    SceneSet set;
    set.params = dparams;
    set.tasks = dtasks;
    set.nwaves = L.nwaves;

    PartPlan pp;
    pp.nparts = (uint32_t *)at(L.nparts);
    pp.off = (uint32_t *)at(L.off);
    pp.scan_partials = (uint32_t *)at(L.ppartials);
    pp.padded = L.waves_padded;
    pp.measure = (uint32_t *)at(L.measure);
    pp.total = (unsigned long long *)at(L.total);
    pp.nscenes = 1;
    pp.cap = L.cap;
    pp.tasks = (uint32_t *)at(L.ptasks);
    pp.queue = (uint32_t *)at(L.queue);
    pp.buf = (float4 *)at(L.buf);
    pp.done = (uint32_t *)at(L.done);
    pp.walk = (int4 *)at(L.walk);
    pp.own = (uint32_t *)at(L.own);
    // the previous call's contents are still there: clear what the passes
    // expect at zero (cell counts, part counts, finished-part counters, the
    // per-scene total; the queue counters are cleared by the plan itself)
    launch_fill_u32(st, c.counts, 0u, c.ncells_padded);
    launch_fill_u32(st, pp.nparts, 0u, L.waves_padded);
    launch_fill_u32(st, pp.done, 0u, L.nwaves);
    launch_fill_u32(st, (uint32_t *)pp.total, 0u, 2);

    launch_init_keys(st, a, c, set);
    launch_scan(st, c, false);
    launch_scatter(st, a, c, set);
    launch_sort_big_cells(st, a, c);
    launch_reorder(st, a, c, set);
    launch_plan_parts(st, a, c, set, pp);
    float4 *force = (float4 *)at(L.force);
    const NeighborLists nl{(uint32_t *)at(L.nlist), (uint32_t *)at(L.ncount), n};
    launch_density(st, a, c, set, nullptr, &pp, kStencilBlock, &nl);
    launch_force_integrate(st, a, c, set, force, /*next_keys=*/false, &pp, kStencilBlock, &nl);

    CellArrays c16;
    c16.counts = (uint32_t *)at(L.counts16);
    c16.cstart = (uint32_t *)at(L.cstart16);
    c16.partials = (uint32_t *)at(L.partials16);
    c16.big = nullptr;
    c16.big_cap = 0;
    c16.ncells_padded = 65536;
    uint32_t *perm16 = (uint32_t *)at(L.perm16);
    launch_fill_u32(st, c16.counts, 0u, 65536);
    launch_h16_rank(st, a, c16, n);
    launch_scan(st, c16, true);
    launch_h16_scatter(st, a, c16, n, perm16);
    void *aos_out = at(L.aos_out);
    float *xf = transforms ? (float *)at(L.xf) : nullptr;
    launch_pack_out(st, a, c16, perm16, force, at(L.aos_in), n, P.xf_scale, aos_out, xf);
    e = hipGetLastError();
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);  // the scene constants' copy reads P
        return hip_fail(e, "kernel launch");
    }
    e = hipMemcpyAsync(particles, aos_out, sizeof(sph_particle_v1) * (size_t)n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && transforms)
        e = hipMemcpyAsync(transforms, xf, 64ull * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, "download particles");
    return SPH_OK;
}

}  // namespace sph

extern "C" int sph_hip_update_particles_noalloc(sph_particle_v1 *particles, float *transforms, size_t n,
                                                const sph_settings_v1 *settings, float dt,
                                                const sph_box_v1 *box)
{
    return sph::update_particles_noalloc(particles, transforms, n, settings, dt, box);
}
