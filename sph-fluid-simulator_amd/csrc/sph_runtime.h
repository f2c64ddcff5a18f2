// sph_runtime.h -- internal host-side pieces shared by the runtime files.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/sph_hip.h"
#include "sph_kernels.h"

namespace sph {

int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);
const char *last_error();
int ensure_device();

void settings_init(sph_settings_v1 *s, float mass, float restDensity, float gasConst,
                   float viscosity, float h, float g, float tension);
void box_reference_cpu(sph_box_v1 *b);
void box_reference_gpu(sph_box_v1 *b);
int check_settings(const sph_settings_v1 *s);
StepParams make_params(const sph_settings_v1 &s, const sph_box_v1 &box, float dt);
int make_grid(StepParams &P, const double lo[3], const double hi[3], double cell_scale,
              uint64_t max_cells);
// Stencil reach in cells of a grid cell scale (cell edge in units of
// h (1 + 2^-10)): 1 for 0 (the default) and scales >= 1, ceil(1 / scale)
// for scales in [1 / kMaxReach, 1), 0 (invalid) otherwise.
int grid_reach(double cell_scale);

// Growable device allocation (hipMalloc on growth only; a failed growth
// leaves the buffer as it was).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t want);
    void release();
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
};

// All device arrays of one particle set (see sph_kernels.h for the layout).
struct Workspace {
    DevBuf posA, velA, posB, velB, cellcA, rankA, cellcB, perm, rhoB, posR, velR;
    DevBuf counts, cstart, partials, big;
    DevBuf force;
    DevBuf params, tasks;  // scene table of the batch living in this workspace
    // split wave tasks: parts per task (padded), their scan, scan partials,
    // walk lengths and their total, part tasks, queue counters, partial sums,
    // finished-part counters
    DevBuf pnparts, poff, ppartials, pmeasure, ptotal, ptasks, pqueue, pbuf, pdone, pwalk, pown;
    // neighbour lists the density pass writes and the force pass reads
    DevBuf nlist, ncount;
    // cells left in arrival order (above the in-cell sort's limit), summed
    // over steps
    DevBuf unordered;
    uint32_t parts_padded = 0;
    uint32_t ncells_padded = 0;
    uint32_t big_cap = 0;
    uint32_t nwaves = 0;
    uint32_t nscenes = 0;
    uint32_t part_cap = 0;  // part-task capacity of the current scene set
    int stencil_block = kStencilBlock;  // threads per block of the density and force passes
    int reserve_particles(uint32_t n);
    int reserve_cells(uint64_t ncells);
    // Uploads the scene table (p_begin/p_end/cell_base filled in) and builds
    // one 64-slot wave task per started 64 slots of every scene.
    int set_scenes(const std::vector<StepParams> &scenes, hipStream_t st);
    ParticleArrays particles() const;
    CellArrays cells() const;
    SceneSet scene_set() const;
    PartPlan part_plan() const;
    NeighborLists lists() const;
    void release_all();
};

// HIP events recorded between kernels on one stream.
struct StepTimer {
    std::vector<hipEvent_t> events;
    std::vector<const char *> names;  // names[k] labels events[k-1] -> events[k]
    void mark(hipStream_t st, const char *name)
    {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, st);
        events.push_back(e);
        names.push_back(name);
    }
    void clear()
    {
        for (hipEvent_t e : events)
            (void)hipEventDestroy(e);
        events.clear();
        names.clear();
    }
    ~StepTimer() { clear(); }
};

// One step on the workspace (keys of posA valid, counts holding the
// per-cell totals).  `timer` (optional) records an event after each kernel.
void enqueue_step(hipStream_t st, Workspace &w, float4 *force_out, bool next_keys, StepTimer *timer);

}  // namespace sph
