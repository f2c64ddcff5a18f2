// sph_batch.h -- resident scenes sharing one set of device arrays (see
// sph_batch.hip).  Used by the single-scene context, the batch API and the
// serving loop.
#pragma once

#include <vector>

#include "sph_runtime.h"

namespace sph {

constexpr uint32_t kMaxParticles = 1u << 30;

struct BatchScene {
    sph_scene_desc_v1 desc;
    StepParams P;  // constants, grid, slot range
    bool uploaded = false;
};

int validate_desc(const sph_scene_desc_v1 &d);
void scene_grid(StepParams &P, const sph_scene_desc_v1 &d, const float *pos, uint32_t n);

struct Batch {
    hipStream_t stream = nullptr;
    Workspace w;
    std::vector<BatchScene> scenes;
    uint32_t total_n = 0;
    bool layout_dirty = true, keys_dirty = true;
    bool use_graph = true, keep_force = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf aos, xf;  // download staging

    Batch() = default;
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;
    ~Batch();

    int init(const sph_scene_desc_v1 *descs, uint32_t k);
    int upload(uint32_t s, const float *pos, const float *vel);
    // Lays out cells and the scene table and computes keys when needed.
    int prepare();
    void enqueue(StepTimer *timer);
    int step(int nsteps);
    int sync();
    int download(uint32_t s, float *pos, float *vel, float *rho, float *pressure, uint32_t *ids);
    int download_particles(uint32_t s, sph_particle_v1 *out, float *transforms);
    int profile(int nsteps, double *ms, int max_k, int *nk, char *names_buf, size_t names_cap);
    void drop_graph();
};

}  // namespace sph
