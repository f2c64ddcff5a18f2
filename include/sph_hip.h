/*
 * sph_hip.h -- C ABI of the MI355X-native SPH step (libsph_hip.so).
 *
 * Drop-in for the per-step operator of lijenicol/SPH-Fluid-Simulator:
 *   void updateParticles(Particle*, glm::mat4*, size_t, const SPHSettings&,
 *                        float deltaTime, bool onGPU)      src/sph.h:19-22
 *   void updateParticlesGPU(Particle*, glm::mat4*, size_t,
 *                           const SPHSettings&, float)     src/kernels/sphGPU.h:8-11
 * with the semantics of the reference's multithreaded CPU path
 * (updateParticlesCPU, src/sph.cpp:195-275), plus a resident API that keeps
 * scene state in HBM across steps.
 *
 * Conventions: every entry point returns SPH_OK (0) or a negative SPH_E_*
 * code and never throws; sph_hip_last_error() returns the calling thread's
 * last message.  Plain pointers and sizes only.  Calls on one context must be
 * serialised by the caller; distinct contexts are independent.  Host arrays
 * of positions/velocities are n x 3 floats, row-major.
 */
#ifndef SPH_HIP_H
#define SPH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_HIP_ABI_VERSION 1

#define SPH_OK 0
#define SPH_E_INVALID -1   /* bad argument */
#define SPH_E_NODEVICE -2  /* no usable HIP device */
#define SPH_E_NOMEM -3     /* device or host allocation failed */
#define SPH_E_HIP -4       /* other HIP runtime error (see last_error) */
#define SPH_E_LIMIT -5     /* size beyond the supported maximum */
#define SPH_E_STATE -6     /* call not valid in the object's current state */

/* Mirrors struct Particle, src/Particle.h:4-10: glm::vec3 position, velocity,
 * acceleration, force; float density, pressure; uint16_t hash; 2 bytes of
 * tail padding.  sizeof == 60, alignment 4. */
typedef struct sph_particle_v1 {
    float position[3];
    float velocity[3];
    float acceleration[3];
    float force[3];
    float density;
    float pressure;
    uint16_t hash;
    uint16_t _pad;
} sph_particle_v1;

/* Mirrors struct SPHSettings, src/SPHSystem.h:13-22 (glm::mat4 first, then
 * 13 floats in declaration order).  sizeof == 116. */
typedef struct sph_settings_v1 {
    float sphereScale[16];
    float poly6, spikyGrad, spikyLap, gasConstant, mass, h2, selfDens,
        restDensity, viscosity, h, g, tension, massPoly6Product;
} sph_settings_v1;

/* Axis-aligned wall box.  Particles reflect off the planes lo + h and hi - h
 * (floor, ceiling, x and z walls) with the given restitution and offset, in
 * the order floor, ceiling, x-low, x-high, z-low, z-high.  An infinite bound
 * disables that wall.  No counterpart struct in the reference: its CPU path
 * hard-codes lo = (-8, 0, -8), hi = (8, +inf, 8), 0.5, 1e-4
 * (src/sph.cpp:139-176), which sph_box_reference_cpu() reproduces. */
typedef struct sph_box_v1 {
    float lo[3];
    float hi[3];
    float restitution;
    float offset;
} sph_box_v1;

/* One scene for the resident API. */
typedef struct sph_scene_desc_v1 {
    uint32_t n;               /* particle count (> 0) */
    float dt;                 /* time step used by sph_hip_step */
    sph_settings_v1 settings; /* from sph_settings_init */
    sph_box_v1 box;
    float cell_scale;         /* grid cell edge in units of h (1 + 2^-10): 0 = 1; >= 1
                                 coarser cells; in [1/3, 1) finer cells with a stencil of
                                 ceil(1 / cell_scale) cells each way (fewer candidates per
                                 particle, more cells); results agree to rounding */
    uint32_t flags;           /* SPH_SCENE_* bits, others must be 0 */
} sph_scene_desc_v1;

/* Keep the per-particle force of the last step in device memory so that
 * sph_hip_download_particles can return it (costs 16 bytes per particle
 * per step of extra traffic). */
#define SPH_SCENE_KEEP_FORCE 1u

/* Every wave of the density and force passes walks its candidates per lane
 * instead of choosing the wave-uniform walk where it pays.  A wave task
 * whose walk is not split into parts gives bit-identical results either
 * way; a split one adds the same terms in another grouping (rounding-level
 * differences).  The flag exists for tests. */
#define SPH_SCENE_LANE_WALK 2u

/* The cell size follows the state: before the scene's first step and every
 * 64 steps after that (every 256 while the scene is clumped well below h on
 * cells of h), the candidates per neighbour of a sample of its particles
 * choose cells of h / 2 (a state uniform at the scale of h: about 40% fewer
 * candidates per particle) or of h (a state clumped below h, where finer
 * cells only add stencil rows).  Scenes of fewer than 131,072 particles
 * keep cells of h.  The choice depends only on the scene's own
 * state and step count, so the scene steps the same alone and in any batch;
 * results move by rounding when the grid changes. */
#define SPH_SCENE_AUTO_GRID 4u

/* ---- host-only helpers (no device work) --------------------------------- */

int sph_abi_version(void);
const char *sph_hip_last_error(void);

/* SPHSettings::SPHSettings(mass, restDensity, gasConst, viscosity, h, g,
 * tension), src/SPHSystem.cpp:8-26: same derived constants, bit for bit. */
int sph_settings_init(sph_settings_v1 *s, float mass, float restDensity,
                      float gasConst, float viscosity, float h, float g,
                      float tension);

/* Fixed box of the reference CPU path, src/sph.cpp:139-176. */
int sph_box_reference_cpu(sph_box_v1 *b);
/* Fixed box of the reference CUDA path, src/kernels/sphGPU.cu:8-10,210-238
 * (half-width 10, restitution 1, offset 1e-5). */
int sph_box_reference_gpu(sph_box_v1 *b);

/* getHash(ivec3), src/neighborTable.cpp:5-12 (18-bit value). */
uint32_t sph_get_hash(int32_t cx, int32_t cy, int32_t cz);

/* Host evaluation of the per-particle classification the kernels use:
 * out[0] = 16-bit bucket (Particle::hash), out[1] = 1 if the particle takes
 * the exact multiplicity path, out[2..4] = reference cell (getCell,
 * src/neighborTable.cpp:14-17).  Exposed for tests. */
int sph_classify_host(const float pos[3], float h, int32_t out[5]);

/* ---- stateless drop-in (replaces updateParticles(..., onGPU)) -----------
 * One step of the reference CPU operator on n particles, in place:
 * particles come back sorted by hash (ties in input order); hash, density,
 * pressure and force hold the values computed from the incoming positions;
 * position and velocity are integrated; acceleration is carried through
 * unchanged; transforms[i] (16 floats, column-major glm::mat4, may be NULL)
 * is translate(position) * scale(h / 2) of the returned particles[i].
 * The reference CPU wall box is used; the _box variant takes any box.
 * n == 0 is a no-op.  Device buffers are cached between calls. */
int sph_hip_update_particles(sph_particle_v1 *particles, float *transforms,
                             size_t n, const sph_settings_v1 *settings,
                             float dt);
int sph_hip_update_particles_box(sph_particle_v1 *particles, float *transforms,
                                 size_t n, const sph_settings_v1 *settings,
                                 float dt, const sph_box_v1 *box);

/* Largest n sph_hip_update_particles_noalloc accepts. */
#define SPH_NOALLOC_MAX_PARTICLES 65536u

/* As sph_hip_update_particles_box, but allocates no device memory: every
 * device buffer of the step comes from one fixed 64 MiB array in device
 * memory, which each GPU the library runs on holds until the process ends.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (harnesses that forbid hipMalloc, allocation-free serving loops).
 * n > SPH_NOALLOC_MAX_PARTICLES fails with SPH_E_LIMIT.  Results are
 * bit-identical to sph_hip_update_particles_box.  All calls on one GPU share
 * its array: the library serialises them with a process-wide lock and each
 * returns after its device work finished; other device work that wrote into
 * the array during a call would race with it.  Runs on the null stream. */
int sph_hip_update_particles_noalloc(sph_particle_v1 *particles,
                                     float *transforms, size_t n,
                                     const sph_settings_v1 *settings,
                                     float dt, const sph_box_v1 *box);

/* Neighbour multisets of the reference scan for the given particles (the
 * sets parallelDensityAndPressures sums over, src/sph.cpp:40-66), computed
 * by the density kernel's own candidate walk.  offsets has n + 1 entries;
 * row i lists input indices j, ascending, each repeated m_ij times.  If idx
 * is NULL or cap < total, only offsets and *total are written. */
int sph_hip_neighbors(const sph_particle_v1 *particles, size_t n,
                      const sph_settings_v1 *settings, uint64_t *offsets,
                      uint32_t *idx, size_t cap, size_t *total);

/* ---- resident single-scene API ------------------------------------------ */

typedef struct sph_ctx sph_ctx;

int sph_hip_create(const sph_scene_desc_v1 *desc, sph_ctx **out);
int sph_hip_destroy(sph_ctx *ctx);
/* Uploads n positions and velocities (vel may be NULL = zero); particle ids
 * are 0..n-1 in upload order. */
int sph_hip_upload(sph_ctx *ctx, const float *pos, const float *vel);
/* Advances nsteps steps of dt on the context's stream (asynchronous). */
int sph_hip_step(sph_ctx *ctx, int nsteps);
int sph_hip_sync(sph_ctx *ctx);
/* Copies the current state in device order; any pointer may be NULL.
 * rho, pressure are the values computed during the last step (from the
 * positions at its start).  Synchronises. */
int sph_hip_download(sph_ctx *ctx, float *pos, float *vel, float *rho,
                     float *pressure, uint32_t *ids);
/* AoS view of the current state (reference Particle layout) plus the
 * transform buffer the instanced renderer reads (16 floats per particle,
 * may be NULL).  hash/density/pressure/force are from the last step. */
int sph_hip_download_particles(sph_ctx *ctx, sph_particle_v1 *out,
                               float *transforms);
/* Tests: how the production walks of the next step see the current state.
 * Runs that step's passes up to density (cell sort, part plan, union /
 * per-lane / split walks, neighbour lists) with a neighbour count switched
 * on, without changing the state or the results of later steps.  Per
 * particle id: count[i] = neighbours with multiplicity (the terms the
 * reference sums, src/sph.cpp:40-66); walk[i] (may be NULL) = bit 0: its
 * wave task walks the union of its lanes' stencils, bit 1: the task is
 * split into parts, bit 2: a neighbour list was recorded; list_count[i]
 * (may be NULL) = list entries or 0xFFFFFFFF; list (may be NULL): 64
 * entries per id, the listed neighbours' ids (each once) in walk order. */
int sph_hip_walk_dump(sph_ctx *ctx, uint32_t *count, uint32_t *walk,
                      uint32_t *list_count, uint32_t *list);
/* The scene's current cell scale (in units of h (1 + 2^-10)) and the
 * candidates per neighbour of a sample of its particles in the last step's
 * state, computed now on the current grid (0 before the first step): the
 * statistic SPH_SCENE_AUTO_GRID decides by.  Either pointer may be NULL. */
int sph_hip_grid_info(sph_ctx *ctx, float *cell_scale, double *cand_per_nbr);
/* The determinism limit: a cell holding more than 16,384 particles is not
 * sorted by particle id and keeps the arrival order of its particles, so
 * the summation order over it, and the last bits of the results near it,
 * can differ between runs.  *out = the number of such cells summed over
 * every cell sort the context ran since creation (steps, autotune trials,
 * walk dumps); 0 means no step so far met the limit.  No counterpart in
 * the reference (its std::sort of the particles by hash,
 * src/sph.cpp:184-191, has no such limit). */
int sph_hip_unordered_cells(sph_ctx *ctx, uint64_t *out);
/* 1 = replay steps through a captured HIP graph (default), 0 = launch. */
int sph_hip_set_graph(sph_ctx *ctx, int enable);
uint32_t sph_hip_count(const sph_ctx *ctx);

/* ---- batch of scenes ------------------------------------------------------
 * Independent scenes (own settings, box, dt, grid, particle ids) advanced
 * together: every pass is one launch over the particles of all scenes.  No
 * counterpart in the reference (it steps one system); each scene's results
 * are bit-identical to stepping it alone with sph_hip_*.  Every scene needs
 * one upload before the first step.  Downloads return the scene's particles
 * in device order with their ids. */
typedef struct sph_batch sph_batch;

int sph_batch_create(const sph_scene_desc_v1 *descs, uint32_t nscenes,
                     sph_batch **out);
int sph_batch_destroy(sph_batch *b);
int sph_batch_upload(sph_batch *b, uint32_t scene, const float *pos,
                     const float *vel);
int sph_batch_step(sph_batch *b, int nsteps);
int sph_batch_sync(sph_batch *b);
int sph_batch_download(sph_batch *b, uint32_t scene, float *pos, float *vel,
                       float *rho, float *pressure, uint32_t *ids);
int sph_batch_set_graph(sph_batch *b, int enable);
/* sph_hip_download_particles for one scene of a batch (forces need
 * SPH_SCENE_KEEP_FORCE on some scene of the batch). */
int sph_batch_download_particles(sph_batch *b, uint32_t scene,
                                 sph_particle_v1 *out, float *transforms);
/* sph_hip_grid_info for one scene of a batch. */
int sph_batch_grid_info(sph_batch *b, uint32_t scene, float *cell_scale,
                        double *cand_per_nbr);
/* sph_hip_unordered_cells for a batch (all its scenes). */
int sph_batch_unordered_cells(sph_batch *b, uint64_t *out);
/* sph_hip_walk_dump for one scene of a batch. */
int sph_batch_walk_dump(sph_batch *b, uint32_t scene, uint32_t *count,
                        uint32_t *walk, uint32_t *list_count, uint32_t *list);
uint32_t sph_batch_scene_count(const sph_batch *b);
int sph_batch_profile(sph_batch *b, int nsteps, double *ms, int max_k,
                      int *nk, char *names_buf, size_t names_cap);
/* sph_batch_step with the GPU time of every step: ms[k] (may be NULL) for
 * step k.  Synchronises. */
int sph_batch_step_timed(sph_batch *b, int nsteps, float *ms);

/* ---- autotuner --------------------------------------------------------------
 * Times trial_steps steps from the current state for each stencil block
 * size (64, 128, 256 threads), restoring the state before every trial and
 * after the last, and keeps the fastest for the following steps.  The knob
 * changes no result bit.  trial_ms[k] is the mean step time of trial k
 * (k = 0, 1, 2 for 64, 128, 256 threads; 3-5 unused).  The reference has no
 * counterpart (its thread count is hardware_concurrency, src/sph.cpp:199).
 *
 * The _grid variants also tune the grid's cell size multiple: every scene
 * is re-gridded with cell scale 1 (trials 0-2) and 1/2 (trials 3-5, a
 * 5x5x5 stencil of half-size cells: about half the candidates where the
 * density is even at the scale of h, more rows where particles cluster
 * tighter than h), and the fastest pair is kept, also for later uploads
 * (it replaces the scenes' own cell_scale).  The grid changes the
 * order of each particle's sum, so results change by rounding (each scene
 * stays bit-identical to the same scene stepped alone on that grid). */
typedef struct sph_tune_v1 {
    int32_t stencil_block;       /* chosen threads per block */
    int32_t lane_walk;           /* always 0 (reserved) */
    int32_t trials;
    float cell_scale;            /* _grid variants: chosen cell scale (1 or 0.5), else 0 */
    double ms_per_step;          /* mean step time of the chosen setting */
    double default_ms_per_step;  /* same for 128 threads */
    double trial_ms[6];
} sph_tune_v1;

int sph_batch_autotune(sph_batch *b, int trial_steps, sph_tune_v1 *out);
int sph_hip_autotune(sph_ctx *ctx, int trial_steps, sph_tune_v1 *out);
int sph_batch_autotune_grid(sph_batch *b, int trial_steps, sph_tune_v1 *out);
int sph_hip_autotune_grid(sph_ctx *ctx, int trial_steps, sph_tune_v1 *out);

/* ---- serving loop -----------------------------------------------------------
 * Continuous batching of scene-advance requests.  No counterpart in the
 * reference (its only caller advances one system per frame,
 * src/SPHSystem.cpp:110-117).  submit() queues a request (the initial state
 * is copied, so the caller's arrays may be reused at once); tick() runs one
 * scheduling round: it admits queued requests first-come first-served while
 * their device memory fits the budget and a scene slot is free, advances
 * every admitted scene by the same number of steps (up to
 * max_steps_per_tick, never past any request's end), and retires finished
 * requests, whose final state waits in host memory for fetch().  Admitted
 * scenes share every kernel launch; each result is bit-identical to the
 * same scene stepped alone with sph_hip_*.  A request that can never fit
 * the budget is refused at submit with SPH_E_LIMIT. */
typedef struct sph_server sph_server;

typedef struct sph_server_config_v1 {
    uint64_t memory_budget;      /* device bytes for admitted scenes; 0 = 3/4 of free memory */
    uint32_t max_scenes;         /* admitted scenes at once; 0 = 1024 */
    uint32_t max_steps_per_tick; /* batch steps per tick; 0 = 8, at most 256 */
    uint32_t flags;              /* SPH_SERVER_* bits, others must be 0 */
    uint32_t max_queued;         /* queued requests at once; 0 = no limit; a submit
                                    beyond it fails with SPH_E_STATE (counted as rejected) */
} sph_server_config_v1;

/* Step with stream launches instead of one captured graph per step. */
#define SPH_SERVER_NO_GRAPH 1u
/* Strict first-come first-served admission.  By default, when the head of
 * the queue does not fit, a later request that fits is admitted ahead of it
 * if it will finish (in batch steps) before enough memory frees up for the
 * head, so backfilling never delays the head. */
#define SPH_SERVER_NO_BACKFILL 2u

#define SPH_REQ_QUEUED 0
#define SPH_REQ_ACTIVE 1
#define SPH_REQ_DONE 2
#define SPH_REQ_FETCHED 3

typedef struct sph_server_stats_v1 {
    uint64_t submitted, admitted, completed, queued, active;
    uint64_t steps;              /* batch steps run */
    uint64_t used_bytes, budget_bytes;
    double particle_steps;       /* sum over batch steps of admitted particles */
    double scene_steps;          /* sum over batch steps of admitted scenes */
    double busy_ms;              /* GPU time of all batch steps */
    double wall_s;               /* seconds since create */
    double step_ms_p50, step_ms_p99, step_ms_max;  /* per batch step */
    uint64_t rejected;           /* submits refused: queue at max_queued */
    uint64_t backfilled;         /* requests admitted ahead of the queue's head */
} sph_server_stats_v1;

int sph_server_create(const sph_server_config_v1 *cfg, sph_server **out);
int sph_server_destroy(sph_server *s);
int sph_server_submit(sph_server *s, const sph_scene_desc_v1 *desc,
                      const float *pos, const float *vel, uint32_t steps,
                      uint64_t *request_id);
/* *ran_steps (may be NULL) receives the batch steps run in this round. */
int sph_server_tick(sph_server *s, uint32_t *ran_steps);
/* *state: SPH_REQ_*; *done_steps (may be NULL): steps run so far; times
 * (may be NULL): seconds since create of submit, admission and completion. */
int sph_server_status(const sph_server *s, uint64_t request_id, int *state,
                      uint32_t *done_steps, double times[3]);
/* Final state of a finished request (n x 3 positions and velocities, n
 * densities, pressures and ids, any may be NULL) in device order; frees it. */
int sph_server_fetch(sph_server *s, uint64_t request_id, float *pos,
                     float *vel, float *rho, float *pressure, uint32_t *ids);
int sph_server_stats(const sph_server *s, sph_server_stats_v1 *out);
/* Tests only: the next admission's relayout (SPH_SERVER_FAIL_RELAYOUT) or
 * the next scene upload (SPH_SERVER_FAIL_UPLOAD) fails once, as a failed
 * device allocation would.  The failing tick returns the error; its joining
 * requests go back to the head of the queue and the running scenes keep
 * their state, so the next tick admits them again. */
#define SPH_SERVER_FAIL_RELAYOUT 1
#define SPH_SERVER_FAIL_UPLOAD 2
int sph_server_inject_failure(sph_server *s, int stage);
/* GPU time of every batch step so far, oldest first (up to cap values). */
int sph_server_step_times(const sph_server *s, float *ms, size_t cap,
                          size_t *n);

/* ---- measurement ---------------------------------------------------------
 * Runs nsteps un-graphed steps with HIP events around every kernel on the
 * context's stream.  names: nul-separated list written into names_buf;
 * ms[k]: mean duration of kernel k per step; *nk: number of kernels. */
int sph_hip_profile(sph_ctx *ctx, int nsteps, double *ms, int max_k, int *nk,
                    char *names_buf, size_t names_cap);
/* Device-to-device copy kernel over `bytes` bytes, `iters` times; *ms is
 * the mean kernel time.  Calibrates bandwidth and HBM counters. */
int sph_hip_copy_bench(size_t bytes, int iters, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* SPH_HIP_H */
