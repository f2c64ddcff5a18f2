// sph_batch.h -- resident scenes sharing one set of device arrays (see
// sph_batch.hip).  Used by the single-scene context, the batch API and the
// serving loop.
#pragma once

#include <vector>

#include "sph_runtime.h"

namespace sph {

constexpr uint32_t kMaxParticles = 1u << 30;

// Automatic cell size (SPH_SCENE_AUTO_GRID): before a scene's first step
// and every kRegridEvery steps after that (kRegridClumped below), the
// sampled candidates per neighbour of its current state choose its cell
// scale: a state uniform at the scale of h takes cells of h / 2 (about 40%
// fewer candidates), one clumped below h takes cells of h (fewer stencil
// rows).  The choice depends on the scene's state and history alone.
constexpr uint64_t kRegridEvery = 64;
// Thresholds from same-state timings of both grids along the 1M-particle
// dam break (20 steps from the states at steps 0, 64, 128, ...): cells of
// h / 2 were faster at steps 0 and 64 (ratios 7.32 / 7.07 on cells of h,
// 4.24 / 4.12 on cells of h / 2) and slower from step 128 on (6.62 / 3.90);
// the two ratios of one state differ by a factor of about 1.7.  Smaller
// lattices read less (their surfaces): a 52^3 one 7.2, an 8,000-particle
// cube 6.3.
constexpr double kToFineRatio = 6.85;  // at cell scale 1: above this, go to 1/2
constexpr double kToCoarseRatio = 4.0;  // at cell scale 1/2: below this, go to 1
// A scene on cells of h that read below kClumpedRatio (clumped well below
// h; the dam break from step ~250 on) is evaluated every kRegridClumped
// steps only: each evaluation sorts the state once more and waits for the
// device, about 0.8 ms on the collapsed 1M-particle flow.
constexpr double kClumpedRatio = 4.0;
constexpr uint64_t kRegridClumped = 256;
// Smaller scenes keep cells of h: on the batched workload (256 scenes of
// 1k-64k particles) cells of h / 2 made the 500-step job 26% slower (eight
// times the cells to scan, and no faster density pass at these sizes).
constexpr uint32_t kAutoGridMinParticles = 1u << 17;

struct BatchScene {
    sph_scene_desc_v1 desc;
    StepParams P;  // constants, grid, slot range
    double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};  // bounds the grid covers
    bool uploaded = false;
    uint64_t steps = 0;  // steps since the upload
    double scale = 0.0;  // cell scale of the current grid (0: desc.cell_scale)
    double last_ratio = 0.0;  // candidates per neighbour at the last evaluation
    uint64_t evaluated = ~0ull;  // step count of the last evaluation (the upload's and the first step's are one)
};

int validate_desc(const sph_scene_desc_v1 &d);
// Bounds of b's grid from its box and particles, then its grid (scene_regrid
// with the scene's own cell scale).
void scene_grid(BatchScene &b, const float *pos, uint32_t n);
// b's grid over its bounds with cell scale `scale` (0: desc.cell_scale).
int scene_regrid(BatchScene &b, double scale);
// A scene entry for desc (validated): constants, flags; no slot range yet.
int make_scene(const sph_scene_desc_v1 &d, BatchScene &out);

struct Batch {
    hipStream_t stream = nullptr;
    Workspace w;
    std::vector<BatchScene> scenes;
    uint32_t total_n = 0;
    bool layout_dirty = true, keys_dirty = true;
    bool use_graph = true, keep_force = false;
    uint64_t max_capacity = 0;  // particle capacity relayout may grow to (0: unbounded)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf aos, xf;  // download staging
    DevBuf gstat;    // grid statistic: candidates, neighbours per scene
    bool stat_ok = false;    // B order and cell ranges are those of the last step
    bool hold_grid = false;  // steps neither count nor re-grid (autotune trials)
    std::vector<hipEvent_t> events;
    // Tests only: a failure the next relayout (kFailRelayout) or upload
    // (kFailUpload) reports at its last fallible point, once.
    static constexpr int kFailRelayout = 1, kFailUpload = 2;
    int inject_failure = 0;

    Batch() = default;
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;
    ~Batch();

    int init(const sph_scene_desc_v1 *descs, uint32_t k);
    // A batch with no scenes yet (the serving loop sets them with relayout).
    int init_empty(bool keep_force_);
    // Replaces the scene list with `next`: next[k] continues old scene
    // from[k] (its particle state moves with device copies, in the same
    // order, so it steps on bit-identically) or, for from[k] < 0, is a new
    // scene to upload before stepping.  Densities, pressures and forces of
    // continued scenes are not carried over (the next step recomputes them).
    // On failure the batch keeps its old scenes and their state.
    int relayout(std::vector<BatchScene> next, const std::vector<int> &from);
    int upload(uint32_t s, const float *pos, const float *vel);
    // Lays out cells and the scene table and computes keys when needed.
    int prepare();
    void enqueue(StepTimer *timer);
    int step(int nsteps);
    // nsteps steps on the current layout (no grid decisions)
    int step_plain(int nsteps);
    // Captures the step into the graph when there is none (use_graph).
    int capture();
    // Sampled candidates per neighbour of every scene (ratio[s]): fresh =
    // from the cell order of the current state (sorted here), else from
    // the last step's (0 if there is none).
    int grid_stat(std::vector<double> &ratio, bool fresh);
    // Evaluates the automatic grid of the scenes whose step count is at a
    // re-grid point (a multiple of kRegridEvery, of kRegridClumped while
    // clumped; the upload included) and re-grids those whose choice changed.
    int auto_grid();
    int grid_info(uint32_t s, float *cell_scale, double *cand_per_nbr);
    // Cells the in-cell sort left in arrival order (more than
    // kBigCellSortMax particles), summed over every sort since creation.
    int unordered_cells(uint64_t *out);
    // step() with an event after every step; ms[k] (may be NULL) receives
    // the GPU time of step k.  Synchronises.
    int step_timed(int nsteps, float *ms);
    // Times trial_steps steps from the current state for every stencil
    // block size (64, 128, 256), restoring the state before each trial and
    // after the last, and keeps the fastest.  Results do not depend on it.
    // grid: also try cell scales 1 and 1/2 for every scene (changes results by rounding)
    int autotune(int trial_steps, sph_tune_v1 *out, bool grid = false);
    int sync();
    int download(uint32_t s, float *pos, float *vel, float *rho, float *pressure, uint32_t *ids);
    int download_particles(uint32_t s, sph_particle_v1 *out, float *transforms);
    int profile(int nsteps, double *ms, int max_k, int *nk, char *names_buf, size_t names_cap);
    // Test view of the production walks on the current state (see
    // sph_batch_walk_dump): runs the step's passes up to density with the
    // neighbour count switched on, into temporary B-side arrays.
    int walk_dump(uint32_t s, uint32_t *count, uint32_t *walk, uint32_t *list_count, uint32_t *list);
    void drop_graph();
};

}  // namespace sph
