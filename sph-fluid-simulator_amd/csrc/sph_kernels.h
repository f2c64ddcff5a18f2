// sph_kernels.h -- host-side launchers for the HIP kernels of one SPH step.
//
// Data layout in HBM (structure of arrays, one entry per particle):
//   "A" arrays: the state at the start of a step, in the order the previous
//               step wrote it (= the previous step's cell order)
//     posA   float4  x, y, z, w = bits of the 16-bit bucket H16(getCell(x))
//     velA   float4  vx, vy, vz, w = bits of the particle id
//     cellcA u32     packed grid cell + slow-multiplicity flag
//     rankA  u32     scratch (the drop-in output's 16-bit bucket ranks)
//   "B" arrays: the same particles in grid-cell order (stable by A index)
//     posB, velB, cellcB  as above
//     rhoB   f32     density                          (density pass)
//     posR   float4  x, y, z, 0.5 / density           (density pass)
//     velR   float4  vx, vy, vz, pressure             (density pass)
//   posR / velR are what the force pass gathers: one record for the
//   candidate filter and one more per neighbour.
//   cells: counts[ncells] (between steps: the particles of A per cell; the
//          scatter counts them back to zero), cstart[ncells + 1]
//
// Several independent scenes can share the arrays ("batch"): scene s owns
// the slot range [p_begin, p_end) of both orders and the cell range
// [cell_base, cell_base + nx ny nz).  Per-particle kernels run one 64-slot
// wave per task (scene, first slot), so every wave reads one scene's
// constants; a single scene is a batch of one.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_common.h"

namespace sph {

constexpr int kBlock = 256;           // threads per block, particle kernels
constexpr int kStencilBlock = 128;    // default threads per block, density / force
constexpr int kForceChunk = 8;        // candidates filtered per bitmask in the force pass
constexpr int kScanItems = 16;        // cells per thread in the scan
constexpr uint32_t kScanChunk = kBlock * kScanItems;  // cells per scan block
constexpr uint32_t kSmallCell = 64;        // cells up to this size: stable rank by counting
constexpr uint32_t kBigCellSortMax = 16384;  // larger cells keep arrival order
constexpr int kBigSortBlock = 1024;        // threads of the big-cell sort
constexpr int kBigSortBlocks = 256;        // blocks walking the big-cell list
constexpr uint32_t kPartBits = 8;          // part index bits of a part task
constexpr uint32_t kMaxParts = 1u << kPartBits;  // parts of one split wave task, at most
constexpr uint32_t kPartCandidates = 2048; // candidates per part, at least
constexpr uint32_t kPartRoomPerWave = 3;   // spare part tasks per wave task of a scene, at least
constexpr uint32_t kPartRoomMin = 49152;   // ... raised to this many per scene (~10 rounds of the
                                           // resident waves, so tasks stay short in mid-size scenes)
constexpr uint32_t kPartRoomMaxPerWave = 16;  // ... but at most this many per wave task
constexpr uint32_t kLaneWalkWeight = 5;    // per-lane walk candidates count this many times
constexpr uint32_t kSlowWalkWeight = 8;    // ... in a wave with slow lanes (m_ij counted per neighbour)

struct ParticleArrays {
    float4 *posA, *velA, *posB, *velB;
    uint32_t *cellcA, *rankA, *cellcB, *perm;
    float *rhoB;
    float4 *posR, *velR;
};

struct CellArrays {
    uint32_t *counts;  // ncells_padded entries: A's particles per cell between steps
    uint32_t *cstart;  // ncells_padded + 1 entries
    uint32_t *partials;
    uint32_t *big;     // [0] = number of big cells, then their indices (may be null)
    uint32_t big_cap;  // capacity of the index list
    unsigned long long *unordered = nullptr;  // += 1 per cell and step kept in arrival order (may be null)
    uint32_t ncells_padded;  // multiple of kScanChunk
};

struct WaveTask {
    uint32_t scene;
    uint32_t first;  // first slot of the wave
};

// Device view of the scenes of a batch.
struct SceneSet {
    const StepParams *params;  // one per scene
    const WaveTask *tasks;     // one per 64-slot wave
    uint32_t nwaves;
};

// Spare part tasks of a scene of `waves` wave tasks (beyond one per task):
// the scene's part size is its total walk over this room, so a scene
// splits into at most waves + part_room(waves) part tasks.  Depends on the
// scene alone.
SPH_HD inline uint64_t part_room(uint64_t waves)
{
    const uint64_t lo = kPartRoomPerWave * waves, hi = kPartRoomMaxPerWave * waves;
    const uint64_t want = hi < kPartRoomMin ? hi : (uint64_t)kPartRoomMin;
    return lo > want ? lo : want;
}

// Split wave tasks.  A wave task whose candidate walk is long (the union's
// candidates, or kLaneWalkWeight x the largest own stencil of a per-lane
// walk) runs as P parts, each an equal slice of every lane's candidate
// sequence, with P = ceil(length / part size) up to kMaxParts; the part size
// is kPartCandidates, raised when the scene's total walk over part_room
// would exceed it (a scene's split depends on the scene alone, so it is the
// same in any batch).  Unsplit tasks are dealt to the waves statically;
// parts are separate tasks pulled from a queue afterwards; the last one to
// finish adds the partial sums in part order, so results are
// deterministic.  An unsplit task sums in candidate order (and its two
// walks agree bit for bit); a split one differs from it by rounding only.
struct PartPlan {
    uint32_t *nparts;    // per wave task: parts of a split task, 0 for an unsplit one; padded
                         // to a kScanChunk multiple (zero outside use)
    uint32_t *off;       // exclusive scan of nparts; off[nwaves] = number of queued part tasks
    uint32_t *scan_partials;
    uint32_t padded;     // padded length of nparts
    uint32_t *measure;   // walk length per wave task
    unsigned long long *total;  // sum of the walk lengths per scene (zero between steps)
    uint32_t nscenes;
    uint32_t cap;        // part-task capacity: sum over scenes of waves + part_room(waves)
    uint32_t *tasks;     // (wave task << kPartBits) | part, cap entries
    uint32_t *queue;     // pulled-task counters: [0] density, [1] force
    float4 *buf;         // partial sums: (part task) * 64 + lane, cap * 64 entries
    uint32_t *done;      // parts finished per wave task (zero between passes)
    int4 *walk;          // per wave task, two entries: union box (x0, x1, y0, y1), (z0, z1, union
                         // count, 1 = union walk / 0 = per-lane walk)
    uint32_t *own;       // per slot of a per-lane-walk wave task: length of its own stencil
};

// Plans the parts of every wave task of S from the B-order cells and
// cell ranges, and clears the task queues.
void launch_plan_parts(hipStream_t st, const ParticleArrays &a, const CellArrays &c, const SceneSet &S,
                       const PartPlan &pp);

// Debug outputs of the density pass: per-particle neighbour count (with
// multiplicity) and, when out != nullptr, the neighbour ids of row s at
// out[off[s] ...].  With `production` set, only the count is written, by
// the production walks of the step (part plan, union / per-lane / split
// parts, neighbour lists) instead of a separate whole-stencil walk.
struct NeighborDump {
    uint32_t *count;
    const uint64_t *off;
    uint32_t *out;
    bool production = false;
};

// Neighbour lists.  The density pass records, for every particle of an
// unsplit wave task that walks its candidates per lane, the slots of its
// neighbours (each once, in walk order: rows z-major, then y, slots
// ascending) when there are at most kListMax of them; the force pass then
// visits exactly those slots instead of filtering the stencil again.  The
// order is the per-lane walk's, so the force sums are the same bit for bit.
// Entry k of slot s lives at list[k stride + s]: the lanes of a wave
// that append their k-th neighbour together write one contiguous run, so
// list lines are written whole.
constexpr uint32_t kListMax = 64;
constexpr uint32_t kNoList = 0xFFFFFFFFu;  // count of a slot without a list

__host__ __device__ inline size_t list_at(uint32_t k, uint32_t s, uint32_t stride)
{
    return (size_t)k * stride + s;
}

struct NeighborLists {
    uint32_t *list;   // kListMax * stride entries; null: no lists
    uint32_t *count;  // per slot: list entries, or kNoList
    uint32_t stride;  // slot capacity of the arrays
};

inline uint32_t blocks_for(uint32_t n, int block = kBlock) { return (n + block - 1) / block; }

void launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t value, size_t n);
void launch_unpack_aos(hipStream_t st, const void *aos, uint32_t n, float4 *posA, float4 *velA);
void launch_bbox(hipStream_t st, const float4 *pos, uint32_t n, int32_t *bbox6);
void launch_init_keys(hipStream_t st, const ParticleArrays &a, const CellArrays &c, const SceneSet &S);
// Scan of the cell counts; also lists the cells above kSmallCell in c.big.
// clear_counts: zero the counts (else the scatter counts them back down).
void launch_scan(hipStream_t st, const CellArrays &c, bool clear_counts);
void launch_scatter(hipStream_t st, const ParticleArrays &a, const CellArrays &c, const SceneSet &S);
// Sorts the A indices of every listed big cell (deterministic in-cell order).
void launch_sort_big_cells(hipStream_t st, const ParticleArrays &a, const CellArrays &c);
void launch_reorder(hipStream_t st, const ParticleArrays &a, const CellArrays &c, const SceneSet &S);
// `block`: threads per block of the stencil passes, 64, 128 or 256 (the
// dump modes always use kStencilBlock); results do not depend on it.
// `parts` (may be null: every wave task unsplit, one wave each) is the
// plan launch_plan_parts made for this step.
// `lists` (may be null: none) receives the neighbour lists of the step;
// the force pass reads the same.
void launch_density(hipStream_t st, const ParticleArrays &a, const CellArrays &c, const SceneSet &S,
                    const NeighborDump *dump, const PartPlan *parts = nullptr, int block = kStencilBlock,
                    const NeighborLists *lists = nullptr);
void launch_force_integrate(hipStream_t st, const ParticleArrays &a, const CellArrays &c,
                            const SceneSet &S, float4 *force_out, bool next_keys,
                            const PartPlan *parts = nullptr, int block = kStencilBlock,
                            const NeighborLists *lists = nullptr);
// Sampled candidates per neighbour of every scene (one wave task in
// kGridStatEvery of each scene), from the current B order and cell ranges:
// stat[2 s] = candidates, stat[2 s + 1] = neighbours of scene s (cleared
// first).  A lane with more than kGridStatTests candidates tests every
// k-th of them (k = ceil(candidates / kGridStatTests)) and counts each
// neighbour found k times.
constexpr uint32_t kGridStatEvery = 16;
constexpr uint32_t kGridStatTests = 256;
void launch_grid_stat(hipStream_t st, const ParticleArrays &a, const CellArrays &c, const SceneSet &S,
                      unsigned long long *stat, uint32_t nscenes);
// Drop-in output (single scene): order slots by the pre-step 16-bit bucket
// (ties by id), and write the reference AoS layout plus transforms.
void launch_h16_rank(hipStream_t st, const ParticleArrays &a, const CellArrays &c16, uint32_t n);
void launch_h16_scatter(hipStream_t st, const ParticleArrays &a, const CellArrays &c16, uint32_t n,
                        uint32_t *perm16);
void launch_pack_out(hipStream_t st, const ParticleArrays &a, const CellArrays &c16,
                     const uint32_t *perm16, const float4 *force, const void *aos_in, uint32_t n,
                     float xf_scale, void *aos_out, float *xf_out);
// Resident download: AoS of slots [first, first + n) of the A state.
void launch_pack_state(hipStream_t st, const ParticleArrays &a, const float4 *force, uint32_t first,
                       uint32_t n, float xf_scale, void *aos_out, float *xf_out);
void launch_copy(hipStream_t st, const uint4 *src, uint4 *dst, size_t n16);

}  // namespace sph
