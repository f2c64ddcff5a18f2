// sph_kernels.h -- host-side launchers for the HIP kernels of one SPH step.
//
// Data layout in HBM (structure of arrays, one entry per particle):
//   "A" arrays: the state at the start of a step, in the order the previous
//               step wrote it (= the previous step's cell order)
//     posA   float4  x, y, z, w = bits of the 16-bit bucket H16(getCell(x))
//     velA   float4  vx, vy, vz, w = bits of the particle id
//     cellcA u32     packed grid cell + slow-multiplicity flag
//     rankA  u32     arrival rank of the particle inside its grid cell
//   "B" arrays: the same particles in grid-cell order (stable by A index)
//     posB, velB, cellcB  as above
//     rhoB   f32     density                          (density pass)
//     posR   float4  x, y, z, 0.5 / density           (density pass)
//     velR   float4  vx, vy, vz, pressure             (density pass)
//   posR / velR are what the force pass gathers: one record for the
//   candidate filter and one more per neighbour.
//   cells: counts[ncells] (zero between steps), cstart[ncells + 1]
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_common.h"

namespace sph {

constexpr int kBlock = 256;           // threads per block, particle kernels
constexpr int kStencilBlock = 128;    // threads per block, density / force
#ifndef SPH_STAGE_CAP
#define SPH_STAGE_CAP 0
#endif
constexpr uint32_t kStageCap = SPH_STAGE_CAP;  // positions staged in LDS per block (0: off)
constexpr int kForceChunk = 16;       // candidates filtered per bitmask in the force pass
constexpr int kScanItems = 16;        // cells per thread in the scan
constexpr uint32_t kScanChunk = kBlock * kScanItems;  // cells per scan block
constexpr uint32_t kStableCellMax = 512;  // larger cells keep arrival order

struct ParticleArrays {
    float4 *posA, *velA, *posB, *velB;
    uint32_t *cellcA, *rankA, *cellcB, *perm;
    float *rhoB;
    float4 *posR, *velR;
};

struct CellArrays {
    uint32_t *counts;  // ncells_padded entries, zero between steps
    uint32_t *cstart;  // ncells_padded + 1 entries
    uint32_t *partials;
    uint32_t ncells_padded;  // multiple of kScanChunk
};

// Debug outputs of the density pass: per-particle neighbour count (with
// multiplicity) and, when out != nullptr, the neighbour ids of row s at
// out[off[s] ...].
struct NeighborDump {
    uint32_t *count;
    const uint64_t *off;
    uint32_t *out;
};

inline uint32_t blocks_for(uint32_t n, int block = kBlock) { return (n + block - 1) / block; }

void launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t value, size_t n);
void launch_unpack_aos(hipStream_t st, const void *aos, uint32_t n, float4 *posA, float4 *velA);
void launch_bbox(hipStream_t st, const float4 *pos, uint32_t n, int32_t *bbox6);
void launch_init_keys(hipStream_t st, const ParticleArrays &a, const CellArrays &c, uint32_t n,
                      const StepParams &P);
void launch_scan(hipStream_t st, const CellArrays &c, uint32_t n_total);
void launch_scatter(hipStream_t st, const ParticleArrays &a, const CellArrays &c, uint32_t n,
                    const StepParams &P);
void launch_reorder(hipStream_t st, const ParticleArrays &a, const CellArrays &c, uint32_t n,
                    const StepParams &P);
void launch_density(hipStream_t st, const ParticleArrays &a, const CellArrays &c, uint32_t n,
                    const StepParams &P, const NeighborDump *dump);
void launch_force_integrate(hipStream_t st, const ParticleArrays &a, const CellArrays &c,
                            uint32_t n, const StepParams &P, float4 *force_out, bool next_keys);
// Drop-in output: order slots by the pre-step 16-bit bucket (ties by id),
// and write the reference AoS layout plus transforms.
void launch_h16_rank(hipStream_t st, const ParticleArrays &a, const CellArrays &c16, uint32_t n);
void launch_h16_scatter(hipStream_t st, const ParticleArrays &a, const CellArrays &c16, uint32_t n,
                        uint32_t *perm16);
void launch_pack_out(hipStream_t st, const ParticleArrays &a, const CellArrays &c16,
                     const uint32_t *perm16, const float4 *force, const void *aos_in, uint32_t n,
                     const StepParams &P, void *aos_out, float *xf_out);
// Resident download: AoS of the current A state (device order).
void launch_pack_state(hipStream_t st, const ParticleArrays &a, const float4 *force, uint32_t n,
                       const StepParams &P, void *aos_out, float *xf_out);
void launch_copy(hipStream_t st, const uint4 *src, uint4 *dst, size_t n16);

}  // namespace sph
