// sph_kernels.hip -- hand-written CDNA4 (gfx950) kernels of one SPH step.
//
// One step is, in launch order:
//   scan          exclusive scan of per-cell counts -> cstart, counts := 0,
//                 and the list of cells holding more than kSmallCell
//   scatter       perm[cstart[cell] + rank] = A index
//   sort_big      sorts the A indices inside each listed big cell
//   reorder       stable cell order (ties by A index), gather A -> B
//   density       poly6 density + pressure        (src/sph.cpp:28-76)
//   force_integrate
//                 spiky pressure + viscosity force (src/sph.cpp:80-129),
//                 symplectic Euler + walls          (src/sph.cpp:133-181),
//                 then the next step's cell key and in-cell rank
// The hash / sort / table passes of the reference (src/sph.cpp:17-24,
// 184-192, src/neighborTable.cpp:19-37) are replaced by the dense-grid
// counting sort; the reference's 16-bit bucket semantics are kept through
// the multiplicity rule documented in sph_common.h.
//
// Floating-point contraction is off for the whole file: the squared
// distance must be fl(fl(fl(dx*dx) + fl(dy*dy)) + fl(dz*dz)) as glm forms it,
// and integration reproduces the reference's rounding step by step.  The
// force pass uses explicit fmaf() where the reference's rounding is not the
// contract (tolerance-checked sums).
#include "sph_kernels.h"

#include <algorithm>
#include <type_traits>

#pragma clang fp contract(off)

namespace sph {

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Re-map so each XCD walks a contiguous range of particles: neighbouring
// cells then share one XCD's L2.
__device__ inline uint32_t xcd_block()
{
    uint32_t b = blockIdx.x, nb = gridDim.x;
    uint32_t per = nb >> 3, rem = nb & 7u;
    uint32_t xcd = b & 7u, k = b >> 3;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

// The wave task of the calling wave: its scene (wave-uniform, so the
// scene's constants are scalar loads) and its slot.  Returns false for
// waves past the task list and lanes past the scene's last slot.
struct Lane {
    const StepParams *P;
    uint32_t slot;
    uint32_t first;  // first slot of the wave (wave-uniform)
    uint32_t task;   // index of the wave task (wave-uniform)
};

template <int BLOCK>
__device__ inline bool lane_of(const SceneSet &S, Lane &out)
{
    const uint32_t w = __builtin_amdgcn_readfirstlane(xcd_block() * (BLOCK / 64) + (threadIdx.x >> 6));
    if (w >= S.nwaves)
        return false;
    const WaveTask t = S.tasks[w];
    const uint32_t scene = __builtin_amdgcn_readfirstlane(t.scene);
    const uint32_t first = __builtin_amdgcn_readfirstlane(t.first);
    out.P = S.params + scene;
    out.first = first;
    out.slot = first + (threadIdx.x & 63u);
    return out.slot < out.P->p_end;
}

// Wave task w (wave-uniform) with every lane live: lanes past the scene's
// last slot mirror the wave's first slot (always a valid slot) and report
// active = false, so wave-wide votes and shuffles see defined data.
__device__ inline void wave_task(const SceneSet &S, uint32_t w, Lane &out, bool &active)
{
    const WaveTask t = S.tasks[w];
    const uint32_t scene = __builtin_amdgcn_readfirstlane(t.scene);
    const uint32_t first = __builtin_amdgcn_readfirstlane(t.first);
    out.P = S.params + scene;
    out.first = first;
    out.task = w;
    const uint32_t slot = first + (threadIdx.x & 63u);
    active = slot < out.P->p_end;
    out.slot = active ? slot : first;
}

// As lane_of, but with every lane of a live wave live (see wave_task).
// Returns false only for waves past the task list (the whole wave).
template <int BLOCK>
__device__ inline bool wave_lanes(const SceneSet &S, Lane &out, bool &active)
{
    const uint32_t w = __builtin_amdgcn_readfirstlane(xcd_block() * (BLOCK / 64) + (threadIdx.x >> 6));
    if (w >= S.nwaves)
        return false;
    wave_task(S, w, out, active);
    return true;
}

inline uint32_t blocks_for_waves(uint32_t nwaves, int block) { return (nwaves + block / 64 - 1) / (block / 64); }

// Waves the device holds at once: CUs x 4 SIMDs x 8 wave slots.
uint32_t resident_waves()
{
    static uint32_t cached = 0;
    if (!cached) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cached = 32u * (uint32_t)cus;
    }
    return cached;
}

// ---------------------------------------------------------------- utilities

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t value, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride)
        p[i] = value;
}

void launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t value, size_t n)
{
    if (n == 0)
        return;
    size_t nb = (n + kBlock - 1) / kBlock;
    if (nb > 4096)
        nb = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)nb), dim3(kBlock), 0, st, p, value, n);
}

// Reference Particle AoS (60 bytes) -> posA/velA with id = input index.
__global__ void k_unpack_aos(const float *__restrict__ aos, uint32_t n, float4 *__restrict__ posA,
                             float4 *__restrict__ velA)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float *p = aos + (size_t)i * 15;
    posA[i] = make_float4(p[0], p[1], p[2], 0.0f);
    velA[i] = make_float4(p[3], p[4], p[5], __uint_as_float(i));
}

void launch_unpack_aos(hipStream_t st, const void *aos, uint32_t n, float4 *posA, float4 *velA)
{
    hipLaunchKernelGGL(k_unpack_aos, dim3(blocks_for(n)), dim3(kBlock), 0, st,
                       (const float *)aos, n, posA, velA);
}

// Order-preserving float <-> int map for atomic min/max.
__device__ inline int32_t ordered_int(float f)
{
    int32_t i = __float_as_int(f);
    return i >= 0 ? i : (int32_t)(i ^ 0x7FFFFFFF);
}

// Bounding box of finite positions: bbox6 = {min x,y,z, max x,y,z} as
// ordered ints (pre-filled with INT_MAX / INT_MIN).
__global__ void k_bbox(const float4 *__restrict__ pos, uint32_t n, int32_t *__restrict__ bbox6)
{
    __shared__ int32_t red[6][kBlock / 64];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t v[6] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN, INT32_MIN};
    if (i < n) {
        float4 p = pos[i];
        float c[3] = {p.x, p.y, p.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (c[k] - c[k] == 0.0f) {  // finite
                int32_t o = ordered_int(c[k]);
                v[k] = o;
                v[k + 3] = o;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int32_t x = v[k];
        for (int d = 32; d >= 1; d >>= 1) {
            int32_t y = __shfl_xor(x, d, 64);
            x = (k < 3) ? min(x, y) : max(x, y);
        }
        if ((threadIdx.x & 63) == 0)
            red[k][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        int k = threadIdx.x;
        int32_t x = red[k][0];
        for (int w = 1; w < kBlock / 64; ++w)
            x = (k < 3) ? min(x, red[k][w]) : max(x, red[k][w]);
        if (k < 3)
            atomicMin(&bbox6[k], x);
        else
            atomicMax(&bbox6[k], x);
    }
}

void launch_bbox(hipStream_t st, const float4 *pos, uint32_t n, int32_t *bbox6)
{
    hipLaunchKernelGGL(k_bbox, dim3(blocks_for(n)), dim3(kBlock), 0, st, pos, n, bbox6);
}

// Cell key and bucket of the positions in posA, counted into their cells
// (the count is not read back here: the scatter ranks the particles).
__global__ __launch_bounds__(kBlock) void k_init_keys(float4 *__restrict__ posA,
                                                     uint32_t *__restrict__ cellcA,
                                                     uint32_t *__restrict__ counts, SceneSet S)
{
    Lane L;
    if (!lane_of<kBlock>(S, L))
        return;
    const StepParams &P = *L.P;
    const uint32_t i = L.slot;
    float4 p = posA[i];
    Classified c = classify(p.x, p.y, p.z, P);
    posA[i].w = __uint_as_float(c.h16);
    cellcA[i] = c.cellc;
    atomicAdd(&counts[linear_cell(P, c.cellc)], 1u);
}

void launch_init_keys(hipStream_t st, const ParticleArrays &a, const CellArrays &c, const SceneSet &S)
{
    hipLaunchKernelGGL(k_init_keys, dim3(blocks_for_waves(S.nwaves, kBlock)), dim3(kBlock), 0, st,
                       a.posA, a.cellcA, c.counts, S);
}

// ------------------------------------------------------------- cell scan
// Three passes over the (padded) cell array: per-block totals, a one-block
// scan of the totals, then the block-local exclusive scan plus offset.  The
// last pass appends every cell above kSmallCell to the big-cell list and,
// when asked, clears the counts (the cell scan leaves them: the scatter
// counts them back down to zero while it ranks the particles).

__device__ inline uint32_t wave_inclusive_scan(uint32_t x)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d)
            x += y;
    }
    return x;
}

// Exclusive scan across the block of one value per thread; returns the
// exclusive prefix and writes the block total.
__device__ inline uint32_t block_exclusive_scan(uint32_t x, uint32_t *total)
{
    __shared__ uint32_t wsum[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(x);
    if (lane == 63)
        wsum[wave] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        uint32_t s = wsum[w];
        off += (w < wave) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - x;
}

__global__ void k_scan_reduce(const uint32_t *__restrict__ counts, uint32_t *__restrict__ partials)
{
    const uint4 *src = (const uint4 *)(counts + (size_t)blockIdx.x * kScanChunk);
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems / 4; ++k) {
        uint4 v = src[k * kBlock + threadIdx.x];
        s += v.x + v.y + v.z + v.w;
    }
    uint32_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = tot;
}

// One block: exclusive scan of nparts block totals in place; cstart_end
// receives the grand total.  Also clears the big-cell count.
__global__ void k_scan_partials(uint32_t *__restrict__ partials, uint32_t nparts,
                                uint32_t *__restrict__ cstart_end, uint32_t *__restrict__ big)
{
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nparts; base += kBlock) {
        uint32_t i = base + threadIdx.x;
        uint32_t x = i < nparts ? partials[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(x, &tot);
        if (i < nparts)
            partials[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        *cstart_end = carry;
        if (big)
            big[0] = 0u;
    }
}

__global__ void k_scan_final(uint32_t *__restrict__ counts, const uint32_t *__restrict__ partials,
                             uint32_t *__restrict__ cstart, uint32_t *__restrict__ big, uint32_t big_cap,
                             bool clear)
{
    size_t base = (size_t)blockIdx.x * kScanChunk + (size_t)threadIdx.x * kScanItems;
    uint4 *src = (uint4 *)(counts + base);
    uint32_t v[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems / 4; ++k) {
        uint4 q = src[k];
        v[4 * k + 0] = q.x;
        v[4 * k + 1] = q.y;
        v[4 * k + 2] = q.z;
        v[4 * k + 3] = q.w;
    }
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        local += v[k];
    uint32_t tot;
    uint32_t run = block_exclusive_scan(local, &tot) + partials[blockIdx.x];
    uint4 *dst = (uint4 *)(cstart + base);
#pragma unroll
    for (int k = 0; k < kScanItems / 4; ++k) {
        uint4 q;
        q.x = run;
        run += v[4 * k + 0];
        q.y = run;
        run += v[4 * k + 1];
        q.z = run;
        run += v[4 * k + 2];
        q.w = run;
        run += v[4 * k + 3];
        dst[k] = q;
        if (clear)
            src[k] = make_uint4(0u, 0u, 0u, 0u);
    }
    if (big) {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (v[k] > kSmallCell) {
                const uint32_t at = atomicAdd(&big[0], 1u);
                if (at < big_cap)
                    big[1 + at] = (uint32_t)(base + k);
            }
        }
    }
}

void launch_scan(hipStream_t st, const CellArrays &c, bool clear_counts)
{
    uint32_t nparts = c.ncells_padded / kScanChunk;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nparts), dim3(kBlock), 0, st, c.counts, c.partials);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kBlock), 0, st, c.partials, nparts,
                       c.cstart + c.ncells_padded, c.big);
    hipLaunchKernelGGL(k_scan_final, dim3(nparts), dim3(kBlock), 0, st, c.counts, c.partials,
                       c.cstart, c.big, c.big_cap, clear_counts);
}

// ---------------------------------------------------------- scatter/reorder

// Each particle takes a distinct place in its cell's range by counting the
// cell's count down (the order within a cell is fixed afterwards by the
// reorder), which also leaves the counts at zero for the next step.  The
// one atomic whose result a thread waits on is here, in a kernel that does
// nothing else, rather than at the end of the force pass.  Lanes of a wave
// that share a cell (A is in the previous step's cell order, so a wave
// holds few cells) take their places with one atomic: the lowest lane of
// each group subtracts the group's size, and every lane of the group adds
// its position among them.
__global__ __launch_bounds__(kBlock) void k_scatter(const uint32_t *__restrict__ cellcA,
                                                   uint32_t *__restrict__ counts,
                                                   const uint32_t *__restrict__ cstart,
                                                   uint32_t *__restrict__ perm, SceneSet S)
{
    Lane L;
    if (!lane_of<kBlock>(S, L))
        return;
    const uint32_t i = L.slot;
    const uint32_t lin = linear_cell(*L.P, cellcA[i]);
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    uint64_t rem = __ballot(1);  // the live lanes
    uint32_t leader = lane, size = 1, offset = 0;
    while (rem != 0ull) {
        const uint32_t l = (uint32_t)__builtin_ctzll(rem);
        const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)lin, (int)l);
        const uint64_t group = __ballot(lin == c) & rem;
        if (lin == c) {
            leader = l;
            size = (uint32_t)__builtin_popcountll(group);
            offset = (uint32_t)__builtin_popcountll(group & below);
        }
        rem &= ~group;
    }
    uint32_t top = 0;
    if (lane == leader)
        top = atomicSub(&counts[lin], size);  // one instruction for every group of the wave
    top = (uint32_t)__shfl((int)top, (int)leader, 64);
    perm[cstart[lin] + top - size + offset] = i;
}

void launch_scatter(hipStream_t st, const ParticleArrays &a, const CellArrays &c, const SceneSet &S)
{
    hipLaunchKernelGGL(k_scatter, dim3(blocks_for_waves(S.nwaves, kBlock)), dim3(kBlock), 0, st,
                       a.cellcA, c.counts, c.cstart, a.perm, S);
}

// Big cells: one block sorts the A indices of one listed cell (bitonic,
// ascending), so slot order inside the cell is the stable order.  Thread t
// holds elements t + kBigSortBlock j in registers: compare-exchanges at
// strides >= kBigSortBlock stay in the thread, strides < 64 are wave
// shuffles, and only strides 64 .. kBigSortBlock / 2 go through LDS (22 of
// the 91 stages of an 8192-element sort).  Blocks walk the list with a
// stride; every block exits when the list ends.
constexpr int kBigSortRegs = (int)(kBigCellSortMax / kBigSortBlock);

// Stages at stride Q kBigSortBlock: element j pairs with j | Q in the same
// thread (Q and the element count RR compile-time, so the register indices
// are fixed).
template <int RR, uint32_t Q>
__device__ inline void sort_stage_in_thread(uint32_t (&x)[RR], uint32_t t, uint32_t size)
{
#pragma unroll
    for (int j = 0; j < RR; ++j) {
        if (((uint32_t)j & Q) == 0u && (j | (int)Q) < RR) {
            const int jj = j | (int)Q;
            const bool up = ((t + (uint32_t)kBigSortBlock * (uint32_t)j) & size) == 0u;
            const uint32_t lo = min(x[j], x[jj]), hi = max(x[j], x[jj]);
            x[j] = up ? lo : hi;
            x[jj] = up ? hi : lo;
        }
    }
}

// Sorts the len (<= RR kBigSortBlock, > 64) entries of perm at b; m is the
// power of two >= len.  Called by every thread of the block.
template <int RR>
__device__ inline void sort_cell(uint32_t *__restrict__ perm, uint32_t b, uint32_t len, uint32_t m,
                                 uint32_t *__restrict__ sh)
{
    constexpr uint32_t T = kBigSortBlock;
    const uint32_t t = threadIdx.x;
    // m >= 128 is a multiple of 64: a wave is wholly live or wholly idle
    const bool live = t < m;
    uint32_t x[RR];
#pragma unroll
    for (int j = 0; j < RR; ++j) {
        const uint32_t k = t + T * (uint32_t)j;
        x[j] = k < len ? perm[b + k] : 0xFFFFFFFFu;
    }
    for (uint32_t size = 2; size <= m; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= T) {
                switch (stride / T) {
                case 1: sort_stage_in_thread<RR, 1>(x, t, size); break;
                case 2: sort_stage_in_thread<RR, 2>(x, t, size); break;
                case 4: sort_stage_in_thread<RR, 4>(x, t, size); break;
                default: sort_stage_in_thread<RR, 8>(x, t, size); break;
                }
            } else if (stride < 64u) {
                if (live) {
#pragma unroll
                    for (int j = 0; j < RR; ++j) {
                        const uint32_t k = t + T * (uint32_t)j;
                        const uint32_t y = (uint32_t)__shfl_xor((int)x[j], (int)stride, 64);
                        const bool up = (k & size) == 0u, low = (k & stride) == 0u;
                        x[j] = (low == up) ? min(x[j], y) : max(x[j], y);
                    }
                }
            } else {
                if (live) {
#pragma unroll
                    for (int j = 0; j < RR; ++j)
                        sh[t + T * (uint32_t)j] = x[j];
                }
                __syncthreads();
                if (live) {
#pragma unroll
                    for (int j = 0; j < RR; ++j) {
                        const uint32_t k = t + T * (uint32_t)j;
                        const uint32_t y = sh[k ^ stride];
                        const bool up = (k & size) == 0u, low = (k & stride) == 0u;
                        x[j] = (low == up) ? min(x[j], y) : max(x[j], y);
                    }
                }
                __syncthreads();
            }
        }
    }
#pragma unroll
    for (int j = 0; j < RR; ++j) {
        const uint32_t k = t + T * (uint32_t)j;
        if (k < len)
            perm[b + k] = x[j];
    }
}

__global__ __launch_bounds__(kBigSortBlock) void k_sort_big_cells(
    uint32_t *__restrict__ perm, const uint32_t *__restrict__ cstart, const uint32_t *__restrict__ big,
    uint32_t big_cap, unsigned long long *__restrict__ unordered)
{
    static_assert(kBigSortRegs == 16 && kSmallCell >= 64, "sort layout");
    __shared__ uint32_t sh[kBigCellSortMax];
    const uint32_t nbig = min(big[0], big_cap);
    for (uint32_t e = blockIdx.x; e < nbig; e += gridDim.x) {
        const uint32_t cell = big[1 + e];
        const uint32_t b = cstart[cell], len = cstart[cell + 1] - b;
        if (len > kBigCellSortMax) {  // uniform: keeps arrival order, counted
            if (threadIdx.x == 0 && unordered)
                atomicAdd(unordered, 1ull);
            continue;
        }
        uint32_t m = 1;
        while (m < len)
            m <<= 1;
        switch (m > (uint32_t)kBigSortBlock ? m / kBigSortBlock : 1u) {  // elements per thread
        case 1: sort_cell<1>(perm, b, len, m, sh); break;
        case 2: sort_cell<2>(perm, b, len, m, sh); break;
        case 4: sort_cell<4>(perm, b, len, m, sh); break;
        case 8: sort_cell<8>(perm, b, len, m, sh); break;
        default: sort_cell<16>(perm, b, len, m, sh); break;
        }
    }
}

void launch_sort_big_cells(hipStream_t st, const ParticleArrays &a, const CellArrays &c)
{
    if (!c.big)
        return;
    hipLaunchKernelGGL(k_sort_big_cells, dim3(kBigSortBlocks), dim3(kBigSortBlock), 0, st, a.perm,
                       c.cstart, c.big, c.big_cap, c.unordered);
}

// Slot s holds A index i; its stable destination is the cell start plus the
// number of smaller A indices in the same cell.  Cells above kSmallCell
// were sorted by k_sort_big_cells, so their destination is s itself.
__global__ __launch_bounds__(kBlock) void k_reorder(
    const uint32_t *__restrict__ perm, const uint32_t *__restrict__ cellcA,
    const uint32_t *__restrict__ cstart, const float4 *__restrict__ posA,
    const float4 *__restrict__ velA, float4 *__restrict__ posB, float4 *__restrict__ velB,
    uint32_t *__restrict__ cellcB, SceneSet S)
{
    Lane L;
    if (!lane_of<kBlock>(S, L))
        return;
    const uint32_t s = L.slot;
    const uint32_t i = perm[s];
    const uint32_t c = cellcA[i];
    const uint32_t lin = linear_cell(*L.P, c);
    const uint32_t b = cstart[lin], e = cstart[lin + 1];
    uint32_t d;
    if (e - b <= kSmallCell) {
        uint32_t r = 0;
        for (uint32_t t = b; t < e; ++t)
            r += (perm[t] < i) ? 1u : 0u;
        d = b + r;
    } else {
        d = s;
    }
    posB[d] = posA[i];
    velB[d] = velA[i];
    cellcB[d] = c;
}

void launch_reorder(hipStream_t st, const ParticleArrays &a, const CellArrays &c, const SceneSet &S)
{
    hipLaunchKernelGGL(k_reorder, dim3(blocks_for_waves(S.nwaves, kBlock)), dim3(kBlock), 0, st,
                       a.perm, a.cellcA, c.cstart, a.posA, a.velA, a.posB, a.velB, a.cellcB, S);
}

// ------------------------------------------------------------ stencil walk

// fl(fl(fl(dx*dx) + fl(dy*dy)) + fl(dz*dz)), dx = x_j - x_i: glm::length2.
__device__ inline float dist2_ref(float4 pi, float4 q, float &dx, float &dy, float &dz)
{
    dx = q.x - pi.x;
    dy = q.y - pi.y;
    dz = q.z - pi.z;
    return dx * dx + dy * dy + dz * dz;
}

// Correctly rounded sqrt for d2 >= 2^-96 (any d2 a neighbour pair can have
// short of coincident points; d2 = 0 also comes out exact): v_sqrt_f32 is
// within 1 ulp, and the neighbouring candidate whose residual
// d2 - s * s' changes sign is the correctly rounded root (the tiny-input
// rescaling of the library sequence is not needed in this range).
__device__ inline float sqrt_rn(float x)
{
    float s = __builtin_amdgcn_sqrtf(x);
    float down = __uint_as_float(__float_as_uint(s) - 1u);
    float up = __uint_as_float(__float_as_uint(s) + 1u);
    float r_down = fmaf(-down, s, x);
    float r_up = fmaf(-up, s, x);
    s = (r_down <= 0.0f) ? down : s;
    s = (r_up > 0.0f) ? up : s;
    return s;
}

struct RefCell {
    int x, y, z;
};

__device__ inline RefCell ref_cell(float4 p, const StepParams &P)
{
    return RefCell{trunc_to_int(p.x / P.h), trunc_to_int(p.y / P.h), trunc_to_int(p.z / P.h)};
}

// The (2R + 1)^2 rows of the (2R + 1)^3 grid-cell stencil around packed
// cell cc (R = P.reach), clipped to the grid: for row (y, z) the x-run of
// cells [gx - R, gx + R] is one contiguous range [b, e) of the cell-ordered
// arrays.  Calls row(b, e) for each row, z-major then y, both ascending.
// The bounds of the next row are loaded while the current row runs.
template <class Row>
__device__ inline void for_each_row(const uint32_t *__restrict__ cstart, uint32_t cc,
                                    const StepParams &P, Row row)
{
    const int gx = cell_x(cc), gy = cell_y(cc), gz = cell_z(cc), R = P.reach;
    const int x0 = max(gx - R, 0), x1 = min(gx + R, P.nx - 1);
    const int y0 = max(gy - R, 0), y1 = min(gy + R, P.ny - 1);
    const int z0 = max(gz - R, 0), z1 = min(gz + R, P.nz - 1);
    const int nrows = (y1 - y0 + 1) * (z1 - z0 + 1);
    auto base = [&](int y, int z) {
        return P.cell_base + (uint32_t)P.nx * ((uint32_t)y + (uint32_t)P.ny * (uint32_t)z);
    };
    int y = y0, z = z0;
    uint32_t r0 = base(y, z);
    uint32_t b = cstart[r0 + (uint32_t)x0], e = cstart[r0 + (uint32_t)x1 + 1u];
    for (int k = 0; k < nrows; ++k) {
        if (++y > y1) {
            y = y0;
            ++z;
        }
        uint32_t bn = 0, en = 0;
        if (k + 1 < nrows) {
            r0 = base(y, z);
            bn = cstart[r0 + (uint32_t)x0];
            en = cstart[r0 + (uint32_t)x1 + 1u];
        }
        row(b, e);
        b = bn;
        e = en;
    }
}

// The lane's place in its grid cell, for cutting its stencil to the cells
// a neighbour can occupy.  A neighbour j of lane i has fl(d^2) < fl(h^2),
// so its true distance is below h (1 + 2^-21), i.e. below rho = h / H cell
// edges.  The cell offset d along an axis lies at least gap(d) cell edges
// away (d - f for d > 0, -d - 1 + f for d < 0, 0 for d = 0, f = the lane's
// fraction of its cell), so a cell whose squared gaps add up to rho^2 or
// more holds no neighbour.  f comes from the same double expression as the
// cell index, and the float work around it errs by far less than the slack
// taken off every gap (2^-12 edges) and added to the reach (2^-10 of
// rho^2): the cut drops only candidates whose float squared distance
// exceeds h^2.
struct LaneTrim {
    float fx, fy, fz;  // fraction of the cell, clamped to [0, 1]
    float r2;          // rho^2 (1 + 2^-10)
};

constexpr float kTrimSlack = 1.0f / 4096.0f;

__device__ inline LaneTrim lane_trim(float4 p, uint32_t cc, const StepParams &P)
{
    auto frac = [&](float x, double o, int g) {
        const double q = ((double)x - o) * P.grid_inv - (double)g;
        return fminf(fmaxf((float)q, 0.0f), 1.0f);  // NaN -> 0
    };
    LaneTrim t;
    t.fx = frac(p.x, P.grid_o[0], cell_x(cc));
    t.fy = frac(p.y, P.grid_o[1], cell_y(cc));
    t.fz = frac(p.z, P.grid_o[2], cell_z(cc));
    const double rh = (double)P.h * P.grid_inv;
    t.r2 = (float)(rh * rh * (1.0 + 1.0 / 1024.0));
    return t;
}

// gap(d) less the slack, at least 0
__device__ inline float trim_gap(int d, float f)
{
    const float g = d > 0 ? (float)d - f : (d < 0 ? (float)(-d - 1) + f : 0.0f);
    return fmaxf(g - kTrimSlack, 0.0f);
}

// for_each_row cut by the lane's trim: row (y, z) is skipped when its
// gaps alone reach rho, else its x-run is cut to the cells within the
// remaining chord.  Calls row(b, e) for every non-empty row, in
// for_each_row's order, so the passing candidates come in the same order
// as on the whole stencil.  The bounds of the next row are loaded while the
// current row runs.
template <class Row>
__device__ inline void for_each_cut_row(const uint32_t *__restrict__ cstart, uint32_t cc, const LaneTrim &T,
                                        const StepParams &P, Row row)
{
    const int gx = cell_x(cc), gy = cell_y(cc), gz = cell_z(cc), R = P.reach;
    const int y0 = max(gy - R, 0), y1 = min(gy + R, P.ny - 1);
    const int z0 = max(gz - R, 0), z1 = min(gz + R, P.nz - 1);
    const int nrows = (y1 - y0 + 1) * (z1 - z0 + 1);
    auto bounds = [&](int y, int z, uint32_t &b, uint32_t &e) {
        const float ay = trim_gap(y - gy, T.fy), az = trim_gap(z - gz, T.fz);
        const float rem = T.r2 - (ay * ay + az * az);
        b = e = 0u;
        if (rem > 0.0f) {
            const float w = __builtin_amdgcn_sqrtf(rem) + 2.0f * kTrimSlack;
            const int x0 = max(gx + max((int)floorf(T.fx - w), -R), 0);
            const int x1 = min(gx + min((int)floorf(T.fx + w), R), P.nx - 1);
            if (x0 <= x1) {
                const uint32_t r0 = P.cell_base + (uint32_t)P.nx * ((uint32_t)y + (uint32_t)P.ny * (uint32_t)z);
                b = cstart[r0 + (uint32_t)x0];
                e = cstart[r0 + (uint32_t)x1 + 1u];
            }
        }
    };
    int y = y0, z = z0;
    uint32_t b, e;
    bounds(y, z, b, e);
    for (int k = 0; k < nrows; ++k) {
        if (++y > y1) {
            y = y0;
            ++z;
        }
        uint32_t bn = 0, en = 0;
        if (k + 1 < nrows)
            bounds(y, z, bn, en);
        if (b < e)
            row(b, e);
        b = bn;
        e = en;
    }
}

// for_each_row restricted to the positions [lo, hi) of the lane's own
// candidate sequence (its rows in order, slots ascending).
template <class Row>
__device__ inline void for_each_row_span(const uint32_t *__restrict__ cstart, uint32_t cc, const StepParams &P,
                                         uint32_t lo, uint32_t hi, Row row)
{
    uint32_t off = 0;
    for_each_row(cstart, cc, P, [&](uint32_t b, uint32_t e) {
        const uint32_t len = e - b;
        const uint32_t s0 = max(lo, off), s1 = min(hi, off + len);
        if (s0 < s1)
            row(b + (s0 - off), b + (s1 - off));
        off += len;
    });
}

__device__ inline int wave_min_i(int x)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        x = min(x, __shfl_xor(x, d, 64));
    return __builtin_amdgcn_readfirstlane(x);
}

__device__ inline int wave_max_i(int x)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        x = max(x, __shfl_xor(x, d, 64));
    return __builtin_amdgcn_readfirstlane(x);
}

// Most cells a union walk's wave may span in y and in z: the union's
// (span + 2R)^2 rows fit the 64 lanes that stage them (R = 1: 4, 2: 4, 3: 2).
__device__ inline int union_span(int R) { return min(4, 8 - 2 * R); }

// A box of grid cells [x0, x1] x [y0, y1] x [z0, z1] (inside the grid)
// whose rows a wave walks with one candidate index for all of its lanes.
struct UnionRows {
    int x0, x1, y0, y1, z0, z1;
};

__device__ inline uint32_t row_base(const StepParams &P, int y, int z)
{
    return P.cell_base + (uint32_t)P.nx * ((uint32_t)y + (uint32_t)P.ny * (uint32_t)z);
}

// Chooses between the two ways a wave walks its candidates:
//  * per lane: each lane walks its own (2R + 1)^3 stencil (gathers);
//  * union: all lanes walk the union of their stencils together, so the
//    candidate index is wave-uniform and its record is a scalar load
//    broadcast to every lane.
// Union is taken when all lanes sit in one cell (the dense regime: one
// cell holds many waves), or when their cells span at most 16 cells in x
// and union_span(R) in y and z (so the union has at most 64 rows) and the
// union holds at most 3x the largest own stencil (a per-lane walk costs
// about three times as much per candidate).  A candidate the union adds
// to a lane's own stencil lies in a cell R + 1 or more cells away, beyond
// R whole cells of edge >= h (1 + 2^-10) / R along some axis, so its
// float squared distance is above h^2 and it never passes: both walks
// visit a lane's passing candidates in the same order (rows z-major, then
// y, slots ascending) and give bit-identical sums.
// Length of the lane's own (2R + 1)^3 stencil, the sum of its rows'
// lengths as for_each_row visits them; the bounds of one z-layer's rows are
// loaded together (all in flight at once) instead of a row ahead.
__device__ inline uint32_t stencil_length(const uint32_t *__restrict__ cstart, uint32_t cc, const StepParams &P)
{
    const int gx = cell_x(cc), gy = cell_y(cc), gz = cell_z(cc), R = P.reach;
    const int x0 = max(gx - R, 0), x1 = min(gx + R, P.nx - 1);
    const int y0 = max(gy - R, 0), y1 = min(gy + R, P.ny - 1);
    const int z0 = max(gz - R, 0), z1 = min(gz + R, P.nz - 1);
    uint32_t own = 0;
    for (int z = z0; z <= z1; ++z) {
        uint32_t b[2 * kMaxReach + 1], e[2 * kMaxReach + 1];
#pragma unroll
        for (int k = 0; k < 2 * kMaxReach + 1; ++k) {
            b[k] = e[k] = 0u;
            if (y0 + k <= y1) {  // 2R + 1 rows (fewer at the grid's border)
                const uint32_t r0 = P.cell_base + (uint32_t)P.nx * ((uint32_t)(y0 + k) + (uint32_t)P.ny * (uint32_t)z);
                b[k] = cstart[r0 + (uint32_t)x0];
                e[k] = cstart[r0 + (uint32_t)x1 + 1u];
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * kMaxReach + 1; ++k)
            own += e[k] - b[k];
    }
    return own;
}

// own: the lane's own candidate count when this computed it (the
// comparison needs it unless all lanes share one cell), else kNoOwn.
constexpr uint32_t kNoOwn = 0xFFFFFFFFu;

__device__ inline bool plan_union(const uint32_t *__restrict__ cstart, uint32_t cc, const StepParams &P,
                                  UnionRows &U, uint32_t &count, uint32_t &own)
{
    if (P.lane_walk)
        return false;
    const uint32_t cell = cc & ~kSlowBit;
    const uint32_t c0 = __builtin_amdgcn_readfirstlane(cell);
    int x0, x1, y0, y1, z0, z1;
    const bool one = __all(cell == c0);
    if (one) {
        x0 = x1 = cell_x(c0);
        y0 = y1 = cell_y(c0);
        z0 = z1 = cell_z(c0);
    } else {
        const int gx = cell_x(cell), gy = cell_y(cell), gz = cell_z(cell);
        x0 = wave_min_i(gx);
        x1 = wave_max_i(gx);
        y0 = wave_min_i(gy);
        y1 = wave_max_i(gy);
        z0 = wave_min_i(gz);
        z1 = wave_max_i(gz);
        const int span = union_span(P.reach);
        if (x1 - x0 > 15 || y1 - y0 >= span || z1 - z0 >= span)
            return false;
    }
    const int R = P.reach;
    U.x0 = max(x0 - R, 0);
    U.x1 = min(x1 + R, P.nx - 1);
    U.y0 = max(y0 - R, 0);
    U.y1 = min(y1 + R, P.ny - 1);
    U.z0 = max(z0 - R, 0);
    U.z1 = min(z1 + R, P.nz - 1);
    uint32_t un = 0;
    for (int z = U.z0; z <= U.z1; ++z)
        for (int y = U.y0; y <= U.y1; ++y) {
            const uint32_t r0 = row_base(P, y, z);
            un += cstart[r0 + (uint32_t)U.x1 + 1u] - cstart[r0 + (uint32_t)U.x0];
        }
    count = __builtin_amdgcn_readfirstlane(un);
    if (one)
        return true;
    own = stencil_length(cstart, cc, P);
    const uint32_t own_max = (uint32_t)wave_max_i((int)own);
    return count <= 3ull * own_max;
}

// The walk a wave task takes and, for the union walk, its box and count:
// from the step's part plan when there is one (decided once per wave task
// by k_plan_parts), else decided here.  Waves with a slow lane take the
// per-lane walk.
__device__ inline bool walk_of(const PartPlan &pp, uint32_t wt, const uint32_t *__restrict__ cstart, uint32_t cc,
                               const StepParams &P, bool slow, UnionRows &U, uint32_t &count)
{
    if (pp.walk) {
        const int4 a = pp.walk[2u * wt], b = pp.walk[2u * wt + 1u];
        U.x0 = __builtin_amdgcn_readfirstlane(a.x);
        U.x1 = __builtin_amdgcn_readfirstlane(a.y);
        U.y0 = __builtin_amdgcn_readfirstlane(a.z);
        U.y1 = __builtin_amdgcn_readfirstlane(a.w);
        U.z0 = __builtin_amdgcn_readfirstlane(b.x);
        U.z1 = __builtin_amdgcn_readfirstlane(b.y);
        count = (uint32_t)__builtin_amdgcn_readfirstlane(b.z);
        return __builtin_amdgcn_readfirstlane(b.w) != 0;
    }
    uint32_t own = kNoOwn;
    return plan_union(cstart, cc, P, U, count, own) && !__any(slow);
}

// Length of the lane's own candidate sequence (its (2R + 1)^3 stencil).
__device__ inline uint32_t own_length(const PartPlan &pp, uint32_t s, const uint32_t *__restrict__ cstart, uint32_t cc,
                                      const StepParams &P)
{
    if (pp.own)
        return pp.own[s];
    return stencil_length(cstart, cc, P);
}

// Scalar union walk: the positions [lo, hi) of the union's flattened
// candidate sequence as wave-uniform runs (slots ascending), each split at
// the wave's own slots [first, first + 64) so only that run carries the
// self test.  run(j0, j1, self) walks slots [j0, j1); the candidate
// records are then scalar loads broadcast to the lanes through SGPRs.
template <class Run>
__device__ inline void scalar_union_walk(const uint32_t *__restrict__ cstart, const UnionRows &U,
                                         const StepParams &P, uint32_t lo, uint32_t hi, uint32_t first, Run run)
{
    uint32_t off = 0;
    for (int z = U.z0; z <= U.z1; ++z)
        for (int y = U.y0; y <= U.y1; ++y) {
            const uint32_t r0 = row_base(P, y, z);
            const uint32_t b = __builtin_amdgcn_readfirstlane(cstart[r0 + (uint32_t)U.x0]);
            const uint32_t e = __builtin_amdgcn_readfirstlane(cstart[r0 + (uint32_t)U.x1 + 1u]);
            const uint32_t s0 = max(lo, off), s1 = min(hi, off + (e - b));
            if (s0 < s1) {
                const uint32_t j0 = b + (s0 - off), j1 = b + (s1 - off);
                const uint32_t m0 = min(max(j0, first), j1), m1 = min(max(j0, first + 64u), j1);
                if (j0 < m0)
                    run(j0, m0, std::false_type{});
                if (m0 < m1)
                    run(m0, m1, std::true_type{});
                if (m1 < j1)
                    run(m1, j1, std::false_type{});
            }
            off += e - b;
        }
}

// Union walk staged through LDS.  The wave's union candidates, flattened
// in row order (rows as for_each_row visits them, slots ascending), are
// fetched 64 at a time -- one coalesced load per lane and record array, a
// chunk ahead of their use -- into a per-wave LDS double buffer, then read
// back by all lanes with one broadcast LDS read per record, so the
// candidate loop never waits on a global load and never pays per-row
// overhead.  Lanes past the end stage a sentinel (position +inf, slot ~0)
// that fails every test, so chunks run in groups of four unchecked.
template <int NV>
struct StageBuf {
    float4 rec[2][NV][64];
    uint32_t slot[2][64];
};

// Walks the positions [lo, hi) of the flattened sequence.
template <int NV, class Visit>
__device__ inline void staged_union_walk(const uint32_t *__restrict__ cstart, const UnionRows &U,
                                         const StepParams &P, uint32_t lo, uint32_t hi,
                                         const float4 *const (&src)[NV], StageBuf<NV> &sb, Visit visit)
{
    const uint32_t lane = threadIdx.x & 63u;
    const int ny = U.y1 - U.y0 + 1;
    const int nrows = ny * (U.z1 - U.z0 + 1);  // at most (union_span + 2R)^2 <= 64 lanes
    uint32_t b = 0, len = 0;
    if ((int)lane < nrows) {
        const uint32_t r0 = row_base(P, U.y0 + (int)lane % ny, U.z0 + (int)lane / ny);
        b = cstart[r0 + (uint32_t)U.x0];
        len = cstart[r0 + (uint32_t)U.x1 + 1u] - b;
    }
    const uint32_t end = wave_inclusive_scan(len);  // flattened end of the lane's row
    const uint32_t total = min(__builtin_amdgcn_readlane(end, 63), hi);
    const uint32_t shift = b - (end - len);         // slot = flattened position + shift
    auto fetch = [&](uint32_t f0, float4(&r)[NV], uint32_t &slot) {
        const uint32_t f = f0 + lane;
        uint32_t row = 0;
        for (int k = 0; k < nrows; ++k)
            row += f >= (uint32_t)__builtin_amdgcn_readlane(end, k) ? 1u : 0u;
        const uint32_t sh = (uint32_t)__shfl((int)shift, (int)min(row, 63u), 64);
        if (f < total) {
            slot = f + sh;
#pragma unroll
            for (int v = 0; v < NV; ++v)
                r[v] = src[v][slot];
        } else {
            slot = 0xFFFFFFFFu;
            r[0] = make_float4(__builtin_inff(), __builtin_inff(), __builtin_inff(), 0.0f);
#pragma unroll
            for (int v = 1; v < NV; ++v)
                r[v] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    };
    const uint32_t nch = total > lo ? (total - lo + 63u) >> 6 : 0u;
    float4 nr[NV];
    uint32_t ns = 0;
    if (nch)
        fetch(lo, nr, ns);
    for (uint32_t c = 0; c < nch; ++c) {
        const uint32_t bs = c & 1u;
#pragma unroll
        for (int v = 0; v < NV; ++v)
            sb.rec[bs][v][lane] = nr[v];
        sb.slot[bs][lane] = ns;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (c + 1u < nch)
            fetch(lo + ((c + 1u) << 6), nr, ns);
        const uint32_t cnt = min(64u, total - lo - (c << 6));
        const uint32_t cnt4 = (cnt + 3u) & ~3u;
        for (uint32_t k = 0; k < cnt4; k += 4u) {
            // all four records and slots in registers before the first use
            float4 q[4][NV];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    q[i][v] = sb.rec[bs][v][k + i];
            const uint4 j4 = *reinterpret_cast<const uint4 *>(&sb.slot[bs][k]);
            visit(j4.x, q[0]);
            visit(j4.y, q[1]);
            visit(j4.z, q[2]);
            visit(j4.w, q[3]);
        }
    }
}

// ------------------------------------------------------------ part plan

// Walk length of every wave task (the same decision the stencil passes
// make: union count, or the largest own stencil), and their sum per scene:
// the block's waves add theirs in LDS first, so runs of one scene cost one
// global atomic per block.
constexpr int kPlanBlock = 1024;

__global__ __launch_bounds__(kPlanBlock) void k_plan_parts(const uint32_t *__restrict__ cellcB,
                                                           const uint32_t *__restrict__ cstart, SceneSet S,
                                                           uint32_t *__restrict__ measure,
                                                           unsigned long long *__restrict__ total,
                                                           int4 *__restrict__ walk, uint32_t *__restrict__ own_len)
{
    __shared__ uint32_t s_scene[kPlanBlock / 64], s_len[kPlanBlock / 64];
    const uint32_t wave = threadIdx.x >> 6;
    Lane L;
    bool active;
    if (wave_lanes<kPlanBlock>(S, L, active)) {
        const StepParams &P = *L.P;
        const uint32_t cc = cellcB[L.slot];
        UnionRows U{0, 0, 0, 0, 0, 0};
        uint32_t ucount = 0, len, own = kNoOwn;
        const bool use_union = plan_union(cstart, cc, P, U, ucount, own) && !__any((cc & kSlowBit) != 0u);
        if (use_union) {
            len = ucount;
        } else {
            // Per-lane walks cost about three times as much per candidate
            // as the wave-uniform walk (divergent gathers, no shared loads),
            // so they count three times toward the part size; with slow
            // lanes (m_ij counted from the 27 buckets per neighbour, the
            // wave waiting on them) more again.
            if (__builtin_amdgcn_readfirstlane(own) == kNoOwn)  // wave-uniform
                own = stencil_length(cstart, cc, P);
            const uint32_t weight = __any((cc & kSlowBit) != 0u) ? kSlowWalkWeight : kLaneWalkWeight;
            len = weight * (uint32_t)wave_max_i((int)own);
            if (active)
                own_len[L.slot] = own;
        }
        if ((threadIdx.x & 63u) == 0u) {
            walk[2u * L.task] = make_int4(U.x0, U.x1, U.y0, U.y1);
            walk[2u * L.task + 1u] = make_int4(U.z0, U.z1, (int)ucount, use_union ? 1 : 0);
        }
        if ((threadIdx.x & 63u) == 0u) {
            measure[L.task] = len;
            s_scene[wave] = S.tasks[L.task].scene;
            s_len[wave] = len;
        }
    } else if ((threadIdx.x & 63u) == 0u) {
        s_scene[wave] = 0xFFFFFFFFu;
        s_len[wave] = 0u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t scene = s_scene[0];
        unsigned long long sum = 0;
        for (int k = 0; k < kPlanBlock / 64; ++k) {
            if (s_scene[k] != scene) {
                if (scene != 0xFFFFFFFFu)
                    atomicAdd(&total[scene], sum);
                scene = s_scene[k];
                sum = 0;
            }
            sum += s_len[k];
        }
        if (scene != 0xFFFFFFFFu)
            atomicAdd(&total[scene], sum);
    }
}

// Parts per wave task: the part size grows beyond kPartCandidates when the
// scene's total walk over its part room would exceed it.
__global__ void k_count_parts(const uint32_t *__restrict__ measure, SceneSet S,
                              const unsigned long long *__restrict__ total, uint32_t *__restrict__ nparts)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= S.nwaves)
        return;
    const uint32_t scene = S.tasks[g].scene;
    const StepParams &P = S.params[scene];
    const unsigned long long waves = (P.p_end - P.p_begin + 63u) / 64u;
    const unsigned long long room = part_room(waves);
    const unsigned long long size = max((unsigned long long)kPartCandidates, (total[scene] + room - 1u) / room);
    const unsigned long long p = (measure[g] + size - 1u) / size;
    const uint32_t np = (uint32_t)min(p, (unsigned long long)kMaxParts);
    nparts[g] = np > 1u ? np : 0u;  // an unsplit task is not queued
}

__global__ void k_emit_parts(const uint32_t *__restrict__ off, uint32_t ntasks, uint32_t *__restrict__ tasks,
                             uint32_t *__restrict__ queue, unsigned long long *__restrict__ total,
                             uint32_t nscenes)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) {
        queue[0] = 0u;
        queue[1] = 0u;
    }
    for (uint32_t k = g; k < nscenes; k += gridDim.x * blockDim.x)
        total[k] = 0ull;  // for the next step's plan
    if (g >= ntasks)
        return;
    const uint32_t b = off[g], e = off[g + 1];
    for (uint32_t p = 0; b + p < e; ++p)
        tasks[b + p] = (g << kPartBits) | p;
}

void launch_plan_parts(hipStream_t st, const ParticleArrays &a, const CellArrays &c, const SceneSet &S,
                       const PartPlan &pp)
{
    hipLaunchKernelGGL(k_plan_parts, dim3(blocks_for_waves(S.nwaves, kPlanBlock)), dim3(kPlanBlock), 0, st,
                       a.cellcB, c.cstart, S, pp.measure, pp.total, pp.walk, pp.own);
    hipLaunchKernelGGL(k_count_parts, dim3(blocks_for(S.nwaves)), dim3(kBlock), 0, st, pp.measure, S, pp.total,
                       pp.nparts);
    CellArrays cs;
    cs.counts = pp.nparts;
    cs.cstart = pp.off;
    cs.partials = pp.scan_partials;
    cs.big = nullptr;
    cs.big_cap = 0;
    cs.ncells_padded = pp.padded;
    launch_scan(st, cs, /*clear_counts=*/true);  // nparts must be zero for the next step
    hipLaunchKernelGGL(k_emit_parts, dim3(blocks_for(S.nwaves)), dim3(kBlock), 0, st, pp.off, S.nwaves, pp.tasks,
                       pp.queue, pp.total, pp.nscenes);
}

// Runs task(w, part, nparts) for the work the calling wave takes.  Without
// a plan each wave runs wave task w of its own index, unsplit.  With one,
// each wave first runs its static, XCD-contiguous share of the unsplit
// tasks -- wave v takes tasks v, v + G, ... that are not split, with no
// queue traffic, neighbouring tasks sharing an L2 -- then pulls the parts
// of split tasks from the queue counter until that list is exhausted
// (every wave exits then), which also evens out the static share.  All 64
// lanes must be live on entry.
template <int BLOCK, class Task>
__device__ inline void for_each_part(const SceneSet &S, const PartPlan &pp, uint32_t *queue, Task task)
{
    const uint32_t first = __builtin_amdgcn_readfirstlane(xcd_block() * (BLOCK / 64) + (threadIdx.x >> 6));
    if (!queue) {
        if (first < S.nwaves)
            task(first, 0u, 1u);
        return;
    }
    const uint32_t total = __builtin_amdgcn_readfirstlane(pp.off[S.nwaves]);  // parts of split tasks
    const uint32_t G = gridDim.x * (BLOCK / 64);
    for (uint32_t w = first; w < S.nwaves; w += G)
        if (total == 0u || __builtin_amdgcn_readfirstlane(pp.off[w + 1] - pp.off[w]) == 0u)
            task(w, 0u, 1u);
    if (total == 0u)
        return;
    for (;;) {
        uint32_t t = 0;
        if ((threadIdx.x & 63u) == 0u)
            t = atomicAdd(queue, 1u);
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= total)
            return;
        const uint32_t packed = __builtin_amdgcn_readfirstlane(pp.tasks[t]);
        const uint32_t w = packed >> kPartBits, part = packed & (kMaxParts - 1u);
        const uint32_t nparts = __builtin_amdgcn_readfirstlane(pp.off[w + 1] - pp.off[w]);
        task(w, part, nparts);
    }
}

__device__ inline float load_agent(const float *p)
{
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t *>(p), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT));
}

// Part `part` of wave task w stores its N partial sums; the last part to
// finish gets every part's sums added in part order in v and returns true
// (the others return false).  Every lane of the wave calls it.
template <int N>
__device__ inline bool combine_parts(const PartPlan &pp, uint32_t w, uint32_t part, uint32_t nparts, float (&v)[N])
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t base = __builtin_amdgcn_readfirstlane(pp.off[w]);
    float *mine = reinterpret_cast<float *>(pp.buf + ((size_t)base + part) * 64u + lane);
#pragma unroll
    for (int k = 0; k < N; ++k)
        mine[k] = v[k];
    __threadfence();
    uint32_t prev = 0;
    if (lane == 0u)
        prev = atomicAdd(&pp.done[w], 1u);
    prev = __builtin_amdgcn_readfirstlane(prev);
    if (prev + 1u != nparts)
        return false;
    __threadfence();
    for (uint32_t q = 0; q < nparts; ++q) {
        const float *src = reinterpret_cast<const float *>(pp.buf + ((size_t)base + q) * 64u + lane);
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const float x = load_agent(src + k);
            v[k] = q == 0u ? x : v[k] + x;
        }
    }
    if (lane == 0u)
        pp.done[w] = 0u;  // ready for the next pass
    return true;
}

// --------------------------------------------------------------- density
// MODE 0: density + pressure.  MODE 1: also the neighbour count per
// particle.  MODE 2: also the neighbour ids (multiplicity-repeated).  MODE
// 3: MODE 0 plus the neighbour count (with multiplicity) of the production
// walks.  Writes rho, posR = (x, 0.5 / rho) and velR = (v, p) for the force
// pass, and in MODE 0 / 3 the neighbour lists of unsplit per-lane tasks.
template <int MODE, int BLOCK>
__global__ __launch_bounds__(BLOCK, 6) void k_density(
    const float4 *__restrict__ pos, const float4 *__restrict__ vel,
    const uint32_t *__restrict__ cellc, const uint32_t *__restrict__ cstart,
    float *__restrict__ rho, float4 *__restrict__ posR, float4 *__restrict__ velR, SceneSet S,
    NeighborDump dump, PartPlan pp, NeighborLists nl)
{
    constexpr bool kCount = MODE >= 1;
    for_each_part<BLOCK>(S, pp, pp.queue, [&](uint32_t wt, uint32_t part, uint32_t nparts) {
        Lane L;
        bool active;
        wave_task(S, wt, L, active);
        // a copy: read once (scalar loads) before the task's first store
        const StepParams P = *L.P;
        const uint32_t s = L.slot;
        const float4 pi = pos[s];
        const uint32_t cc = cellc[s];
        const bool slow = (cc & kSlowBit) != 0u;
        // the reference cell (three IEEE divisions) only matters for m_ij
        RefCell ci{0, 0, 0};
        if (slow)
            ci = ref_cell(pi, P);
        const float h2 = P.h2, mp6 = P.mp6;
        float acc = 0.0f;
        uint32_t cnt = 0;
        uint32_t listed = kNoList;
        uint64_t w = (MODE == 2 && active) ? dump.off[s] : 0;
        // true when q passes (and then adds its term)
        auto body = [&](uint32_t j, float4 q) -> bool {
            float dx, dy, dz;
            float d2 = dist2_ref(pi, q, dx, dy, dz);
            if (d2 < h2) {
                int m = 1;
                if (slow) {
                    int cx = ci.x, cy = ci.y, cz = ci.z;
                    asm volatile("" : "+v"(cx), "+v"(cy), "+v"(cz));
                    m = bucket_multiplicity(cx, cy, cz, __float_as_uint(q.w) & 0xFFFFu);
                }
                float t = h2 - d2;
                float term = mp6 * ((t * t) * t);
                acc += (float)m * term;
                if (kCount)
                    cnt += (uint32_t)m;
                if (MODE == 2 && active) {
                    uint32_t id = __float_as_uint(vel[j].w);
                    for (int k = 0; k < m; ++k)
                        dump.out[w++] = id;
                }
                return true;
            }
            return false;
        };
        // Waves with a slow-multiplicity lane (rare where the union pays)
        // take the per-lane walk, so the union walk needs no m_ij.  A split
        // task walks its slice [part, part + 1) / nparts of the sequence.
        UnionRows U;
        uint32_t ucount = 0;
        if ((MODE == 0 || MODE == 3) && walk_of(pp, wt, cstart, cc, P, slow, U, ucount)) {
            // Wave-uniform candidate j: scalar loads, four records at a time
            // (density is too short per candidate to feed from LDS).
            const uint32_t lo = (uint32_t)((uint64_t)ucount * part / nparts);
            const uint32_t hi = (uint32_t)((uint64_t)ucount * (part + 1u) / nparts);
            auto cand = [&](uint32_t j, float4 q, auto self_tag) {
                float dx, dy, dz;
                const float d2 = dist2_ref(pi, q, dx, dy, dz);
                // Branch-free: h2 - d2 > 0 exactly when d2 < h2 (the float
                // difference of unequal values is nonzero), so clamping it
                // at 0 turns a failing candidate into a +0 term, which
                // leaves acc unchanged bit for bit.  (float)1 * term == term.
                float t = fmaxf(h2 - d2, 0.0f);
                if (decltype(self_tag)::value)
                    t = (j != s) ? t : 0.0f;
                acc = acc + mp6 * ((t * t) * t);
                if (kCount)
                    cnt += t > 0.0f ? 1u : 0u;
            };
            scalar_union_walk(cstart, U, P, lo, hi, L.first, [&](uint32_t j0, uint32_t j1, auto self_tag) {
                const float4 *p = pos + j0;
                uint32_t j = j0;
                for (; j + 4u <= j1; j += 4u, p += 4) {
                    const float4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
                    cand(j, q0, self_tag);
                    cand(j + 1u, q1, self_tag);
                    cand(j + 2u, q2, self_tag);
                    cand(j + 3u, q3, self_tag);
                }
                for (; j < j1; ++j, ++p)
                    cand(j, p[0], self_tag);
            });
        } else if ((MODE == 0 || MODE == 3) && nl.list && nparts == 1u) {
            // Per-lane walk that also records the lane's neighbours, in
            // walk order (rows in order, slots ascending).
            uint32_t nn = 0;
            // Without slow lanes every neighbour counts once, and the sum
            // runs branch-free: a failing candidate, or the particle
            // itself, adds +0, which leaves acc unchanged bit for bit (as
            // in the union walk), so both forms give the same sums.
            const bool any_slow = __any(slow);
            // On fine grids (R >= 2: rows of a few candidates) each
            // neighbour is appended as it passes; on cells of h (longer
            // rows, most candidates failing) a row's passes are collected
            // in a bitmask and appended afterwards, which issues fewer
            // stores.  The order is the same.
            const bool direct = P.reach >= 2;
            uint32_t *lp = nl.list + s;  // list entry nn (list_at's layout)
            auto fast = [&](float4 q) -> bool {
                float dx, dy, dz;
                const float d2 = dist2_ref(pi, q, dx, dy, dz);
                const float t = fmaxf(h2 - d2, 0.0f);
                acc = acc + mp6 * ((t * t) * t);
                if (kCount)
                    cnt += t > 0.0f ? 1u : 0u;
                return t > 0.0f;
            };
            auto append = [&](uint32_t j) {
                if (nn < kListMax) {
                    *lp = j;
                    lp += nl.stride;
                }
                ++nn;
            };
            // candidates [j0, j1), appended as they pass
            auto run_direct = [&](uint32_t j0, uint32_t j1) {
                uint32_t j = j0;
                for (; j + 4u <= j1; j += 4u) {  // four candidates in flight
                    const float4 q0 = pos[j], q1 = pos[j + 1u], q2 = pos[j + 2u], q3 = pos[j + 3u];
                    if (fast(q0))
                        append(j);
                    if (fast(q1))
                        append(j + 1u);
                    if (fast(q2))
                        append(j + 2u);
                    if (fast(q3))
                        append(j + 3u);
                }
                for (; j < j1; ++j)
                    if (fast(pos[j]))
                        append(j);
            };
            // candidates [j0, j1) of the segment starting at g, into its mask
            auto run_mask = [&](uint32_t j0, uint32_t j1, uint32_t g, uint64_t &mask) {
                uint32_t j = j0;
                for (; j + 4u <= j1; j += 4u) {
                    const float4 q0 = pos[j], q1 = pos[j + 1u], q2 = pos[j + 2u], q3 = pos[j + 3u];
                    mask |= fast(q0) ? 1ull << (j - g) : 0ull;
                    mask |= fast(q1) ? 2ull << (j - g) : 0ull;
                    mask |= fast(q2) ? 4ull << (j - g) : 0ull;
                    mask |= fast(q3) ? 8ull << (j - g) : 0ull;
                }
                for (; j < j1; ++j)
                    mask |= fast(pos[j]) ? 1ull << (j - g) : 0ull;
            };
            // Once a lane has more neighbours than a list holds, the force
            // pass walks this task anyway: the wave stops recording (and
            // only sums) for the rest of its rows.
            // The cut stencil: it drops only candidates that fail, each of
            // which adds +0 below, so the sums equal the whole stencil's.
            bool listing = true;
            const LaneTrim T = lane_trim(pi, cc, P);
            for_each_cut_row(cstart, cc, T, P, [&](uint32_t b, uint32_t e) {
                if (!any_slow && direct) {
                    // the row split at the lane's own slot (when it holds
                    // it), so no candidate needs a self test
                    const uint32_t m = (s >= b && s < e) ? s : e;
                    run_direct(b, m);
                    run_direct(min(m + 1u, e), e);
                    return;
                }
                for (uint32_t g = b; g < e; g += 64u) {
                    const uint32_t ge = min(g + 64u, e);
                    uint64_t mask = 0;
                    uint32_t j = g;
                    if (!any_slow) {
                        // the segment split at the lane's own slot
                        const uint32_t m = (s >= g && s < ge) ? s : ge;
                        run_mask(g, m, g, mask);
                        run_mask(min(m + 1u, ge), ge, g, mask);
                        j = ge;
                    }
                    for (; j + 1u < ge; j += 2u) {  // two candidates in flight
                        const float4 q0 = pos[j];
                        const float4 q1 = pos[j + 1u];
                        if (j != s && body(j, q0))
                            mask |= 1ull << (j - g);
                        if (j + 1u != s && body(j + 1u, q1))
                            mask |= 2ull << (j - g);
                    }
                    if (j < ge && j != s && body(j, pos[j]))
                        mask |= 1ull << (j - g);
                    if (!listing)
                        continue;
                    while (mask != 0ull) {
                        const uint32_t k = (uint32_t)__builtin_ctzll(mask);
                        mask &= mask - 1ull;
                        if (nn < kListMax)
                            nl.list[list_at(nn, s, nl.stride)] = g + k;
                        ++nn;
                    }
                    listing = !__any(nn > kListMax);
                }
            });
            if (!any_slow && direct)
                listing = !__any(nn > kListMax);
            listed = listing ? nn : kNoList;
        } else {
            uint32_t lo = 0, hi = 0xFFFFFFFFu;
            if (nparts > 1u) {
                const uint32_t own = own_length(pp, s, cstart, cc, P);
                lo = (uint32_t)((uint64_t)own * part / nparts);
                hi = (uint32_t)((uint64_t)own * (part + 1u) / nparts);
            }
            for_each_row_span(cstart, cc, P, lo, hi, [&](uint32_t b, uint32_t e) {
                uint32_t j = b;
                for (; j + 1u < e; j += 2u) {  // two candidates in flight
                    const float4 q0 = pos[j];
                    const float4 q1 = pos[j + 1u];
                    if (j != s)
                        body(j, q0);
                    if (j + 1u != s)
                        body(j + 1u, q1);
                }
                if (j < e && j != s)
                    body(j, pos[j]);
            });
        }
        if (nparts > 1u) {
            if constexpr (kCount) {
                float v[2] = {acc, (float)cnt};  // counts stay far below 2^24: exact
                if (!combine_parts<2>(pp, wt, part, nparts, v))
                    return;
                acc = v[0];
                cnt = (uint32_t)v[1];
            } else {
                float v[1] = {acc};
                if (!combine_parts<1>(pp, wt, part, nparts, v))
                    return;
                acc = v[0];
            }
        }
        if (!active)
            return;
        const float r = acc + P.self_dens;
        const float p = P.gas * (r - P.rho0);
        rho[s] = r;
        posR[s] = make_float4(pi.x, pi.y, pi.z, 0.5f / r);
        const float4 vs = vel[s];
        velR[s] = make_float4(vs.x, vs.y, vs.z, p);
        if (nl.count)
            nl.count[s] = listed;
        if (kCount)
            dump.count[s] = cnt;
    });
}

// Waves of `kernel` (at `block` threads) the device holds at once, from its
// register and LDS use.
static uint32_t occupancy_waves(const void *kernel, int block)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0) != hipSuccess || nb <= 0)
        return resident_waves();
    return std::min<uint32_t>(resident_waves(), (uint32_t)nb * (uint32_t)(block / 64) * (resident_waves() / 32u));
}

// Waves of a stencil pass: one per wave task without a plan; with one, a
// persistent grid of the waves the device holds at once for this kernel
// (a wave launched beyond them would start only as the first ones finish,
// with a full share of a static assignment), and at most 4 per wave task
// (more would only pull from an empty queue: a small scene splits into
// about as many parts as it has wave tasks).
inline uint32_t stencil_waves(const SceneSet &S, const PartPlan *pp, uint32_t resident)
{
    if (!pp)
        return S.nwaves;
    return std::min<uint32_t>(std::min<uint32_t>(pp->cap, resident), 4u * std::max(S.nwaves, 1u));
}

template <int MODE, int BLOCK>
static void launch_density_block(hipStream_t st, const ParticleArrays &a, const CellArrays &c,
                                 const SceneSet &S, const NeighborDump &d, const PartPlan *pp,
                                 const NeighborLists *lists)
{
    PartPlan p{};
    if (pp)
        p = *pp;
    NeighborLists nl{nullptr, nullptr, 0u};
    if (lists)
        nl = *lists;
    static const uint32_t resident = occupancy_waves(reinterpret_cast<const void *>(&k_density<MODE, BLOCK>), BLOCK);
    hipLaunchKernelGGL((k_density<MODE, BLOCK>), dim3(std::max(1u, blocks_for_waves(stencil_waves(S, pp, resident), BLOCK))),
                       dim3(BLOCK), 0, st, a.posB, a.velB, a.cellcB, c.cstart, a.rhoB, a.posR, a.velR, S, d, p, nl);
}

void launch_density(hipStream_t st, const ParticleArrays &a, const CellArrays &c, const SceneSet &S,
                    const NeighborDump *dump, const PartPlan *parts, int block, const NeighborLists *lists)
{
    NeighborDump d = dump ? *dump : NeighborDump{nullptr, nullptr, nullptr};
    if (dump && dump->production)  // the step's own walks, plus the count
        return launch_density_block<3, kStencilBlock>(st, a, c, S, d, parts, lists);
    if (dump)  // the dump modes walk every lane's whole stencil in one task
        return dump->out ? launch_density_block<2, kStencilBlock>(st, a, c, S, d, nullptr, nullptr)
                         : launch_density_block<1, kStencilBlock>(st, a, c, S, d, nullptr, nullptr);
    switch (block) {
    case 64:
        return launch_density_block<0, 64>(st, a, c, S, d, parts, lists);
    case 256:
        return launch_density_block<0, 256>(st, a, c, S, d, parts, lists);
    default:
        return launch_density_block<0, kStencilBlock>(st, a, c, S, d, parts, lists);
    }
}

// ----------------------------------------------------- force + integrate
// Per neighbour (reference, src/sph.cpp:103-118):
//   pressure = -dir m (p_i + p_j) / (2 rho_j) spikyGrad (h - r)^2
//   visco    = mu m (v_j - v_i) / rho_j spikyLap (h - r)
// with dir = (x_j - x_i) / r.  Both are evaluated as one scalar coefficient
// times a vector, from 0.5 / rho_j precomputed by the density pass.
__device__ inline void integrate_and_walls(float4 &x, float4 &v, float fx, float fy, float fz,
                                           float rho, const StepParams &P)
{
    // src/sph.cpp:145-150: a = F / rho + (0, g, 0); v += a dt; x += v dt
    float ax = fx / rho + 0.0f;
    float ay = fy / rho + P.g;
    float az = fz / rho + 0.0f;
    v.x += ax * P.dt;
    v.y += ay * P.dt;
    v.z += az * P.dt;
    x.x += v.x * P.dt;
    x.y += v.y * P.dt;
    x.z += v.z * P.dt;
    // src/sph.cpp:153-176, planes L = lo + h, U = hi - h: floor, ceiling,
    // x-low, x-high, z-low, z-high.
    const float e = P.restitution, off = P.offset;
    if (x.y < P.wall_lo[1]) {
        x.y = -x.y + 2.0f * P.wall_lo[1] + off;
        v.y = -v.y * e;
    }
    if (x.y > P.wall_hi[1]) {
        x.y = -x.y + 2.0f * P.wall_hi[1] - off;
        v.y = -v.y * e;
    }
    if (x.x < P.wall_lo[0]) {
        x.x = -x.x + 2.0f * P.wall_lo[0] + off;
        v.x = -v.x * e;
    }
    if (x.x > P.wall_hi[0]) {
        x.x = -x.x + 2.0f * P.wall_hi[0] - off;
        v.x = -v.x * e;
    }
    if (x.z < P.wall_lo[2]) {
        x.z = -x.z + 2.0f * P.wall_lo[2] + off;
        v.z = -v.z * e;
    }
    if (x.z > P.wall_hi[2]) {
        x.z = -x.z + 2.0f * P.wall_hi[2] - off;
        v.z = -v.z * e;
    }
}

// The force of one passing neighbour, accumulated into (fx, fy, fz):
//   cp = -m spikyGrad (p_i + p_j) m_ij (0.5 / rho_j) (h - r)^2 / r
//   cv = 2 mu m spikyLap (h - r) m_ij (0.5 / rho_j)
//   F += cp d + cv (v_j - v_i)
// q = (x_j, 0.5 / rho_j), w = (v_j, p_j), d = x_j - x_i, am = m_ij (0.5 / rho_j).
__device__ inline void force_pair(float d2, float dx, float dy, float dz, float am, float4 w, float4 vi,
                                  float p_i, float h, float neg_m_spg, float mu2_m_lap, float &fx,
                                  float &fy, float &fz)
{
    // r = sqrt(d2) correctly rounded and hr = h - r exactly as the
    // reference forms them: near the support edge h - r cancels and any
    // error in r would be amplified.
    const float r = sqrt_rn(d2);
    const float rinv = __builtin_amdgcn_rcpf(r);
    const float hr = h - r;
    const float cp = neg_m_spg * (p_i + w.w) * am * (hr * hr) * rinv;
    const float cv = mu2_m_lap * hr * am;
    fx = fmaf(cp, dx, fx);
    fy = fmaf(cp, dy, fy);
    fz = fmaf(cp, dz, fz);
    fx = fmaf(cv, w.x - vi.x, fx);
    fy = fmaf(cv, w.y - vi.y, fy);
    fz = fmaf(cv, w.z - vi.z, fz);
}

// Two walks, chosen per wave by plan_union, with identical per-lane sums:
//  * union: one wave-uniform candidate at a time (staged through LDS), the
//    pair term predicated on the lane's own test;
//  * per lane: candidates filtered kForceChunk at a time into a per-lane
//    bitmask, the terms then run only over the set bits, so a wave pays for
//    the most neighbours any of its lanes has in the chunk.
template <bool WRITE_FORCE, bool NEXT_KEYS, int BLOCK>
__global__ __launch_bounds__(BLOCK, 5) void k_force_integrate(
    const float4 *__restrict__ posB, const float4 *__restrict__ velB,
    const float4 *__restrict__ posR, const float4 *__restrict__ velR,
    const uint32_t *__restrict__ cellcB, const uint32_t *__restrict__ cstart,
    const float *__restrict__ rhoB, float4 *__restrict__ posA, float4 *__restrict__ velA,
    uint32_t *__restrict__ cellcA, uint32_t *__restrict__ counts,
    float4 *__restrict__ force_out, SceneSet S, PartPlan pp, NeighborLists nl)
{
    __shared__ StageBuf<2> stage[BLOCK / 64];
    for_each_part<BLOCK>(S, pp, pp.queue ? pp.queue + 1 : nullptr, [&](uint32_t wt, uint32_t part,
                                                                      uint32_t nparts) {
        Lane L;
        bool active;
        wave_task(S, wt, L, active);
        const StepParams &P = *L.P;
        const uint32_t s = L.slot;
        float4 pi = posR[s];        // w = 0.5 / rho_i
        const float4 vr = velR[s];  // w = p_i
        float4 vi = make_float4(vr.x, vr.y, vr.z, velB[s].w);  // w = particle id
        const uint32_t cc = cellcB[s];
        const bool slow = (cc & kSlowBit) != 0u;
        // the reference cell (three IEEE divisions) only matters for m_ij
        RefCell ci{0, 0, 0};
        if (slow)
            ci = ref_cell(pi, P);
        const float p_i = vr.w;
        const float h = P.h, h2 = P.h2, neg_m_spg = P.neg_m_spg;
        const float mu2_m_lap = 2.0f * P.mu_m_lap;
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;

        // The term of neighbour j in the order the walks visit it.
        auto pair = [&](uint32_t j, float4 q, float4 w) {
            float dx, dy, dz;
            const float d2 = dist2_ref(pi, q, dx, dy, dz);
            float am = q.w;  // 0.5 / rho_j
            if (slow) {
                int cx = ci.x, cy = ci.y, cz = ci.z;
                asm volatile("" : "+v"(cx), "+v"(cy), "+v"(cz));
                am = q.w * (float)bucket_multiplicity(cx, cy, cz, __float_as_uint(posB[j].w) & 0xFFFFu);
            }
            force_pair(d2, dx, dy, dz, am, w, vi, p_i, h, neg_m_spg, mu2_m_lap, fx, fy, fz);
        };
        // Every lane holds its neighbour list (the density pass walked this
        // task per lane, unsplit): visit the listed slots four at a time,
        // the next group of four slots loaded while the current one runs;
        // entries past a lane's list stand for its own slot and add nothing.
        const uint32_t listed = (nl.count && nparts == 1u) ? nl.count[s] : kNoList;
        const bool use_list = nl.count && nparts == 1u && __all(listed != kNoList);
        UnionRows U;
        uint32_t ucount = 0;
        if (use_list) {
            const uint32_t groups = ((uint32_t)wave_max_i((int)listed) + 3u) >> 2;
            const uint32_t *ent = nl.list + s;
            const size_t stride = nl.stride;
            // groups <= kListMax / 4, so every entry loaded lies inside the
            // list array; entries a lane has not written are never used
            auto group = [&](uint32_t g) {
                const uint32_t *e0 = ent + (size_t)(4u * g) * stride;
                return make_uint4(e0[0], e0[stride], e0[2 * stride], e0[3 * stride]);
            };
            uint4 next = group(0);
            for (uint32_t g = 0; g < groups; ++g) {
                const uint32_t k = 4u * g;
                const uint4 cur = next;
                if (g + 1u < groups)
                    next = group(g + 1u);
                const uint32_t j0 = k < listed ? cur.x : s, j1 = k + 1u < listed ? cur.y : s;
                const uint32_t j2 = k + 2u < listed ? cur.z : s, j3 = k + 3u < listed ? cur.w : s;
                const float4 q0 = posR[j0], w0 = velR[j0], q1 = posR[j1], w1 = velR[j1];
                const float4 q2 = posR[j2], w2 = velR[j2], q3 = posR[j3], w3 = velR[j3];
                if (k < listed)
                    pair(j0, q0, w0);
                if (k + 1u < listed)
                    pair(j1, q1, w1);
                if (k + 2u < listed)
                    pair(j2, q2, w2);
                if (k + 3u < listed)
                    pair(j3, q3, w3);
            }
        } else if (walk_of(pp, wt, cstart, cc, P, slow, U, ucount)) {  // as in k_density
            const float4 *const src[2] = {posR, velR};
            const uint32_t lo = (uint32_t)((uint64_t)ucount * part / nparts);
            const uint32_t hi = (uint32_t)((uint64_t)ucount * (part + 1u) / nparts);
            staged_union_walk<2>(cstart, U, P, lo, hi, src, stage[threadIdx.x >> 6],
                                 [&](uint32_t j, const float4(&r)[2]) {
                const float4 q = r[0], w = r[1];  // (x_j, 0.5 / rho_j), (v_j, p_j)
                float dx, dy, dz;
                const float d2 = dist2_ref(pi, q, dx, dy, dz);
                if (d2 < h2 && j != s)
                    force_pair(d2, dx, dy, dz, q.w, w, vi, p_i, h, neg_m_spg, mu2_m_lap, fx, fy, fz);  // am = q.w * 1
            });
        } else {
            uint32_t lo = 0, hi = 0xFFFFFFFFu;
            if (nparts > 1u) {
                const uint32_t own = own_length(pp, s, cstart, cc, P);
                lo = (uint32_t)((uint64_t)own * part / nparts);
                hi = (uint32_t)((uint64_t)own * (part + 1u) / nparts);
            }
            for_each_row_span(cstart, cc, P, lo, hi, [&](uint32_t b, uint32_t e) {
                for (uint32_t jj = b; jj < e; jj += kForceChunk) {
                    uint32_t mask = 0u;
#pragma unroll
                    for (int k = 0; k < kForceChunk; ++k) {
                        const uint32_t j = jj + (uint32_t)k;
                        const float4 q = posR[j < e ? j : e - 1u];
                        float dx, dy, dz;
                        const float d2 = dist2_ref(pi, q, dx, dy, dz);
                        const bool pass = (j < e) && (j != s) && (d2 < h2);
                        mask |= pass ? (1u << k) : 0u;
                    }
                    while (mask != 0u) {
                        const uint32_t k = (uint32_t)__builtin_ctz(mask);
                        mask &= mask - 1u;
                        const uint32_t j = jj + k;
                        pair(j, posR[j], velR[j]);  // (x_j, 0.5 / rho_j), (v_j, p_j)
                    }
                }
            });
        }
        if (nparts > 1u) {
            float v[3] = {fx, fy, fz};
            if (!combine_parts<3>(pp, wt, part, nparts, v))
                return;
            fx = v[0];
            fy = v[1];
            fz = v[2];
        }
        if (!active)
            return;
        if (WRITE_FORCE)
            force_out[s] = make_float4(fx, fy, fz, 0.0f);
        integrate_and_walls(pi, vi, fx, fy, fz, rhoB[s], P);
        if (NEXT_KEYS) {
            Classified c = classify(pi.x, pi.y, pi.z, P);
            pi.w = __uint_as_float(c.h16);
            cellcA[s] = c.cellc;
            atomicAdd(&counts[linear_cell(P, c.cellc)], 1u);  // no result: nothing waits on it
        } else {
            pi.w = 0.0f;
        }
        posA[s] = pi;
        velA[s] = vi;
    });
}

template <int BLOCK>
static void launch_force_block(hipStream_t st, const ParticleArrays &a, const CellArrays &c, const SceneSet &S,
                               float4 *force_out, bool next_keys, const PartPlan *pp, const NeighborLists *lists)
{
    PartPlan p{};
    if (pp)
        p = *pp;
    NeighborLists nl{nullptr, nullptr, 0u};
    if (lists)
        nl = *lists;
    static const uint32_t resident =
        occupancy_waves(reinterpret_cast<const void *>(&k_force_integrate<false, true, BLOCK>), BLOCK);
    dim3 g(std::max(1u, blocks_for_waves(stencil_waves(S, pp, resident), BLOCK))), b(BLOCK);
#define SPH_LAUNCH_FI(WF, NK)                                                                          \
    hipLaunchKernelGGL((k_force_integrate<WF, NK, BLOCK>), g, b, 0, st, a.posB, a.velB, a.posR, a.velR, \
                       a.cellcB, c.cstart, a.rhoB, a.posA, a.velA, a.cellcA, c.counts,         \
                       force_out, S, p, nl)
    if (force_out && next_keys)
        SPH_LAUNCH_FI(true, true);
    else if (force_out)
        SPH_LAUNCH_FI(true, false);
    else if (next_keys)
        SPH_LAUNCH_FI(false, true);
    else
        SPH_LAUNCH_FI(false, false);
#undef SPH_LAUNCH_FI
}

void launch_force_integrate(hipStream_t st, const ParticleArrays &a, const CellArrays &c,
                            const SceneSet &S, float4 *force_out, bool next_keys, const PartPlan *parts,
                            int block, const NeighborLists *lists)
{
    switch (block) {
    case 64:
        return launch_force_block<64>(st, a, c, S, force_out, next_keys, parts, lists);
    case 256:
        return launch_force_block<256>(st, a, c, S, force_out, next_keys, parts, lists);
    default:
        return launch_force_block<kStencilBlock>(st, a, c, S, force_out, next_keys, parts, lists);
    }
}

// ------------------------------------------------------- grid statistic
// Candidates per neighbour of a sample of every scene: its wave tasks
// 0, kGridStatEvery, 2 kGridStatEvery, ... (counted from the scene's own
// first slot, so the sample is the same in any batch) walk their lanes' whole
// stencils over the current cell ranges and B order, counting the
// candidates and those within h (the particle itself excluded); each
// wave adds its sums to stat[2 scene] and stat[2 scene + 1].  The ratio
// tells a state uniform at the scale of h (a jittered lattice of spacing
// h / 2 reads 7.2 at cell scale 1, 4.2 at cell scale 1/2) from one clumped
// below it (near 1).  Lanes with more than kGridStatTests candidates (the
// clumps of a collapsed flow, thousands each) test an evenly strided subset
// and scale their count, so the sample costs a few hundred tests a lane
// however dense the state; the uniform states the choice turns on have
// fewer candidates and are counted exactly.
__global__ __launch_bounds__(kBlock) void k_grid_stat(const float4 *__restrict__ pos,
                                                     const uint32_t *__restrict__ cellc,
                                                     const uint32_t *__restrict__ cstart, SceneSet S,
                                                     unsigned long long *__restrict__ stat)
{
    const uint32_t w = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
    if (w >= S.nwaves)
        return;
    Lane L;
    bool active;
    wave_task(S, w, L, active);
    const StepParams &P = *L.P;
    if (((L.first - P.p_begin) / 64u) % kGridStatEvery != 0u)
        return;  // the sample counts tasks from the scene's own first slot
    const uint32_t s = L.slot;
    uint32_t cand = 0, nbr = 0;
    if (active) {
        const float4 pi = pos[s];
        const float h2 = P.h2;
        const uint32_t cc = cellc[s];
        for_each_row(cstart, cc, P, [&](uint32_t b, uint32_t e) { cand += e - b; });
        const uint32_t stride = max((cand + kGridStatTests - 1u) / kGridStatTests, 1u);
        uint32_t q = 0;  // offset of the next tested candidate in the current row
        for_each_row(cstart, cc, P, [&](uint32_t b, uint32_t e) {
            for (; q < e - b; q += stride) {
                const uint32_t j = b + q;
                float dx, dy, dz;
                nbr += (j != s && dist2_ref(pi, pos[j], dx, dy, dz) < h2) ? 1u : 0u;
            }
            q -= e - b;
        });
        nbr *= stride;
    }
    unsigned long long c = cand, nb = nbr;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        c += __shfl_xor(c, d, 64);
        nb += __shfl_xor(nb, d, 64);
    }
    if ((threadIdx.x & 63u) == 0u) {
        const uint32_t scene = S.tasks[w].scene;
        atomicAdd(&stat[2u * scene], c);
        atomicAdd(&stat[2u * scene + 1u], nb);
    }
}

void launch_grid_stat(hipStream_t st, const ParticleArrays &a, const CellArrays &c, const SceneSet &S,
                      unsigned long long *stat, uint32_t nscenes)
{
    launch_fill_u32(st, (uint32_t *)stat, 0u, 4ull * nscenes);
    if (S.nwaves)
        hipLaunchKernelGGL(k_grid_stat, dim3(blocks_for_waves(S.nwaves, kBlock)), dim3(kBlock), 0, st, a.posB,
                           a.cellcB, c.cstart, S, stat);
}

// ------------------------------------------------------- drop-in output
// glm::translate(x) * glm::scale(vec3(s)) with glm's column arithmetic
// (src/sph.cpp:179-180): T[3] = I0 x + I1 y + I2 z + I3, then
// M[c] = T0 S[c][0] + T1 S[c][1] + T2 S[c][2] + T3 S[c][3].
__device__ inline void write_transform(float *__restrict__ xf, float x, float y, float z, float sc)
{
    const float pos[3] = {x, y, z};
    float T3[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        float a0 = (r == 0) ? 1.0f : 0.0f, a1 = (r == 1) ? 1.0f : 0.0f, a2 = (r == 2) ? 1.0f : 0.0f;
        float a3 = (r == 3) ? 1.0f : 0.0f;
        T3[r] = ((a0 * pos[0] + a1 * pos[1]) + a2 * pos[2]) + a3;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float s0 = (c == 0) ? sc : 0.0f, s1 = (c == 1) ? sc : 0.0f;
        const float s2 = (c == 2) ? sc : 0.0f, s3 = (c == 3) ? 1.0f : 0.0f;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float t0 = (r == 0) ? 1.0f : 0.0f, t1 = (r == 1) ? 1.0f : 0.0f, t2 = (r == 2) ? 1.0f : 0.0f;
            xf[c * 4 + r] = ((t0 * s0 + t1 * s1) + t2 * s2) + T3[r] * s3;
        }
    }
}

__global__ void k_h16_rank(const float4 *__restrict__ posB, uint32_t *__restrict__ counts16,
                           uint32_t *__restrict__ rank16, uint32_t n)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    uint32_t h = __float_as_uint(posB[s].w) & 0xFFFFu;
    rank16[s] = atomicAdd(&counts16[h], 1u);
}

void launch_h16_rank(hipStream_t st, const ParticleArrays &a, const CellArrays &c16, uint32_t n)
{
    // rankA is free once the force pass ran without next keys
    hipLaunchKernelGGL(k_h16_rank, dim3(blocks_for(n)), dim3(kBlock), 0, st, a.posB, c16.counts,
                       a.rankA, n);
}

__global__ void k_h16_scatter(const float4 *__restrict__ posB, const uint32_t *__restrict__ rank16,
                              const uint32_t *__restrict__ cstart16, uint32_t *__restrict__ perm16,
                              uint32_t n)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    uint32_t h = __float_as_uint(posB[s].w) & 0xFFFFu;
    perm16[cstart16[h] + rank16[s]] = s;
}

void launch_h16_scatter(hipStream_t st, const ParticleArrays &a, const CellArrays &c16, uint32_t n,
                        uint32_t *perm16)
{
    hipLaunchKernelGGL(k_h16_scatter, dim3(blocks_for(n)), dim3(kBlock), 0, st, a.posB, a.rankA,
                       c16.cstart, perm16, n);
}

// Output slot o: the particle perm16[o] lands at its bucket start plus the
// number of bucket members with a smaller id (ties of std::sort resolved by
// input order).  Buckets above kBigCellSortMax keep arrival order.
__global__ void k_pack_out(const uint32_t *__restrict__ perm16, const uint32_t *__restrict__ cstart16,
                           const float4 *__restrict__ posB, const float4 *__restrict__ velB,
                           const float *__restrict__ rhoB, const float4 *__restrict__ velR,
                           const float4 *__restrict__ force, const float4 *__restrict__ posA,
                           const float4 *__restrict__ velA, const float *__restrict__ aos_in,
                           float *__restrict__ aos_out, float *__restrict__ xf_out, uint32_t n,
                           float sc)
{
    uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n)
        return;
    const uint32_t s = perm16[o];
    const uint32_t h = __float_as_uint(posB[s].w) & 0xFFFFu;
    const uint32_t id = __float_as_uint(velB[s].w);
    const uint32_t b = cstart16[h], e = cstart16[h + 1];
    uint32_t d;
    if (e - b <= kBigCellSortMax) {
        uint32_t r = 0;
        for (uint32_t t = b; t < e; ++t)
            r += (__float_as_uint(velB[perm16[t]].w) < id) ? 1u : 0u;
        d = b + r;
    } else {
        d = o;
    }
    const float4 x = posA[s], v = velA[s], f = force[s];
    const float *src = aos_in + (size_t)id * 15;
    float *dst = aos_out + (size_t)d * 15;
    dst[0] = x.x;
    dst[1] = x.y;
    dst[2] = x.z;
    dst[3] = v.x;
    dst[4] = v.y;
    dst[5] = v.z;
    dst[6] = src[6];
    dst[7] = src[7];
    dst[8] = src[8];
    dst[9] = f.x;
    dst[10] = f.y;
    dst[11] = f.z;
    dst[12] = rhoB[s];
    dst[13] = velR[s].w;
    dst[14] = __uint_as_float(h);  // uint16 hash + zero padding (little endian)
    if (xf_out)
        write_transform(xf_out + (size_t)d * 16, x.x, x.y, x.z, sc);
}

void launch_pack_out(hipStream_t st, const ParticleArrays &a, const CellArrays &c16,
                     const uint32_t *perm16, const float4 *force, const void *aos_in, uint32_t n,
                     float xf_scale, void *aos_out, float *xf_out)
{
    hipLaunchKernelGGL(k_pack_out, dim3(blocks_for(n)), dim3(kBlock), 0, st, perm16, c16.cstart,
                       a.posB, a.velB, a.rhoB, a.velR, force, a.posA, a.velA,
                       (const float *)aos_in, (float *)aos_out, xf_out, n, xf_scale);
}

// Resident download: slots [first, first + n) of the A arrays (device
// order).  Bucket, density and pressure belong to the last step (B slot s is
// the same particle).
__global__ void k_pack_state(const float4 *__restrict__ posA, const float4 *__restrict__ velA,
                             const float4 *__restrict__ posB, const float *__restrict__ rhoB,
                             const float4 *__restrict__ velR, const float4 *__restrict__ force,
                             float *__restrict__ aos_out, float *__restrict__ xf_out, uint32_t first,
                             uint32_t n, float sc)
{
    uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n)
        return;
    const uint32_t s = first + o;
    const float4 x = posA[s], v = velA[s];
    const float4 f = force ? force[s] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float *dst = aos_out + (size_t)o * 15;
    dst[0] = x.x;
    dst[1] = x.y;
    dst[2] = x.z;
    dst[3] = v.x;
    dst[4] = v.y;
    dst[5] = v.z;
    dst[6] = 0.0f;
    dst[7] = 0.0f;
    dst[8] = 0.0f;
    dst[9] = f.x;
    dst[10] = f.y;
    dst[11] = f.z;
    dst[12] = rhoB[s];
    dst[13] = velR[s].w;
    dst[14] = __uint_as_float(__float_as_uint(posB[s].w) & 0xFFFFu);
    if (xf_out)
        write_transform(xf_out + (size_t)o * 16, x.x, x.y, x.z, sc);
}

void launch_pack_state(hipStream_t st, const ParticleArrays &a, const float4 *force, uint32_t first,
                       uint32_t n, float xf_scale, void *aos_out, float *xf_out)
{
    hipLaunchKernelGGL(k_pack_state, dim3(blocks_for(n)), dim3(kBlock), 0, st, a.posA, a.velA,
                       a.posB, a.rhoB, a.velR, force, (float *)aos_out, xf_out, first, n, xf_scale);
}

// ------------------------------------------------------------------ copy

__global__ void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride)
        dst[i] = src[i];
}

void launch_copy(hipStream_t st, const uint4 *src, uint4 *dst, size_t n16)
{
    size_t nb = (n16 + kBlock - 1) / kBlock;
    if (nb > 65536)
        nb = 65536;
    hipLaunchKernelGGL(k_copy, dim3((uint32_t)nb), dim3(kBlock), 0, st, src, dst, n16);
}

}  // namespace sph
