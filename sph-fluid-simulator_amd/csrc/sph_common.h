// sph_common.h -- data layout and per-particle classification shared by the
// HIP kernels and the host runtime (compiled for both sides).
//
// Reference semantics restated here (lijenicol/SPH-Fluid-Simulator):
//   getHash / getCell        src/neighborTable.cpp:5-17
//   16-bit bucket truncation src/Particle.h:9, src/sph.cpp:43,93
//
// Neighbour search.  The reference scans the 27 16-bit hash buckets around a
// particle's cell; a candidate j inside the support radius is counted once
// for every offset o in {-1,0,1}^3 with H16(C_i + o) == H16(C_j), i.e. with
// multiplicity
//     m_ij = #{ o : H16(C_i + o) == H16(C_j) }.
// This library instead sorts particles into a dense grid of cells of edge
// H = h * (1 + 2^-10) and scans the 3x3x3 block of grid cells around each
// particle; for every candidate with fl(d^2) < h^2 it applies m_ij exactly.
//   * Completeness: if fl(d^2) < fl(h*h) then |x_j - x_i| < h (1 + 2^-22)
//     per axis, and with H > h (1 + 2^-10) and grid coordinates computed in
//     double precision the two grid cells are adjacent (clamping to the grid
//     keeps adjacency), so every reference neighbour is a candidate.
//   * Multiplicity: when the 27 buckets around C_i are pairwise distinct and
//     q_i = fl(x_i / h) is at least eta(q_i) = 2^-21 + 2^-23 (|q_i| + 1) from
//     every reference cell boundary (nonzero integers), every neighbour j
//     has C_j adjacent to C_i and m_ij == 1 ("fast" particle).  Proof:
//     |x_j - x_i| < h (1 + 2^-22) and each quotient carries a relative error
//     below 2^-24, so |q_j - q_i| < 1 + 2^-22 + 2^-24 (2 |q_i| + 2) <= 1 + eta.
//     Otherwise the particle is flagged "slow" and m_ij is counted from the
//     27 buckets for every neighbour.  The bucket test is exact: two offsets
//     collide iff the six 16-bit vectors
//       A(cx-1)^A(cx), A(cx)^A(cx+1), B(cy-1)^B(cy), ..., Cz(cz)^Cz(cz+1)
//     (A(c) = c * 73856093 mod 2^16, ...) are linearly dependent over GF(2).
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define SPH_HD __host__ __device__
#else
#define SPH_HD
#endif

#pragma clang fp contract(off)

namespace sph {

// Grid-coordinate packing in one 32-bit word: gx 11 bits, gy 10, gz 10,
// bit 31 = slow-multiplicity flag.
constexpr int kGridMaxX = 2048;
constexpr int kGridMaxY = 1024;
constexpr int kGridMaxZ = 1024;
constexpr uint32_t kSlowBit = 0x80000000u;
constexpr double kCellMargin = 1.0 / 1024.0;  // H = h (1 + 2^-10)
constexpr int kMaxReach = 3;  // finest grid: cell edge h (1 + 2^-10) / 3
constexpr float kQMax = 8192.0f;              // |x/h| beyond this: always slow

// Per-scene constants used by every pass.  Derived on the host from the
// reference SPHSettings (src/SPHSystem.h:13-22) plus the wall box and grid.
struct StepParams {
    // reference settings, verbatim
    float h, h2, mass, mp6, self_dens, gas, rho0, visc, spiky_grad, spiky_lap, g;
    float dt;
    // derived products used by the force pass (host-rounded floats)
    float neg_m_spg;  // -(mass * spikyGrad)
    float mu_m_lap;   // (viscosity * mass) * spikyLap
    // wall reflection planes L = lo + h, U = hi - h, restitution, offset
    float wall_lo[3], wall_hi[3];
    float restitution, offset;
    float xf_scale;  // h / 2 (glm::scale(vec3(h / 2.f)))
    // dense grid
    double grid_o[3];
    double grid_inv;  // 1 / H
    int nx, ny, nz;
    // stencil reach R in cells: H >= h (1 + 2^-10) / R, so a neighbour within
    // h lies at most R cells away along each axis (1 <= R <= kMaxReach)
    int reach;
    uint32_t cell_base;  // first linear cell of this scene
    // slots [p_begin, p_end) of the particle arrays belong to this scene
    uint32_t p_begin, p_end;
    // 1: every wave walks its candidates per lane (never the union walk)
    uint32_t lane_walk;
};

SPH_HD inline uint32_t hash16(uint32_t cx, uint32_t cy, uint32_t cz)
{
    return ((cx * 73856093u) ^ (cy * 19349663u) ^ (cz * 83492791u)) & 0xFFFFu;
}

// float -> int toward zero; NaN and out-of-range give INT_MIN, which is what
// the reference's x86 build produces for ivec3(pos / h).
SPH_HD inline int trunc_to_int(float q)
{
    if (q > -2147483648.0f && q < 2147483648.0f)
        return (int)q;
    return (int)0x80000000u;
}

// True when q = x / h lies within eta(q) = 2^-21 + 2^-23 (|q| + 1) of a
// reference cell boundary (a nonzero integer), or |q| >= kQMax, or q is NaN.
// Both distances below are exact in float (Sterbenz) for |q| >= 1.
SPH_HD inline bool near_cell_boundary(float q)
{
    float a = q < 0.0f ? -q : q;
    if (!(a < kQMax))
        return true;
    float k = (float)(int)a;  // exact for a < 2^13
    float eta = 4.76837158203125e-07f + 1.1920928955078125e-07f * (a + 1.0f);  // 2^-21, 2^-23
    return (k >= 1.0f && a - k < eta) || (k + 1.0f - a < eta);
}

// True when two of the 27 buckets around reference cell c coincide.
SPH_HD inline bool buckets_collide(int cx, int cy, int cz)
{
    const uint32_t A = 73856093u, B = 19349663u, C = 83492791u;
    uint32_t x0 = (uint32_t)cx - 1u, x1 = (uint32_t)cx, x2 = (uint32_t)cx + 1u;
    uint32_t y0 = (uint32_t)cy - 1u, y1 = (uint32_t)cy, y2 = (uint32_t)cy + 1u;
    uint32_t z0 = (uint32_t)cz - 1u, z1 = (uint32_t)cz, z2 = (uint32_t)cz + 1u;
    uint32_t v[6] = {((x0 * A) ^ (x1 * A)) & 0xFFFFu, ((x1 * A) ^ (x2 * A)) & 0xFFFFu,
                     ((y0 * B) ^ (y1 * B)) & 0xFFFFu, ((y1 * B) ^ (y2 * B)) & 0xFFFFu,
                     ((z0 * C) ^ (z1 * C)) & 0xFFFFu, ((z1 * C) ^ (z2 * C)) & 0xFFFFu};
    bool dependent = false;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        uint32_t p = v[a];
        dependent = dependent || (p == 0u);
        uint32_t low = p & (0u - p);
#pragma unroll
        for (int b = a + 1; b < 6; ++b)
            v[b] = (v[b] & low) ? (v[b] ^ p) : v[b];
    }
    return dependent;
}

// Grid coordinate along one axis: floor((x - o) / H) clamped to [0, dim).
// Evaluated in double so the completeness bound above holds; NaN maps to 0.
SPH_HD inline int grid_coord(float x, double o, double inv, int dim)
{
    double q = ((double)x - o) * inv;
    if (!(q >= 0.0))
        return 0;
    if (!(q < (double)dim))
        return dim - 1;
    return (int)q;
}

SPH_HD inline uint32_t pack_cell(int gx, int gy, int gz, bool slow)
{
    return (uint32_t)gx | ((uint32_t)gy << 11) | ((uint32_t)gz << 21) | (slow ? kSlowBit : 0u);
}

SPH_HD inline int cell_x(uint32_t c) { return (int)(c & 2047u); }
SPH_HD inline int cell_y(uint32_t c) { return (int)((c >> 11) & 1023u); }
SPH_HD inline int cell_z(uint32_t c) { return (int)((c >> 21) & 1023u); }

SPH_HD inline uint32_t linear_cell(const StepParams &P, uint32_t c)
{
    return P.cell_base + (uint32_t)cell_x(c) +
           (uint32_t)P.nx * ((uint32_t)cell_y(c) + (uint32_t)P.ny * (uint32_t)cell_z(c));
}

struct Classified {
    uint32_t h16;    // H16 of the reference cell (the Particle::hash value)
    uint32_t cellc;  // packed grid cell + slow flag
};

// Reference cell, its 16-bit bucket, the slow flag and the grid cell of a
// position.  x / h is the IEEE quotient the reference's getCell forms.
SPH_HD inline Classified classify(float x, float y, float z, const StepParams &P)
{
    float qx = x / P.h, qy = y / P.h, qz = z / P.h;
    int cx = trunc_to_int(qx), cy = trunc_to_int(qy), cz = trunc_to_int(qz);
    Classified r;
    r.h16 = hash16((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
    bool slow = near_cell_boundary(qx) || near_cell_boundary(qy) ||
                near_cell_boundary(qz) || buckets_collide(cx, cy, cz);
    int gx = grid_coord(x, P.grid_o[0], P.grid_inv, P.nx);
    int gy = grid_coord(y, P.grid_o[1], P.grid_inv, P.ny);
    int gz = grid_coord(z, P.grid_o[2], P.grid_inv, P.nz);
    r.cellc = pack_cell(gx, gy, gz, slow);
    return r;
}

// m_ij counted from the 27 buckets around reference cell (cx, cy, cz).
SPH_HD inline int bucket_multiplicity(int cx, int cy, int cz, uint32_t h16j)
{
    const uint32_t A = 73856093u, B = 19349663u, C = 83492791u;
    uint32_t ax[3], by[3], cz3[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ax[k] = ((uint32_t)cx + (uint32_t)(k - 1)) * A;
        by[k] = ((uint32_t)cy + (uint32_t)(k - 1)) * B;
        cz3[k] = ((uint32_t)cz + (uint32_t)(k - 1)) * C;
    }
    int m = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            uint32_t ab = ax[a] ^ by[b];
#pragma unroll
            for (int c = 0; c < 3; ++c)
                m += (((ab ^ cz3[c]) & 0xFFFFu) == h16j) ? 1 : 0;
        }
    return m;
}

}  // namespace sph
