/*
 * sph_oracle.h -- CPU restatement of the reference SPH step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * It restates, in plain C, the multithreaded CPU path of
 * lijenicol/SPH-Fluid-Simulator:
 *   src/SPHSystem.cpp:8-26      SPHSettings constructor (derived constants)
 *   src/neighborTable.cpp:5-37  getHash / getCell / createNeighborTable
 *   src/sph.cpp:17-24           parallelCalculateHashes
 *   src/sph.cpp:28-76           parallelDensityAndPressures
 *   src/sph.cpp:80-129          parallelForces
 *   src/sph.cpp:133-181         parallelUpdateParticlePositions
 *   src/sph.cpp:184-192         sortParticles
 *   src/sph.cpp:195-275         updateParticlesCPU (thread-per-pass orchestration)
 *
 * The library is compiled twice from one source: ORC_REAL=float gives the
 * orc32_* symbols (bit-level restatement of the reference arithmetic, same
 * 60-byte particle layout), ORC_REAL=double gives orc64_* (the same method
 * evaluated in FP64, used as the drift reference).
 *
 * Deliberate differences from the reference, none of which changes a result
 * the reference defines:
 *   - the wall box is a parameter (orc_box); orcXX_box_reference_cpu() gives
 *     the reference's fixed box (src/sph.cpp:139-176) bit for bit;
 *   - sortParticles uses a stable counting sort on the 16-bit hash instead of
 *     std::sort (unstable).  Any order sorted by hash is a valid std::sort
 *     outcome; stability makes the oracle deterministic;
 *   - the neighbour table is freed with free() (the reference mixes malloc
 *     with delete, src/sph.cpp:274).
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_TABLE_SIZE 262144u   /* src/neighborTable.h:9  */
#define ORC_NO_PARTICLE 0xFFFFFFFFu /* src/neighborTable.h:10 */

/* FP32 flavour: layout identical to the reference Particle (60 bytes). */
typedef struct {
    float position[3], velocity[3], acceleration[3], force[3];
    float density, pressure;
    uint16_t hash;
} orc32_particle;

typedef struct {
    float sphereScale[16];
    float poly6, spikyGrad, spikyLap, gasConstant, mass, h2, selfDens,
        restDensity, viscosity, h, g, tension, massPoly6Product;
} orc32_settings;

typedef struct {
    float lo[3], hi[3];
    float restitution, offset;
} orc32_box;

/* FP64 flavour. */
typedef struct {
    double position[3], velocity[3], acceleration[3], force[3];
    double density, pressure;
    uint16_t hash;
} orc64_particle;

typedef struct {
    double sphereScale[16];
    double poly6, spikyGrad, spikyLap, gasConstant, mass, h2, selfDens,
        restDensity, viscosity, h, g, tension, massPoly6Product;
} orc64_settings;

typedef struct {
    double lo[3], hi[3];
    double restitution, offset;
} orc64_box;

#define ORC_DECLARE(P, T)                                                      \
    int P##_sizeof_particle(void);                                             \
    int P##_sizeof_settings(void);                                             \
    void P##_settings_init(P##_settings *s, float mass, float restDensity,     \
                           float gasConst, float viscosity, float h, float g,  \
                           float tension);                                     \
    void P##_box_reference_cpu(P##_box *b);                                    \
    uint32_t P##_get_hash(int cx, int cy, int cz);                             \
    void P##_get_cell(const T pos[3], T h, int out[3]);                        \
    void P##_calculate_hashes(P##_particle *p, size_t n,                       \
                              const P##_settings *s, int nthreads);            \
    void P##_sort_particles(P##_particle *p, size_t n);                        \
    uint32_t *P##_create_neighbor_table(const P##_particle *sorted, size_t n); \
    void P##_free_table(uint32_t *table);                                      \
    void P##_density_pressure(P##_particle *p, size_t n,                       \
                              const uint32_t *table, const P##_settings *s,    \
                              int nthreads);                                   \
    void P##_forces(P##_particle *p, size_t n, const uint32_t *table,          \
                    const P##_settings *s, int nthreads);                      \
    void P##_integrate(P##_particle *p, size_t n, T *transforms,               \
                       const P##_settings *s, T dt, const P##_box *box,        \
                       int nthreads);                                          \
    void P##_update_particles(P##_particle *p, T *transforms, size_t n,        \
                              const P##_settings *s, T dt, const P##_box *box, \
                              int nthreads);                                   \
    size_t P##_neighbors(const P##_particle *sorted, size_t n,                 \
                         const uint32_t *table, const P##_settings *s,         \
                         uint64_t *offsets, uint32_t *idx, size_t cap);        \
    void P##_density_subset(const P##_particle *sorted, size_t n,              \
                            const uint32_t *table, const P##_settings *s,      \
                            const uint32_t *which, size_t k, T *rho_out,       \
                            uint32_t *count_out);                      \
    size_t P##_neighbors_subset(const P##_particle *sorted, size_t n,          \
                                const uint32_t *table, const P##_settings *s,  \
                                const uint32_t *which, size_t k,               \
                                uint64_t *offsets, uint32_t *idx, size_t cap); \
    void P##_forces_subset(const P##_particle *sorted, size_t n,               \
                           const uint32_t *table, const P##_settings *s,       \
                           const uint32_t *which, size_t k, T *force_out);

ORC_DECLARE(orc32, float)
ORC_DECLARE(orc64, double)

#ifdef __cplusplus
}
#endif

#endif /* SPH_ORACLE_H */
