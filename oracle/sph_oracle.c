/*
 * sph_oracle.c -- CPU restatement of the reference SPH step.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker).  See sph_oracle.h for the
 * list of reference functions restated here and for the deliberate
 * differences.  Build with -ffp-contract=off: the reference is compiled for
 * generic x86-64, where no fused multiply-add is ever formed, and the
 * restatement must round every product and sum separately as it does.
 *
 * Compiled twice: default gives orc32_* (float, reference bit semantics),
 * -DORC_F64 gives orc64_* (the same method in double precision).
 */
#include "sph_oracle.h"

#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#ifdef ORC_F64
typedef double real;
typedef orc64_particle particle;
typedef orc64_settings settings;
typedef orc64_box box_t;
#define FN(name) orc64_##name
#define PI_R ((double)3.14159265f)
#define SQRT_R(x) sqrt(x)
#else
typedef float real;
typedef orc32_particle particle;
typedef orc32_settings settings;
typedef orc32_box box_t;
#define FN(name) orc32_##name
#define PI_R 3.14159265f /* src/SPHSystem.cpp:6 */
/* The reference calls the unqualified double sqrt on a float
 * (src/sph.cpp:110); a correctly rounded double sqrt rounded to float is the
 * correctly rounded float sqrt, so sqrtf gives the same bits. */
#define SQRT_R(x) sqrtf(x)
#endif

int FN(sizeof_particle)(void) { return (int)sizeof(particle); }
int FN(sizeof_settings)(void) { return (int)sizeof(settings); }

/* SPHSettings::SPHSettings -- src/SPHSystem.cpp:8-26.
 * In the float flavour every expression keeps the reference's promotions:
 * pow(h, n) is the C double pow, so poly6, spikyGrad, spikyLap, selfDens are
 * evaluated in double and rounded to float on assignment. */
void FN(settings_init)(settings *s, float mass, float restDensity,
                       float gasConst, float viscosity, float h, float g,
                       float tension)
{
    memset(s, 0, sizeof(*s));
    s->mass = mass;
    s->restDensity = restDensity;
    s->gasConstant = gasConst;
    s->viscosity = viscosity;
    s->h = h;
    s->g = g;
    s->tension = tension;
#ifdef ORC_F64
    double hd = h;
    s->poly6 = 315.0 / (64.0 * PI_R * pow(hd, 9));
    s->spikyGrad = -45.0 / (PI_R * pow(hd, 6));
    s->spikyLap = 45.0 / (PI_R * pow(hd, 6));
    s->h2 = hd * hd;
    s->selfDens = (double)mass * s->poly6 * pow(hd, 6);
    s->massPoly6Product = (double)mass * s->poly6;
    real sc = hd / 2.0;
#else
    s->poly6 = 315.0f / (64.0f * PI_R * pow(h, 9));
    s->spikyGrad = -45.0f / (PI_R * pow(h, 6));
    s->spikyLap = 45.0f / (PI_R * pow(h, 6));
    s->h2 = h * h;
    s->selfDens = mass * s->poly6 * pow(h, 6);
    s->massPoly6Product = mass * s->poly6;
    real sc = h / 2.f;
#endif
    /* glm::scale(glm::vec3(h / 2.f)): column-major diag(sc, sc, sc, 1) */
    s->sphereScale[0] = sc;
    s->sphereScale[5] = sc;
    s->sphereScale[10] = sc;
    s->sphereScale[15] = 1;
}

/* The fixed wall box of the reference CPU path, src/sph.cpp:139-176:
 * floor plane y = h (lo.y = 0), x/z walls at +-(8 - h), no ceiling,
 * restitution 0.5, offset 1e-4. */
void FN(box_reference_cpu)(box_t *b)
{
    b->lo[0] = -8;
    b->lo[1] = 0;
    b->lo[2] = -8;
    b->hi[0] = 8;
    b->hi[1] = (real)INFINITY;
    b->hi[2] = 8;
    b->restitution = (real)0.5f;
#ifdef ORC_F64
    b->offset = 0.0001;
#else
    b->offset = 0.0001f;
#endif
}

/* getHash -- src/neighborTable.cpp:5-12.  The reference multiplies signed
 * ints (wrapping in practice); unsigned arithmetic gives the same bits. */
static inline uint32_t hash_u(uint32_t cx, uint32_t cy, uint32_t cz)
{
    return ((cx * 73856093u) ^ (cy * 19349663u) ^ (cz * 83492791u)) %
           ORC_TABLE_SIZE;
}

uint32_t FN(get_hash)(int cx, int cy, int cz)
{
    return hash_u((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
}

/* float -> int truncation as x86 cvttss2si/cvttsd2si performs it: toward
 * zero, with NaN and out-of-range values giving INT_MIN. */
static inline int trunc_to_int(real q)
{
    if (q > (real)-2147483649.0 && q < (real)2147483648.0)
        return (int)q;
    return INT_MIN;
}

/* getCell -- src/neighborTable.cpp:14-17: ivec3(pos / h), truncated. */
static inline void cell_of(const real pos[3], real h, int out[3])
{
    out[0] = trunc_to_int(pos[0] / h);
    out[1] = trunc_to_int(pos[1] / h);
    out[2] = trunc_to_int(pos[2] / h);
}

void FN(get_cell)(const real pos[3], real h, int out[3]) { cell_of(pos, h, out); }

/* 16-bit bucket of cell c + o: getHash() truncated to uint16_t as every
 * caller stores it (src/Particle.h:9, src/sph.cpp:43,93). */
static inline uint16_t bucket_of(const int c[3], int ox, int oy, int oz)
{
    return (uint16_t)hash_u((uint32_t)c[0] + (uint32_t)ox,
                            (uint32_t)c[1] + (uint32_t)oy,
                            (uint32_t)c[2] + (uint32_t)oz);
}

/* ---- thread-per-pass orchestration, src/sph.cpp:200-272 ---------------- */

typedef void (*range_fn)(void *ctx, size_t start, size_t end);

typedef struct {
    range_fn fn;
    void *ctx;
    size_t start, end;
} range_job;

static void *range_thread(void *arg)
{
    range_job *j = (range_job *)arg;
    j->fn(j->ctx, j->start, j->end);
    return NULL;
}

static int resolve_threads(int nthreads)
{
    if (nthreads > 0)
        return nthreads;
    long c = sysconf(_SC_NPROCESSORS_ONLN);
    return c > 0 ? (int)c : 1;
}

/* Static block partition: boundary i = i * (n / T), last = n
 * (src/sph.cpp:202-209); one thread per block, joined before returning. */
static void run_blocks(range_fn fn, void *ctx, size_t n, int nthreads)
{
    int t = resolve_threads(nthreads);
    if (t <= 1 || n == 0) {
        fn(ctx, 0, n);
        return;
    }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)t);
    range_job *jobs = (range_job *)malloc(sizeof(range_job) * (size_t)t);
    size_t block = n / (size_t)t;
    for (int i = 0; i < t; i++) {
        jobs[i].fn = fn;
        jobs[i].ctx = ctx;
        jobs[i].start = (size_t)i * block;
        jobs[i].end = (i == t - 1) ? n : (size_t)(i + 1) * block;
        pthread_create(&th[i], NULL, range_thread, &jobs[i]);
    }
    for (int i = 0; i < t; i++)
        pthread_join(th[i], NULL);
    free(jobs);
    free(th);
}

/* ---- parallelCalculateHashes, src/sph.cpp:17-24 ------------------------- */

typedef struct {
    particle *p;
    size_t n;
    const uint32_t *table;
    const settings *s;
    real dt;
    const box_t *box;
    real *xf;
} pass_ctx;

static void hashes_range(void *vctx, size_t start, size_t end)
{
    pass_ctx *c = (pass_ctx *)vctx;
    for (size_t i = start; i < end; i++) {
        int cell[3];
        cell_of(c->p[i].position, c->s->h, cell);
        c->p[i].hash = bucket_of(cell, 0, 0, 0);
    }
}

void FN(calculate_hashes)(particle *p, size_t n, const settings *s,
                          int nthreads)
{
    pass_ctx c = {p, n, NULL, s, 0, NULL, NULL};
    run_blocks(hashes_range, &c, n, nthreads);
}

/* ---- sortParticles, src/sph.cpp:184-192 ---------------------------------
 * Stable counting sort on the 16-bit hash (see header). */
void FN(sort_particles)(particle *p, size_t n)
{
    if (n == 0)
        return;
    size_t *start = (size_t *)calloc(65536 + 1, sizeof(size_t));
    particle *tmp = (particle *)malloc(sizeof(particle) * n);
    for (size_t i = 0; i < n; i++)
        start[(size_t)p[i].hash + 1]++;
    for (size_t b = 0; b < 65536; b++)
        start[b + 1] += start[b];
    for (size_t i = 0; i < n; i++)
        tmp[start[p[i].hash]++] = p[i];
    memcpy(p, tmp, sizeof(particle) * n);
    free(tmp);
    free(start);
}

/* ---- createNeighborTable, src/neighborTable.cpp:19-37 ------------------ */
uint32_t *FN(create_neighbor_table)(const particle *sorted, size_t n)
{
    uint32_t *table = (uint32_t *)malloc(sizeof(uint32_t) * ORC_TABLE_SIZE);
    for (size_t i = 0; i < ORC_TABLE_SIZE; ++i)
        table[i] = ORC_NO_PARTICLE;
    uint32_t prev = ORC_NO_PARTICLE;
    for (size_t i = 0; i < n; ++i) {
        uint32_t cur = sorted[i].hash;
        if (cur != prev) {
            table[cur] = (uint32_t)i;
            prev = cur;
        }
    }
    return table;
}

void FN(free_table)(uint32_t *table) { free(table); }

/* Squared distance exactly as glm::length2(pj - pi) forms it:
 * tmp = v * v; tmp.x + tmp.y + tmp.z, each operation rounded. */
static inline real dist2_of(const particle *pi, const particle *pj)
{
    real dx = pj->position[0] - pi->position[0];
    real dy = pj->position[1] - pi->position[1];
    real dz = pj->position[2] - pi->position[2];
    return dx * dx + dy * dy + dz * dz;
}

/* Walks the 27 hash buckets around particle i exactly like the reference
 * loops (src/sph.cpp:40-66 and :90-127) and calls visit() for every
 * candidate inside the support radius, once per bucket it is found in. */
#define FOR_EACH_NEIGHBOR(p, n, table, s, i, pi, pj, dist2, ...)              \
    do {                                                                       \
        int cell_[3];                                                          \
        cell_of((pi)->position, (s)->h, cell_);                                \
        for (int x_ = -1; x_ <= 1; x_++) {                                     \
            for (int y_ = -1; y_ <= 1; y_++) {                                 \
                for (int z_ = -1; z_ <= 1; z_++) {                             \
                    uint16_t cellHash_ = bucket_of(cell_, x_, y_, z_);         \
                    uint32_t j_ = (table)[cellHash_];                          \
                    if (j_ == ORC_NO_PARTICLE)                                 \
                        continue;                                              \
                    while (j_ < (n)) {                                         \
                        if (j_ == (i)) {                                       \
                            j_++;                                              \
                            continue;                                          \
                        }                                                      \
                        const particle *pj = &(p)[j_];                         \
                        if (pj->hash != cellHash_)                             \
                            break;                                             \
                        real dist2 = dist2_of((pi), pj);                       \
                        if (dist2 < (s)->h2) {                                 \
                            __VA_ARGS__                                        \
                        }                                                      \
                        j_++;                                                  \
                    }                                                          \
                }                                                              \
            }                                                                  \
        }                                                                      \
    } while (0)

/* ---- parallelDensityAndPressures, src/sph.cpp:28-76 ---------------------
 * Float flavour: pDensity (float) += float * pow(double, 3): the term is a
 * double and the sum is rounded back to float after every addition. */
static real density_of(const particle *p, size_t n, const uint32_t *table,
                       const settings *s, size_t i, uint32_t *count)
{
    real massPoly6Product = s->mass * s->poly6;
    real pDensity = 0;
    uint32_t cnt = 0;
    const particle *pi = &p[i];
    FOR_EACH_NEIGHBOR(p, n, table, s, i, pi, pj, dist2, {
        pDensity += massPoly6Product * pow(s->h2 - dist2, 3);
        cnt++;
    });
    if (count)
        *count = cnt;
    return pDensity + s->selfDens;
}

static void density_range(void *vctx, size_t start, size_t end)
{
    pass_ctx *c = (pass_ctx *)vctx;
    for (size_t i = start; i < end; i++) {
        particle *pi = &c->p[i];
        pi->density = density_of(c->p, c->n, c->table, c->s, i, NULL);
        pi->pressure = c->s->gasConstant * (pi->density - c->s->restDensity);
    }
}

void FN(density_pressure)(particle *p, size_t n, const uint32_t *table,
                          const settings *s, int nthreads)
{
    pass_ctx c = {p, n, table, s, 0, NULL, NULL};
    run_blocks(density_range, &c, n, nthreads);
}

/* ---- parallelForces, src/sph.cpp:80-129 ---------------------------------
 * Operation order follows the glm expressions term by term:
 *   dist = sqrt(dist2);  dir = v * (1 / sqrt(dot(v, v)))      (normalize)
 *   pressure = ((((-dir) * m) * (p_i + p_j)) / (2 rho_j)) * spikyGrad,
 *              then *= (float)pow(h - dist, 2)
 *   visco    = (((mu * m) * (dv / rho_j)) * spikyLap) * (h - dist)        */
static void force_of(const particle *p, size_t n, const uint32_t *table, const settings *s, size_t i,
                     real f[3])
{
    const particle *pi = &p[i];
    f[0] = f[1] = f[2] = 0;
    FOR_EACH_NEIGHBOR(p, n, table, s, i, pi, pj, dist2, {
        real dist = SQRT_R(dist2);
        real v[3] = {pj->position[0] - pi->position[0],
                     pj->position[1] - pi->position[1],
                     pj->position[2] - pi->position[2]};
        real dd = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        real inv = (real)1 / SQRT_R(dd);
        real psum = pi->pressure + pj->pressure;
        real den = 2 * pj->density;
        real hd = s->h - dist;
        real sq = (real)pow(hd, 2);
        real mum = s->viscosity * s->mass;
        for (int k = 0; k < 3; k++) {
            real dir = v[k] * inv;
            real pf = (((-dir) * s->mass) * psum) / den * s->spikyGrad;
            pf *= sq;
            f[k] += pf;
        }
        for (int k = 0; k < 3; k++) {
            real dv = pj->velocity[k] - pi->velocity[k];
            real vf = ((mum * (dv / pj->density)) * s->spikyLap) * hd;
            f[k] += vf;
        }
    });
}

static void forces_range(void *vctx, size_t start, size_t end)
{
    pass_ctx *c = (pass_ctx *)vctx;
    for (size_t i = start; i < end; i++) {
        real f[3];
        force_of(c->p, c->n, c->table, c->s, i, f);
        c->p[i].force[0] = f[0];
        c->p[i].force[1] = f[1];
        c->p[i].force[2] = f[2];
    }
}

void FN(forces)(particle *p, size_t n, const uint32_t *table,
                const settings *s, int nthreads)
{
    pass_ctx c = {p, n, table, s, 0, NULL, NULL};
    run_blocks(forces_range, &c, n, nthreads);
}

/* ---- parallelUpdateParticlePositions, src/sph.cpp:133-181 ---------------
 * Reflection planes are L = lo + h and U = hi - h; for the reference box
 * (orcXX_box_reference_cpu) these are h, h - 8 and 8 - h bit for bit, and
 * the reflected coordinate -x + 2L + offset / -x + 2U - offset is the
 * reference's expression.  The ceiling (never reached with hi.y = +inf) is
 * tested right after the floor. */
static void integrate_range(void *vctx, size_t start, size_t end)
{
    pass_ctx *c = (pass_ctx *)vctx;
    const settings *s = c->s;
    const box_t *b = c->box;
    real dt = c->dt;
    real sc = s->h / 2;
    real L[3], U[3];
    for (int k = 0; k < 3; k++) {
        L[k] = b->lo[k] + s->h;
        U[k] = b->hi[k] - s->h;
    }
    real e = b->restitution, off = b->offset;
    for (size_t i = start; i < end; i++) {
        particle *p = &c->p[i];
        real acc[3];
        acc[0] = p->force[0] / p->density + 0;
        acc[1] = p->force[1] / p->density + s->g;
        acc[2] = p->force[2] / p->density + 0;
        for (int k = 0; k < 3; k++)
            p->velocity[k] += acc[k] * dt;
        for (int k = 0; k < 3; k++)
            p->position[k] += p->velocity[k] * dt;

        static const int order[3] = {1, 0, 2}; /* y, x, z */
        for (int a = 0; a < 3; a++) {
            int k = order[a];
            if (p->position[k] < L[k]) {
                p->position[k] = -p->position[k] + 2 * L[k] + off;
                p->velocity[k] = -p->velocity[k] * e;
            }
            if (p->position[k] > U[k]) {
                p->position[k] = -p->position[k] + 2 * U[k] - off;
                p->velocity[k] = -p->velocity[k] * e;
            }
        }

        if (c->xf) {
            /* glm::translate(position) * sphereScale, evaluated as glm's
             * column operations: T[3] = I0*x + I1*y + I2*z + I3, then
             * M[c] = T0*S[c][0] + T1*S[c][1] + T2*S[c][2] + T3*S[c][3]. */
            real T[4][4], S[4][4];
            for (int cc = 0; cc < 4; cc++)
                for (int r = 0; r < 4; r++) {
                    T[cc][r] = (cc == r) ? 1 : 0;
                    S[cc][r] = (cc == r) ? (cc == 3 ? 1 : sc) : 0;
                }
            for (int r = 0; r < 4; r++)
                T[3][r] = ((T[0][r] * p->position[0] + T[1][r] * p->position[1]) +
                           T[2][r] * p->position[2]) + T[3][r];
            real *M = c->xf + i * 16;
            for (int cc = 0; cc < 4; cc++)
                for (int r = 0; r < 4; r++)
                    M[cc * 4 + r] = ((T[0][r] * S[cc][0] + T[1][r] * S[cc][1]) +
                                     T[2][r] * S[cc][2]) + T[3][r] * S[cc][3];
        }
    }
}

void FN(integrate)(particle *p, size_t n, real *transforms, const settings *s,
                   real dt, const box_t *box, int nthreads)
{
    pass_ctx c = {p, n, NULL, s, dt, box, transforms};
    run_blocks(integrate_range, &c, n, nthreads);
}

/* ---- updateParticlesCPU, src/sph.cpp:195-275 ---------------------------- */
void FN(update_particles)(particle *p, real *transforms, size_t n,
                          const settings *s, real dt, const box_t *box,
                          int nthreads)
{
    FN(calculate_hashes)(p, n, s, nthreads);
    FN(sort_particles)(p, n);
    uint32_t *table = FN(create_neighbor_table)(p, n);
    FN(density_pressure)(p, n, table, s, nthreads);
    FN(forces)(p, n, table, s, nthreads);
    FN(integrate)(p, n, transforms, s, dt, box, nthreads);
    FN(free_table)(table);
}

/* Neighbour multiset dump (CSR over sorted indices).  Entry j appears once
 * for every bucket the reference finds it in, i.e. m_ij times.  With idx ==
 * NULL (or cap too small) only offsets are filled.  Returns the total. */
size_t FN(neighbors)(const particle *sorted, size_t n, const uint32_t *table,
                     const settings *s, uint64_t *offsets, uint32_t *idx,
                     size_t cap)
{
    size_t total = 0;
    if (offsets)
        offsets[0] = 0;
    for (size_t i = 0; i < n; i++) {
        const particle *pi = &sorted[i];
        FOR_EACH_NEIGHBOR(sorted, n, table, s, i, pi, pj, dist2, {
            if (idx && total < cap)
                idx[total] = (uint32_t)(pj - sorted);
            total++;
        });
        if (offsets)
            offsets[i + 1] = total;
    }
    return total;
}

/* Neighbour rows (sorted indices, each repeated m_ij times) of selected
 * sorted indices only: CSR offsets[k + 1]; returns the total, writing at
 * most cap indices. */
size_t FN(neighbors_subset)(const particle *sorted, size_t n, const uint32_t *table,
                            const settings *s, const uint32_t *which, size_t k,
                            uint64_t *offsets, uint32_t *idx, size_t cap)
{
    size_t total = 0;
    if (offsets)
        offsets[0] = 0;
    for (size_t a = 0; a < k; a++) {
        const size_t i = which[a];
        const particle *pi = &sorted[i];
        FOR_EACH_NEIGHBOR(sorted, n, table, s, i, pi, pj, dist2, {
            if (idx && total < cap)
                idx[total] = (uint32_t)(pj - sorted);
            total++;
        });
        if (offsets)
            offsets[a + 1] = total;
    }
    return total;
}

/* Force of selected sorted indices (k x 3), from the densities and
 * pressures the particles hold. */
void FN(forces_subset)(const particle *sorted, size_t n, const uint32_t *table,
                       const settings *s, const uint32_t *which, size_t k,
                       real *force_out)
{
    for (size_t a = 0; a < k; a++)
        force_of(sorted, n, table, s, which[a], force_out + 3 * a);
}

/* Density (with self density) of selected sorted indices, plus the number
 * of neighbour terms summed for each. */
void FN(density_subset)(const particle *sorted, size_t n,
                        const uint32_t *table, const settings *s,
                        const uint32_t *which, size_t k, real *rho_out,
                        uint32_t *count_out)
{
    for (size_t a = 0; a < k; a++) {
        uint32_t cnt = 0;
        rho_out[a] = density_of(sorted, n, table, s, which[a], &cnt);
        if (count_out)
            count_out[a] = cnt;
    }
}
