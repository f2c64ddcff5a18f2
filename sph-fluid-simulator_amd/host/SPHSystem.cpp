// SPHSystem.cpp -- the reference's SPHSystem facade over the HIP library.
// Construction, reset and the start gate follow src/SPHSystem.cpp:28-143
// of lijenicol/SPH-Fluid-Simulator; stepping runs on the device.
#include "SPHSystem.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

void check(int rc)
{
    if (rc != SPH_OK)
        throw SPHError{rc, sph_hip_last_error()};
}

// glm::translate(x) * glm::scale(vec3(s)) for finite x: diag(s, s, s, 1)
// with x + 0 in the last column (the + 0 of glm's column sum turns -0 into
// +0).  Matches the device kernel for finite positions.
void transform_of(const float x[3], float s, float *m)
{
    std::memset(m, 0, 16 * sizeof(float));
    m[0] = s;
    m[5] = s;
    m[10] = s;
    m[12] = x[0] + 0.0f;
    m[13] = x[1] + 0.0f;
    m[14] = x[2] + 0.0f;
    m[15] = 1.0f;
}

}  // namespace

SPHSettings::SPHSettings(float mass, float restDensity, float gasConst, float viscosity, float h,
                         float g, float tension)
{
    check(sph_settings_init(this, mass, restDensity, gasConst, viscosity, h, g, tension));
}

SPHSystem::SPHSystem(size_t particleCubeWidth_, const SPHSettings &settings_, const bool &runOnGPU_)
    : particles(nullptr), particleCount(0), settings(settings_), particleCubeWidth(particleCubeWidth_),
      isStarted(false), runOnGPU(runOnGPU_), sphereModelMtxs(nullptr)
{
    particleCount = particleCubeWidth * particleCubeWidth * particleCubeWidth;
    particles = static_cast<Particle *>(std::calloc(particleCount ? particleCount : 1, sizeof(Particle)));
    sphereModelMtxs = new float[16 * (particleCount ? particleCount : 1)];
    if (runOnGPU)
        check(sph_box_reference_gpu(&box));
    else
        check(sph_box_reference_cpu(&box));
    initParticles();
    createContext();
    isStarted = false;
}

SPHSystem::~SPHSystem()
{
    if (ctx)
        sph_hip_destroy(ctx);
    delete[] sphereModelMtxs;
    std::free(particles);
}

// src/SPHSystem.cpp:76-108: srand(1024); separation h + 0.01; jitter
// (rand / RAND_MAX * 0.5 - 1) * h / 10 per axis; origin (-1.5, h + 0.1, -1.5);
// index i + (j + w k) w.
void SPHSystem::initParticles()
{
    std::srand(1024);
    const float h = settings.h;
    const float sep = h + 0.01f;
    const size_t w = particleCubeWidth;
    for (size_t i = 0; i < w; i++) {
        for (size_t j = 0; j < w; j++) {
            for (size_t k = 0; k < w; k++) {
                float ranX = (float(std::rand()) / float((RAND_MAX)) * 0.5f - 1) * h / 10;
                float ranY = (float(std::rand()) / float((RAND_MAX)) * 0.5f - 1) * h / 10;
                float ranZ = (float(std::rand()) / float((RAND_MAX)) * 0.5f - 1) * h / 10;
                size_t idx = i + (j + w * k) * w;
                Particle &p = particles[idx];
                std::memset(&p, 0, sizeof p);
                p.position[0] = float(i) * sep + ranX - 1.5f;
                p.position[1] = float(j) * sep + ranY + h + 0.1f;
                p.position[2] = float(k) * sep + ranZ - 1.5f;
                transform_of(p.position, h / 2.f, sphereModelMtxs + 16 * idx);
            }
        }
    }
    particleIds.resize(particleCount);
    std::iota(particleIds.begin(), particleIds.end(), 0u);
}

void SPHSystem::createContext()
{
    if (ctx) {
        sph_hip_destroy(ctx);
        ctx = nullptr;
    }
    if (particleCount == 0)
        return;
    sph_scene_desc_v1 d;
    std::memset(&d, 0, sizeof d);
    d.n = (uint32_t)particleCount;
    d.dt = 0.003f;  // src/SPHSystem.cpp:113
    d.settings = settings;
    d.box = box;
    d.cell_scale = 1.0f;
    d.flags = SPH_SCENE_KEEP_FORCE;
    check(sph_hip_create(&d, &ctx));
    // the device numbers particles 0..n-1 in upload order: that number is
    // the id; the caller's accelerations are kept per id
    std::vector<float> pos(3 * particleCount), vel(3 * particleCount);
    std::vector<float> acc(3 * particleCount);
    for (size_t i = 0; i < particleCount; ++i)
        for (int a = 0; a < 3; ++a) {
            pos[3 * i + a] = particles[i].position[a];
            vel[3 * i + a] = particles[i].velocity[a];
            acc[3 * i + a] = particles[i].acceleration[a];
        }
    accelerationById.swap(acc);
    particleIds.resize(particleCount);
    std::iota(particleIds.begin(), particleIds.end(), 0u);
    check(sph_hip_upload(ctx, pos.data(), vel.data()));
}

void SPHSystem::update(float deltaTime)
{
    (void)deltaTime;  // a fixed step of 0.003 is used, as in the reference
    if (!isStarted || !ctx)
        return;
    check(sph_hip_step(ctx, 1));
    if (!lazyHostSync)
        syncHost();
}

// Device order -> the reference's order: sorted by 16-bit bucket, ties by
// id; acceleration restored from the caller's values.
void SPHSystem::syncHost()
{
    if (!ctx)
        return;
    const size_t n = particleCount;
    std::vector<Particle> dev(n);
    std::vector<float> xf(16 * n);
    std::vector<uint32_t> ids(n);
    check(sph_hip_download_particles(ctx, dev.data(), xf.data()));
    check(sph_hip_download(ctx, nullptr, nullptr, nullptr, nullptr, ids.data()));
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return dev[a].hash != dev[b].hash ? dev[a].hash < dev[b].hash : ids[a] < ids[b];
    });
    for (size_t k = 0; k < n; ++k) {
        const uint32_t s = order[k], id = ids[s];
        particles[k] = dev[s];
        for (int a = 0; a < 3; ++a)
            particles[k].acceleration[a] = accelerationById[3 * (size_t)id + a];
        std::memcpy(sphereModelMtxs + 16 * k, xf.data() + 16 * (size_t)s, 16 * sizeof(float));
        particleIds[k] = id;
    }
}

void SPHSystem::reset()
{
    initParticles();
    createContext();
    isStarted = false;
}

void SPHSystem::startSimulation() { isStarted = true; }

void SPHSystem::setBox(const sph_box_v1 &b)
{
    if (ctx && lazyHostSync)
        syncHost();
    box = b;
    createContext();
}
