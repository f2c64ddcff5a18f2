// SPHSystem.h -- C++ facade with the reference's SPHSystem interface,
// running every step on the HIP device through the C ABI (sph_hip.h).
//
// Mirrors lijenicol/SPH-Fluid-Simulator:
//   struct SPHSettings(mass, restDensity, gasConst, viscosity, h, g, tension)
//                                                   src/SPHSystem.h:13-22
//   class SPHSystem(particleCubeWidth, settings, runOnGPU); update(dt);
//   reset(); startSimulation(); public particles / particleCount
//                                                   src/SPHSystem.h:24-58
// Headless: draw(viewProj, shader) is replaced by transforms(), the per-
// particle glm::mat4 buffer (column-major, 16 floats) that draw() copies
// into the instanced VBO (src/SPHSystem.cpp:119-133).
//
// After update() the device state is authoritative and the host copies
// (particles, transforms) are refreshed from it.  As in the reference
// (src/sph.cpp:184-192), update() leaves the particle array sorted by the
// 16-bit bucket `hash` (ties in the order of ids(), one valid outcome of the
// reference's unstable std::sort) with transforms[i] belonging to
// particles[i]; each particle's `acceleration` comes back as the caller left
// it (the step never reads or writes it).  With lazy host sync enabled,
// update() leaves the host copies stale until syncHost().
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sph_hip.h"

using Particle = sph_particle_v1;  // src/Particle.h:4-10, same 60-byte layout

struct SPHSettings : sph_settings_v1 {
    SPHSettings(float mass, float restDensity, float gasConst, float viscosity, float h, float g,
                float tension);
};

class SPHSystem {
public:
    // runOnGPU keeps the reference's meaning of which path's wall constants
    // apply: false = CPU path box (+-8, restitution 0.5, offset 1e-4),
    // true = CUDA path box (+-10, restitution 1, offset 1e-5).  Both run on
    // the HIP device; there is no CPU fallback.
    SPHSystem(size_t particleCubeWidth, const SPHSettings &settings, const bool &runOnGPU);
    ~SPHSystem();
    SPHSystem(const SPHSystem &) = delete;
    SPHSystem &operator=(const SPHSystem &) = delete;

    Particle *particles;
    size_t particleCount;

    // Advances one step of dt = 0.003 once startSimulation() was called
    // (src/SPHSystem.cpp:110-117: the argument is ignored).
    void update(float deltaTime);
    void reset();
    void startSimulation();

    // Instance transforms of the particles (16 floats each, column-major).
    const float *transforms() const { return sphereModelMtxs; }
    // ids()[i]: the index particles[i] had when the system was built (or
    // reset) -- an extension, the reference's Particle carries no id.
    const uint32_t *ids() const { return particleIds.data(); }

    // Extensions: lazy host copies, explicit sync, an arbitrary wall box.
    void setLazyHostSync(bool lazy) { lazyHostSync = lazy; }
    void syncHost();
    void setBox(const sph_box_v1 &box);
    bool started() const { return isStarted; }

private:
    SPHSettings settings;
    size_t particleCubeWidth;
    bool isStarted;
    bool runOnGPU;
    bool lazyHostSync = false;
    sph_box_v1 box;
    float *sphereModelMtxs;
    sph_ctx *ctx = nullptr;
    std::vector<uint32_t> particleIds;     // id of particles[i]
    std::vector<float> accelerationById;   // acceleration the caller left, per id (3 floats)

    void initParticles();
    void createContext();
};

// Thrown by the facade when the HIP library reports an error.
struct SPHError {
    int code;
    std::string message;
};
