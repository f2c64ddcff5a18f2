// sph_headless.cpp -- headless step loop over the SPHSystem facade: the
// reference application's simulation (src/Tester.cpp:90-91,110-125) without
// window, GUI or renderer.
//
//   sph_headless [--width W] [--steps S] [--gpu-box] [--lazy]
//                [--reset-after R] [--dump FILE]
//
// Builds SPHSystem(W, SPHSettings(0.02, 1000, 1, 1.04, 0.15, -9.8, 0.2),
// runOnGPU), starts it, calls update() S times and prints the mean step
// time.  --reset-after R first runs R updates, calls reset() and update()
// once more (the start gate is closed again: a no-op) before starting
// over.  --dump writes the particle array, the transform buffer and the
// ids: uint64 n, n * 60-byte Particle records, n * 16 floats, n * uint32.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "SPHSystem.h"

int main(int argc, char **argv)
{
    size_t width = 15;
    int steps = 100;
    bool gpuBox = false, lazy = false;
    int resetAfter = -1;
    std::string dump;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--width" && i + 1 < argc)
            width = std::strtoul(argv[++i], nullptr, 10);
        else if (a == "--steps" && i + 1 < argc)
            steps = std::atoi(argv[++i]);
        else if (a == "--gpu-box")
            gpuBox = true;
        else if (a == "--lazy")
            lazy = true;
        else if (a == "--dump" && i + 1 < argc)
            dump = argv[++i];
        else if (a == "--reset-after" && i + 1 < argc)
            resetAfter = std::atoi(argv[++i]);
        else {
            std::fprintf(stderr, "usage: %s [--width W] [--steps S] [--gpu-box] [--lazy] [--reset-after R] "
                                 "[--dump FILE]\n", argv[0]);
            return 2;
        }
    }
    try {
        SPHSettings settings(0.02f, 1000, 1, 1.04f, 0.15f, -9.8f, 0.2f);
        SPHSystem sys(width, settings, gpuBox);
        sys.setLazyHostSync(lazy);
        sys.update(0.016f);  // not started: a no-op, as in the reference
        if (resetAfter >= 0) {
            sys.startSimulation();
            for (int s = 0; s < resetAfter; ++s)
                sys.update(0.016f);
            sys.reset();
            sys.update(0.016f);  // the gate is closed again: a no-op
        }
        sys.startSimulation();
        auto t0 = std::chrono::steady_clock::now();
        for (int s = 0; s < steps; ++s)
            sys.update(0.016f);
        if (lazy)
            sys.syncHost();
        auto t1 = std::chrono::steady_clock::now();
        double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        std::printf("particles %zu steps %d mean %.4f ms/update (host sync %s)\n", sys.particleCount, steps,
                    steps ? ms / steps : 0.0, lazy ? "lazy" : "every update");
        if (!dump.empty()) {
            FILE *f = std::fopen(dump.c_str(), "wb");
            if (!f) {
                std::perror("dump");
                return 1;
            }
            uint64_t n = sys.particleCount;
            std::fwrite(&n, sizeof n, 1, f);
            std::fwrite(sys.particles, sizeof(Particle), n, f);
            std::fwrite(sys.transforms(), 16 * sizeof(float), n, f);
            std::fwrite(sys.ids(), sizeof(uint32_t), n, f);
            std::fclose(f);
        }
    } catch (const SPHError &e) {
        std::fprintf(stderr, "sph_headless: error %d: %s\n", e.code, e.message.c_str());
        return 1;
    }
    return 0;
}
