// sph_server.h -- continuous-batching serving loop over a resident batch.
//
// Requests ("advance this scene by S steps") are queued on submit and
// admitted first-come first-served while the device memory they need fits
// the budget and a scene slot is free.  All admitted scenes live in one
// Batch and advance together: every tick runs up to max_steps_per_tick
// batch steps (one graph launch each), then retires finished requests
// (their state is copied to host memory until fetched) and admits queued
// ones, re-laying out the batch when its membership changed.  A scene's
// arithmetic depends only on its own slots and cells, so each result is
// bit-identical to the same scene stepped alone.
#pragma once

#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "sph_batch.h"

namespace sph {

struct Request {
    uint64_t id = 0;
    BatchScene scene;            // desc + constants + grid of the initial state
    std::vector<float> pos, vel; // initial state (freed on admission), then the result
    std::vector<float> rho, pressure;
    std::vector<uint32_t> ids;
    uint32_t steps = 0, done = 0;
    int state = SPH_REQ_QUEUED;
    uint64_t bytes = 0;  // device memory charged against the budget
    double t_submit = 0, t_admit = 0, t_done = 0;  // seconds since server start
};

struct Server {
    sph_server_config_v1 cfg{};
    Batch batch;
    std::map<uint64_t, std::unique_ptr<Request>> requests;
    std::deque<Request *> queue;
    std::vector<Request *> active;  // active[k] is batch.scenes[k]
    Request retired;  // stands for a batch scene no request owns (state DONE)
    uint64_t next_id = 1, used_bytes = 0;
    std::vector<float> step_ms;  // GPU time of every batch step
    double t0 = 0;
    double particle_steps = 0, scene_steps = 0, busy_ms = 0;
    uint64_t completed = 0, rejected = 0, backfilled = 0;

    ~Server();
    int init(const sph_server_config_v1 *c);
    double now() const;
    static uint64_t bytes_for(const BatchScene &s);
    int submit(const sph_scene_desc_v1 &d, const float *pos, const float *vel, uint32_t steps, uint64_t *id);
    int tick(uint32_t *ran_steps);
    void requeue(const std::vector<Request *> &join);
    void backfill(const std::vector<Request *> &keep, std::vector<Request *> &join);
    int fetch(uint64_t id, float *pos, float *vel, float *rho, float *pressure, uint32_t *ids);
    int stats(sph_server_stats_v1 *out) const;
};

}  // namespace sph
