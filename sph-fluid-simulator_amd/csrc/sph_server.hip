// sph_server.hip -- continuous-batching serving loop (see sph_server.h).
#include "sph_server.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>

namespace sph {

namespace {

// Device bytes per particle of a Workspace: posA, velA, posB, velB, posR,
// velR (16 B each), cellcA, rankA, cellcB, perm, rhoB, the part plan's
// own-stencil length and the neighbour-list count (4 B each), and the
// neighbour list (kListMax x 4 B).
constexpr uint64_t kBytesPerParticle = 6 * 16 + 7 * 4 + 4 * kListMax;
// Per 64-slot wave task, the part plan: part count, offset, walk length and
// done counter (4 B each) and the walk (32 B); per part task (one per wave
// task plus the scene's part room) its entry (4 B) and partial-sum records
// (64 lanes x 16 B).
constexpr uint64_t kBytesPerWave = 16 + 32;
constexpr uint64_t kBytesPerPartTask = 4 + 64 * 16;
// Per grid cell: counts and cstart.
constexpr uint64_t kBytesPerCell = 8;

double wall_seconds()
{
    using clock = std::chrono::steady_clock;
    return std::chrono::duration<double>(clock::now().time_since_epoch()).count();
}

double percentile(std::vector<float> v, double q)
{
    if (v.empty())
        return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t k = (size_t)pos;
    const double f = pos - (double)k;
    return k + 1 < v.size() ? v[k] * (1.0 - f) + v[k + 1] * f : v[k];
}

}  // namespace

Server::~Server()
{
    if (batch.stream)
        (void)hipStreamSynchronize(batch.stream);
}

double Server::now() const { return wall_seconds() - t0; }

uint64_t Server::bytes_for(const BatchScene &s)
{
    const uint64_t cells = (uint64_t)s.P.nx * s.P.ny * s.P.nz;
    const uint64_t waves = ((uint64_t)s.desc.n + 63u) / 64u;
    return kBytesPerParticle * s.desc.n + kBytesPerWave * waves + kBytesPerPartTask * (waves + part_room(waves)) +
           kBytesPerCell * cells + 4096;
}

int Server::init(const sph_server_config_v1 *c)
{
    int rc;
    if (c)
        cfg = *c;
    if (cfg.flags & ~(SPH_SERVER_NO_GRAPH | SPH_SERVER_NO_BACKFILL))
        return fail(SPH_E_INVALID, "server: unknown flags 0x%x", cfg.flags);
    if ((rc = batch.init_empty(false)))
        return rc;
    if (cfg.memory_budget == 0) {
        size_t free_b = 0, total_b = 0;
        hipError_t e = hipMemGetInfo(&free_b, &total_b);
        if (e != hipSuccess)
            return hip_fail(e, "hipMemGetInfo");
        cfg.memory_budget = free_b / 4 * 3;
    }
    if (cfg.max_scenes == 0)
        cfg.max_scenes = 1024;
    if (cfg.max_steps_per_tick == 0)
        cfg.max_steps_per_tick = 8;
    cfg.max_steps_per_tick = std::min<uint32_t>(cfg.max_steps_per_tick, 256u);
    batch.use_graph = (cfg.flags & SPH_SERVER_NO_GRAPH) == 0;
    retired.state = SPH_REQ_DONE;
    batch.max_capacity =
        std::min<uint64_t>(cfg.memory_budget / (kBytesPerParticle + (kBytesPerWave + kBytesPerPartTask) / 64u),
                           kMaxParticles);
    t0 = wall_seconds();
    return SPH_OK;
}

int Server::submit(const sph_scene_desc_v1 &d, const float *pos, const float *vel, uint32_t steps, uint64_t *id)
{
    int rc;
    if (!pos || !id)
        return fail(SPH_E_INVALID, "submit: NULL positions or id");
    if (steps == 0)
        return fail(SPH_E_INVALID, "submit: steps must be > 0");
    if (cfg.max_queued && queue.size() >= cfg.max_queued) {
        ++rejected;
        return fail(SPH_E_STATE, "submit: %zu requests queued (the limit)", queue.size());
    }
    if ((d.flags & SPH_SCENE_KEEP_FORCE) != 0u)
        return fail(SPH_E_INVALID, "submit: the serving loop keeps no forces");
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r)
        return fail(SPH_E_NOMEM, "submit: host allocation failed");
    if ((rc = make_scene(d, r->scene)))
        return rc;
    const size_t n3 = 3ull * d.n;
    try {
        r->pos.assign(pos, pos + n3);
        if (vel)
            r->vel.assign(vel, vel + n3);
        else
            r->vel.assign(n3, 0.0f);
    } catch (...) {
        return fail(SPH_E_NOMEM, "submit: host allocation failed");
    }
    scene_grid(r->scene, r->pos.data(), d.n);
    r->bytes = bytes_for(r->scene);
    if (r->bytes > cfg.memory_budget)
        return fail(SPH_E_LIMIT, "submit: scene needs %llu bytes, budget is %llu",
                    (unsigned long long)r->bytes, (unsigned long long)cfg.memory_budget);
    r->id = next_id++;
    r->steps = steps;
    r->state = SPH_REQ_QUEUED;
    r->t_submit = now();
    *id = r->id;
    queue.push_back(r.get());
    requests.emplace(r->id, std::move(r));
    return SPH_OK;
}

// The head of the queue does not fit.  Memory frees up as running requests
// finish: `shadow` is the number of batch steps until enough has freed for
// the head.  A later request is admitted now when it fits and will finish
// within those steps, so it never delays the head.  Backfilled requests
// keep their queue order among themselves.
void Server::backfill(const std::vector<Request *> &keep, std::vector<Request *> &join)
{
    const Request *head = queue.front();
    if (head->bytes > cfg.memory_budget)
        return;
    std::vector<std::pair<uint64_t, uint64_t>> ends;  // (steps left, bytes) of what runs
    for (const Request *r : keep)
        ends.emplace_back(r->steps - r->done, r->bytes);
    for (const Request *r : join)
        ends.emplace_back(r->steps, r->bytes);
    std::sort(ends.begin(), ends.end());
    uint64_t free_b = cfg.memory_budget - used_bytes, shadow = UINT64_MAX;
    for (const auto &e : ends) {
        free_b += e.second;
        if (free_b >= head->bytes) {
            shadow = e.first;
            break;
        }
    }
    if (shadow == UINT64_MAX || shadow == 0)
        return;
    for (auto it = queue.begin() + 1; it != queue.end() && keep.size() + join.size() < cfg.max_scenes;) {
        Request *r = *it;
        if (used_bytes + r->bytes <= cfg.memory_budget && r->steps <= shadow) {
            used_bytes += r->bytes;
            join.push_back(r);
            ++backfilled;
            it = queue.erase(it);
        } else {
            ++it;
        }
    }
}

// Puts the requests of a failed admission back into the queue, in submit
// order (backfilled ones behind the head they passed), and releases their
// charge.
void Server::requeue(const std::vector<Request *> &join)
{
    for (Request *r : join) {
        r->state = SPH_REQ_QUEUED;
        used_bytes -= r->bytes;
        queue.push_back(r);
    }
    std::sort(queue.begin(), queue.end(), [](const Request *a, const Request *b) { return a->id < b->id; });
}

int Server::tick(uint32_t *ran_steps)
{
    int rc;
    if (ran_steps)
        *ran_steps = 0;
    // Retired scenes leave and queued requests join in one relayout.  The
    // admission is committed only once the relayout and every upload
    // succeeded; otherwise the joining requests go back to the head of the
    // queue and the batch keeps the scenes that were running.
    std::vector<Request *> keep, join;
    for (Request *r : active)
        if (r->state == SPH_REQ_ACTIVE)
            keep.push_back(r);
    while (!queue.empty() && keep.size() + join.size() < cfg.max_scenes &&
           used_bytes + queue.front()->bytes <= cfg.memory_budget) {
        Request *r = queue.front();
        queue.pop_front();
        used_bytes += r->bytes;
        join.push_back(r);
    }
    if (!queue.empty() && !(cfg.flags & SPH_SERVER_NO_BACKFILL) && keep.size() + join.size() < cfg.max_scenes)
        backfill(keep, join);
    if (keep.size() != active.size() || !join.empty()) {
        std::vector<BatchScene> next;
        std::vector<int> from;
        for (size_t k = 0; k < active.size(); ++k)
            if (active[k]->state == SPH_REQ_ACTIVE) {
                next.push_back(batch.scenes[k]);
                from.push_back((int)k);
            }
        for (Request *r : join) {
            next.push_back(r->scene);
            from.push_back(-1);
        }
        if ((rc = batch.relayout(std::move(next), from))) {
            requeue(join);  // the batch still holds the old scenes, as `active` says
            return rc;
        }
        active = keep;
        active.insert(active.end(), join.begin(), join.end());  // matches batch.scenes
        for (size_t k = 0; k < join.size(); ++k) {
            if ((rc = batch.upload((uint32_t)(keep.size() + k), join[k]->pos.data(), join[k]->vel.data()))) {
                // take the joined scenes out again: the kept ones move down
                // to their new slots (no growth, so only device copies)
                std::vector<BatchScene> back(batch.scenes.begin(), batch.scenes.begin() + (ptrdiff_t)keep.size());
                std::vector<int> idx(keep.size());
                for (size_t q = 0; q < keep.size(); ++q)
                    idx[q] = (int)q;
                if (batch.relayout(std::move(back), idx) == SPH_OK)
                    active = keep;
                else  // the joined scenes stay in the batch, marked retired
                    for (size_t q = keep.size(); q < active.size(); ++q)
                        active[q] = &retired;
                requeue(join);
                return rc;
            }
        }
        for (Request *r : join) {
            std::vector<float>().swap(r->pos);
            std::vector<float>().swap(r->vel);
            r->state = SPH_REQ_ACTIVE;
            r->t_admit = now();
        }
    }
    if (active.empty())
        return SPH_OK;

    uint32_t k = cfg.max_steps_per_tick;
    uint64_t n_active = 0;
    for (Request *r : active) {
        k = std::min(k, r->steps - r->done);
        n_active += r->scene.desc.n;
    }
    float ms[256];
    if ((rc = batch.step_timed((int)k, ms)))
        return rc;
    for (uint32_t s = 0; s < k; ++s) {
        step_ms.push_back(ms[s]);
        busy_ms += ms[s];
    }
    particle_steps += (double)k * (double)n_active;
    scene_steps += (double)k * (double)active.size();
    if (ran_steps)
        *ran_steps = k;

    for (size_t j = 0; j < active.size(); ++j) {
        Request *r = active[j];
        r->done += k;
        if (r->done < r->steps)
            continue;
        const size_t n = r->scene.desc.n;
        try {
            r->pos.resize(3 * n);
            r->vel.resize(3 * n);
            r->rho.resize(n);
            r->pressure.resize(n);
            r->ids.resize(n);
        } catch (...) {
            return fail(SPH_E_NOMEM, "server: host allocation failed");
        }
        if ((rc = batch.download((uint32_t)j, r->pos.data(), r->vel.data(), r->rho.data(), r->pressure.data(),
                                 r->ids.data())))
            return rc;
        r->state = SPH_REQ_DONE;
        r->t_done = now();
        used_bytes -= r->bytes;
        ++completed;
    }
    return SPH_OK;
}

int Server::fetch(uint64_t id, float *pos, float *vel, float *rho, float *pressure, uint32_t *ids)
{
    auto it = requests.find(id);
    if (it == requests.end())
        return fail(SPH_E_INVALID, "fetch: unknown request %llu", (unsigned long long)id);
    Request *r = it->second.get();
    if (r->state != SPH_REQ_DONE)
        return fail(SPH_E_STATE, "fetch: request %llu is not finished (or already fetched)",
                    (unsigned long long)id);
    const size_t n = r->scene.desc.n;
    if (pos)
        std::memcpy(pos, r->pos.data(), 12 * n);
    if (vel)
        std::memcpy(vel, r->vel.data(), 12 * n);
    if (rho)
        std::memcpy(rho, r->rho.data(), 4 * n);
    if (pressure)
        std::memcpy(pressure, r->pressure.data(), 4 * n);
    if (ids)
        std::memcpy(ids, r->ids.data(), 4 * n);
    std::vector<float>().swap(r->pos);
    std::vector<float>().swap(r->vel);
    std::vector<float>().swap(r->rho);
    std::vector<float>().swap(r->pressure);
    std::vector<uint32_t>().swap(r->ids);
    r->state = SPH_REQ_FETCHED;
    return SPH_OK;
}

int Server::stats(sph_server_stats_v1 *o) const
{
    std::memset(o, 0, sizeof *o);
    o->submitted = requests.size();
    o->completed = completed;
    o->queued = queue.size();
    for (Request *r : active)
        o->active += r->state == SPH_REQ_ACTIVE ? 1u : 0u;
    o->admitted = o->submitted - o->queued;
    o->steps = step_ms.size();
    o->used_bytes = used_bytes;
    o->budget_bytes = cfg.memory_budget;
    o->particle_steps = particle_steps;
    o->scene_steps = scene_steps;
    o->busy_ms = busy_ms;
    o->wall_s = now();
    o->step_ms_p50 = percentile(step_ms, 0.50);
    o->step_ms_p99 = percentile(step_ms, 0.99);
    o->step_ms_max = step_ms.empty() ? 0.0 : *std::max_element(step_ms.begin(), step_ms.end());
    o->rejected = rejected;
    o->backfilled = backfilled;
    return SPH_OK;
}

}  // namespace sph

using namespace sph;

struct sph_server {
    Server s;
};

extern "C" {

int sph_server_create(const sph_server_config_v1 *cfg, sph_server **out)
{
    if (!out)
        return fail(SPH_E_INVALID, "server create: NULL output");
    *out = nullptr;
    sph_server *o = new (std::nothrow) sph_server();
    if (!o)
        return fail(SPH_E_NOMEM, "server create: host allocation failed");
    int rc = o->s.init(cfg);
    if (rc) {
        delete o;
        return rc;
    }
    *out = o;
    return SPH_OK;
}

int sph_server_destroy(sph_server *s)
{
    delete s;
    return SPH_OK;
}

int sph_server_submit(sph_server *s, const sph_scene_desc_v1 *desc, const float *pos, const float *vel,
                      uint32_t steps, uint64_t *request_id)
{
    if (!s || !desc)
        return fail(SPH_E_INVALID, "submit: NULL argument");
    return s->s.submit(*desc, pos, vel, steps, request_id);
}

int sph_server_tick(sph_server *s, uint32_t *ran_steps)
{
    if (!s)
        return fail(SPH_E_INVALID, "tick: NULL server");
    return s->s.tick(ran_steps);
}

int sph_server_status(const sph_server *s, uint64_t request_id, int *state, uint32_t *done_steps, double times[3])
{
    if (!s || !state)
        return fail(SPH_E_INVALID, "status: NULL argument");
    auto it = s->s.requests.find(request_id);
    if (it == s->s.requests.end())
        return fail(SPH_E_INVALID, "status: unknown request %llu", (unsigned long long)request_id);
    const Request *r = it->second.get();
    *state = r->state;
    if (done_steps)
        *done_steps = r->done;
    if (times) {
        times[0] = r->t_submit;
        times[1] = r->t_admit;
        times[2] = r->t_done;
    }
    return SPH_OK;
}

int sph_server_fetch(sph_server *s, uint64_t request_id, float *pos, float *vel, float *rho, float *pressure,
                     uint32_t *ids)
{
    if (!s)
        return fail(SPH_E_INVALID, "fetch: NULL server");
    return s->s.fetch(request_id, pos, vel, rho, pressure, ids);
}

int sph_server_inject_failure(sph_server *s, int stage)
{
    if (!s || (stage != SPH_SERVER_FAIL_RELAYOUT && stage != SPH_SERVER_FAIL_UPLOAD))
        return fail(SPH_E_INVALID, "inject_failure: NULL server or unknown stage %d", stage);
    s->s.batch.inject_failure = stage == SPH_SERVER_FAIL_RELAYOUT ? Batch::kFailRelayout : Batch::kFailUpload;
    return SPH_OK;
}

int sph_server_stats(const sph_server *s, sph_server_stats_v1 *out)
{
    if (!s || !out)
        return fail(SPH_E_INVALID, "stats: NULL argument");
    return s->s.stats(out);
}

int sph_server_step_times(const sph_server *s, float *ms, size_t cap, size_t *n)
{
    if (!s || !n)
        return fail(SPH_E_INVALID, "step_times: NULL argument");
    const std::vector<float> &v = s->s.step_ms;
    const size_t k = ms ? std::min(cap, v.size()) : 0;
    if (k)
        std::memcpy(ms, v.data(), 4 * k);
    *n = v.size();
    return SPH_OK;
}

}  // extern "C"
