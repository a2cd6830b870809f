// Render server: continuous batching of render requests over resident scenes.
//
// No reference counterpart (the reference renders one interactive view,
// RayTracerCS::Render, RayTracerCS.cpp:241-311); it keeps the reference's
// progressive accumulation (m_NumMuestras + g_uAccumulation,
// RayTracerCS.cpp:246-250, CSPrimaryRays.hlsl:62-63) per stream of requests:
// a request adds one pass of samples to the accumulation buffer of its
// stream (scene, view, size, mode), which stays resident, so the next request
// of the stream continues from it.
//
// Each step: admit arrived requests in arrival order while slots and the
// device-memory budget allow (building their scene if it is not resident and
// evicting the least recently used idle scenes and accumulators when needed;
// one request per stream in flight, since two passes into one accumulator in
// one launch would race), then render a share of every in-flight request's
// tiles in one persistent rt_render_jobs launch, and retire requests whose
// tiles are all done.  A pixel's samples depend only on (seed, frame, sample,
// pixel), so a request's image is the same alone or in any mix.  In dry-run
// mode no device work happens: a step's duration is its estimated rays /
// dry_rate, which lets the scheduling be tested without a GPU.  With a batch
// window, arrived requests wait (while nothing is in flight) up to the window
// for later arrivals to share their launch.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <list>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "rt_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct PoolScene {
    std::vector<float> xyz;
    std::vector<uint32_t> idx;
    uint32_t depth = 20;
    bool resident = false;
    int handle = -1;  // device scene handle while resident (not in dry run)
    uint64_t bytes = 0;
    uint64_t last_use = 0;
    int pins = 0;  // in-flight requests using the scene
};

using StreamKey = std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>;  // scene, view, w, h, mode

struct Accum {
    float *dev = nullptr;
    uint64_t bytes = 0;
    uint64_t last_use = 0;
    uint32_t passes = 0, samples = 0;
    bool busy = false;  // a request of this stream is in flight
};

struct Req {
    uint64_t id = 0;
    rt_request r{};
    std::vector<float> lights;
    StreamKey key;
    uint32_t tiles_total = 0, tiles_done = 0, frame = 0;
    double admitted = 0.0;
    uint64_t rays = 0;
};

}  // namespace

struct rt_server {
    rt_ctx *ctx = nullptr;
    rt_server_config cfg{};
    uint32_t tile_h = 8;  // tile rows; fixed for the server's life
    std::map<uint32_t, PoolScene> pool;
    std::map<StreamKey, Accum> accums;
    std::list<Req> pending;
    std::vector<Req> inflight;
    std::vector<rt_request_result> done;
    uint64_t next_id = 1, tick = 0, used = 0;
    rt_server_stats st{};
    std::string err;
    // launch shape (autotune): whole-image jobs of the first completed
    // requests, searched once there are RT_SERVER_TUNE_REQUESTS of them
    bool tuned = false;
    rt_tuning tune{};
    std::vector<rt_job_desc> tune_jobs;
    std::vector<uint32_t> tune_keys;              // pool key of each job's scene
    std::vector<std::vector<float>> tune_lights;  // each job's lights
};

namespace {

int fail(rt_server *s, int code, const std::string &msg)
{
    s->err = msg;
    return code;
}

uint32_t tiles_x(uint32_t w) { return (w + 15) / 16; }

uint32_t tiles_of(const rt_server *s, uint32_t w, uint32_t h)
{
    return tiles_x(w) * ((h + s->tile_h - 1) / s->tile_h);
}

// Pixels of tiles [b, b + c) of a w x h image (row-major tiles of 16 x tile_h).
uint64_t tile_pixels(const rt_server *s, uint32_t w, uint32_t h, uint32_t b, uint32_t c)
{
    const uint32_t tx = tiles_x(w);
    uint64_t px = 0;
    for (uint32_t t = b; t < b + c; ++t) {
        const uint32_t x0 = (t % tx) * 16, y0 = (t / tx) * s->tile_h;
        px += (uint64_t)std::min(16u, w - x0) * std::min(s->tile_h, h - y0);
    }
    return px;
}

void drop_scene(rt_server *s, PoolScene &p)
{
    if (p.handle >= 0 && s->ctx) rt_scene_destroy(s->ctx, p.handle);
    p.handle = -1;
    p.resident = false;
    s->used -= p.bytes;
    p.bytes = 0;
    s->st.evictions++;
}

void drop_accum(rt_server *s, std::map<StreamKey, Accum>::iterator it)
{
    if (it->second.dev && s->ctx) rt_device_free(s->ctx, it->second.dev);
    s->used -= it->second.bytes;
    s->accums.erase(it);
    s->st.evictions++;
}

// Evicts least recently used idle scenes / accumulators until `need` more
// bytes fit the budget; false when they cannot.  The scene and stream being
// admitted are never evicted.
bool make_room(rt_server *s, uint64_t need, const PoolScene *keep, const StreamKey &keep_key)
{
    while (s->used + need > s->cfg.hbm_budget) {
        uint64_t best = UINT64_MAX;
        PoolScene *ps = nullptr;
        auto ai = s->accums.end();
        for (auto &kv : s->pool) {
            PoolScene &p = kv.second;
            if (p.resident && p.pins == 0 && &p != keep && p.last_use < best) {
                best = p.last_use;
                ps = &p;
            }
        }
        for (auto it = s->accums.begin(); it != s->accums.end(); ++it)
            if (!it->second.busy && it->first != keep_key && it->second.last_use < best) {
                best = it->second.last_use;
                ai = it;
                ps = nullptr;
            }
        if (ai != s->accums.end()) drop_accum(s, ai);
        else if (ps) drop_scene(s, *ps);
        else return false;
    }
    return true;
}

// Makes request q's scene and accumulator resident.  1 = admitted, 0 = wait
// for memory, negative = rejected (RT_ERR_BUDGET: can never fit) or failed.
int admit(rt_server *s, Req &q)
{
    PoolScene &p = s->pool.at(q.r.scene_key);
    const uint32_t nt = (uint32_t)(p.idx.size() / 3), nv = (uint32_t)(p.xyz.size() / 3);
    const uint64_t sfull = rt::scene_bytes(nt, nv, p.depth);
    const uint64_t afull = 16ull * q.r.width * q.r.height;
    if (sfull + afull > s->cfg.hbm_budget) return RT_ERR_BUDGET;
    const uint64_t need = (p.resident ? 0 : sfull) + (s->accums.count(q.key) ? 0 : afull);
    if (!make_room(s, need, &p, q.key)) return 0;
    if (!p.resident) {
        if (!s->cfg.dry_run) {
            int h = -1;
            int rc = rt_scene_create(s->ctx, nt, nv, p.depth, &h);
            if (rc == RT_OK) rc = rt_scene_upload(s->ctx, h, p.xyz.data(), nv, p.idx.data(), nt);
            if (rc == RT_OK) rc = rt_lbvh_build(s->ctx, h, nullptr);
            if (rc != RT_OK) {
                s->err = rt_last_error(s->ctx);
                if (h >= 0) rt_scene_destroy(s->ctx, h);
                return rc;
            }
            p.handle = h;
        }
        p.resident = true;
        p.bytes = sfull;
        s->used += sfull;
        s->st.builds++;
    }
    auto ait = s->accums.find(q.key);
    if (ait == s->accums.end()) {
        Accum a;
        a.bytes = afull;
        if (!s->cfg.dry_run) {
            void *d = nullptr;
            int rc = rt_device_alloc(s->ctx, afull, &d);
            if (rc == RT_OK) rc = rt_device_zero(s->ctx, d, afull);
            if (rc != RT_OK) {
                s->err = rt_last_error(s->ctx);
                if (d) rt_device_free(s->ctx, d);
                return rc;
            }
            a.dev = static_cast<float *>(d);
        }
        s->used += afull;
        ait = s->accums.emplace(q.key, a).first;
    }
    Accum &a = ait->second;
    q.frame = a.passes++;
    a.busy = true;
    p.pins++;
    p.last_use = a.last_use = ++s->tick;
    return 1;
}

bool stream_busy(const rt_server *s, const StreamKey &k)
{
    auto it = s->accums.find(k);
    return it != s->accums.end() && it->second.busy;
}

}  // namespace

extern "C" {

int rt_server_create(rt_ctx *ctx, const rt_server_config *cfg, rt_server **out)
{
    if (!cfg || !out || (!ctx && !cfg->dry_run)) return RT_ERR_ARG;
    if (cfg->max_inflight == 0 || cfg->tiles_per_step == 0 || (cfg->dry_run && !(cfg->dry_rate > 0.0)) ||
        !(cfg->batch_window >= 0.0) ||
        (cfg->tile_order != RT_TILES_ROWS && cfg->tile_order != RT_TILES_CENTRE))
        return RT_ERR_ARG;
    rt_server *s = new rt_server();
    s->ctx = cfg->dry_run ? nullptr : ctx;
    s->cfg = *cfg;
    if (s->ctx) s->tile_h = (uint32_t)s->ctx->tune.block / 16;
    *out = s;
    return RT_OK;
}

void rt_server_destroy(rt_server *s)
{
    if (!s) return;
    if (s->ctx) {
        for (auto &kv : s->pool)
            if (kv.second.handle >= 0) rt_scene_destroy(s->ctx, kv.second.handle);
        for (auto &kv : s->accums)
            if (kv.second.dev) rt_device_free(s->ctx, kv.second.dev);
    }
    delete s;
}

const char *rt_server_last_error(const rt_server *s) { return s ? s->err.c_str() : "null server"; }

int rt_server_add_scene(rt_server *s, uint32_t key, const float *xyz, uint32_t nv, const uint32_t *idx,
                        uint32_t nt, uint32_t depth)
{
    if (!s) return RT_ERR_ARG;
    if (nv == 0 || nt == 0 || !xyz || !idx || depth < 2 || depth > 26)
        return fail(s, RT_ERR_ARG, "rt_server_add_scene: empty scene or bad depth");
    for (uint64_t k = 0; k < 3ull * nt; ++k)
        if (idx[k] >= nv) return fail(s, RT_ERR_ARG, "rt_server_add_scene: index out of range");
    if (s->pool.count(key)) return fail(s, RT_ERR_ARG, "rt_server_add_scene: key already registered");
    PoolScene p;
    p.xyz.assign(xyz, xyz + 3ull * nv);
    p.idx.assign(idx, idx + 3ull * nt);
    p.depth = depth;
    s->pool.emplace(key, std::move(p));
    return RT_OK;
}

int rt_server_submit(rt_server *s, const rt_request *r, uint64_t *id)
{
    if (!s || !r) return RT_ERR_ARG;
    if (!s->pool.count(r->scene_key)) return fail(s, RT_ERR_ARG, "rt_server_submit: unknown scene key");
    if (r->width == 0 || r->height == 0 || r->width > 16384 || r->height > 16384 || r->nlights > 8 ||
        (r->nlights && !r->lights) || (r->mode != RT_MODE_DIRECT && r->mode != RT_MODE_GI) ||
        r->spp == 0 || r->spp > 65536 || r->max_bounces > 64)
        return fail(s, RT_ERR_ARG, "rt_server_submit: bad request");
    Req q;
    q.id = s->next_id++;
    q.r = *r;
    if (r->nlights) q.lights.assign(r->lights, r->lights + 3 * r->nlights);
    q.r.lights = nullptr;
    q.key = StreamKey(r->scene_key, r->view, r->width, r->height, r->mode);
    q.tiles_total = tiles_of(s, r->width, r->height);
    if (id) *id = q.id;
    s->pending.push_back(std::move(q));
    s->st.submitted++;
    return RT_OK;
}

int rt_server_step(rt_server *s, double now, double *elapsed)
{
    if (!s) return RT_ERR_ARG;
    if (elapsed) *elapsed = 0.0;
    if (s->ctx && (uint32_t)s->ctx->tune.block / 16 != s->tile_h)
        return fail(s, RT_ERR_STATE, "rt_server_step: the context's block size changed since rt_server_create");
    const Clock::time_point t0 = Clock::now();
    s->st.wake = 0.0;
    if (s->cfg.batch_window > 0.0 && s->inflight.empty()) {
        // hold arrived requests up to batch_window so that later arrivals share
        // their launch; a full batch starts at once
        size_t arrived = 0;
        double oldest = 0.0;
        for (const Req &q : s->pending) {
            if (q.r.arrival > now) break;
            if (arrived++ == 0) oldest = q.r.arrival;
        }
        if (arrived > 0 && arrived < s->cfg.max_inflight && now < oldest + s->cfg.batch_window) {
            s->st.wake = oldest + s->cfg.batch_window;
            return RT_OK;
        }
    }
    // admission in arrival order; requests of a busy stream wait without
    // holding up the others, a request waiting for memory holds the queue
    for (auto it = s->pending.begin(); it != s->pending.end() && s->inflight.size() < s->cfg.max_inflight;) {
        if (it->r.arrival > now) break;
        if (stream_busy(s, it->key)) {
            ++it;
            continue;
        }
        const int rc = admit(s, *it);
        if (rc == 0) break;
        if (rc < 0) {
            rt_request_result res{};
            res.id = it->id;
            res.arrival = it->r.arrival;
            res.admitted = res.completed = now;
            res.status = rc;
            s->done.push_back(res);
            s->st.rejected++;
            it = s->pending.erase(it);
            continue;
        }
        it->admitted = now;
        s->inflight.push_back(std::move(*it));
        it = s->pending.erase(it);
    }
    if (s->inflight.empty()) return RT_OK;
    // tile shares: the step's budget split evenly over requests with tiles
    // left; what a nearly finished request does not use goes to the others
    const size_t n = s->inflight.size();
    std::vector<uint32_t> share(n, 0);
    uint32_t left = s->cfg.tiles_per_step;
    while (left > 0) {
        uint32_t open = 0;
        for (size_t k = 0; k < n; ++k) open += s->inflight[k].tiles_done + share[k] < s->inflight[k].tiles_total;
        if (open == 0) break;
        const uint32_t each = std::max(1u, left / open);
        for (size_t k = 0; k < n && left > 0; ++k) {
            const Req &q = s->inflight[k];
            const uint32_t give = std::min({q.tiles_total - q.tiles_done - share[k], each, left});
            share[k] += give;
            left -= give;
        }
    }
    std::vector<rt_job_desc> jobs;
    uint64_t est_rays = 0, tiles_step = 0;
    for (size_t k = 0; k < n; ++k) {
        if (share[k] == 0) continue;
        const Req &q = s->inflight[k];
        rt_job_desc j{};
        j.scene = s->pool.at(q.r.scene_key).handle;
        j.cam = q.r.cam;
        j.width = q.r.width;
        j.height = q.r.height;
        j.mode = q.r.mode;
        j.spp = q.r.spp;
        j.max_bounces = q.r.max_bounces;
        j.albedo = q.r.albedo;
        j.sky = q.r.sky;
        j.ambient = q.r.ambient;
        j.lights = q.lights.empty() ? nullptr : q.lights.data();
        j.nlights = (uint32_t)(q.lights.size() / 3);
        j.seed = q.r.seed;
        j.frame = q.frame;
        j.accum = s->accums.at(q.key).dev;
        j.tile_begin = q.tiles_done;
        j.tile_count = share[k];
        j.tile_order = s->cfg.tile_order;
        jobs.push_back(j);
        tiles_step += share[k];
        const uint64_t per_px = q.r.mode == RT_MODE_GI ? 1ull + q.r.max_bounces : 1ull + j.nlights;
        est_rays += tile_pixels(s, q.r.width, q.r.height, q.tiles_done, share[k]) * q.r.spp * per_px;
    }
    double dt;
    uint64_t rays;
    if (s->cfg.dry_run) {
        dt = (double)est_rays / s->cfg.dry_rate;
        rays = est_rays;
        s->st.device_s += dt;
    } else {
        rt_frame_stats fs{};
        const rt_tuning ctx_tune = s->ctx->tune;
        if (s->tuned) s->ctx->tune = s->tune;  // the server's own shape for its launches
        const int rc = rt_render_jobs(s->ctx, jobs.data(), (uint32_t)jobs.size(), RT_FLAG_PERSISTENT, &fs);
        s->ctx->tune = ctx_tune;
        if (rc != RT_OK) return fail(s, rc, rt_last_error(s->ctx));
        rays = fs.primary_rays + fs.shadow_rays + fs.bounce_rays;
        s->st.device_s += fs.trace_ms * 1e-3;
        dt = std::chrono::duration<double>(Clock::now() - t0).count();
    }
    s->st.rays += rays;
    s->st.steps++;
    const double t_done = now + dt;
    // rays are attributed by tile share; finished requests retire
    std::vector<Req> keep;
    keep.reserve(n);
    for (size_t k = 0; k < n; ++k) {
        Req &q = s->inflight[k];
        q.tiles_done += share[k];
        if (tiles_step) q.rays += rays * share[k] / tiles_step;
        if (q.tiles_done < q.tiles_total) {
            keep.push_back(std::move(q));
            continue;
        }
        PoolScene &p = s->pool.at(q.r.scene_key);
        Accum &a = s->accums.at(q.key);
        a.samples += q.r.spp;
        a.busy = false;
        p.pins--;
        p.last_use = a.last_use = ++s->tick;
        rt_request_result res{};
        res.id = q.id;
        res.arrival = q.r.arrival;
        res.admitted = q.admitted;
        res.completed = t_done;
        res.rays = q.rays;
        res.frame = q.frame;
        res.samples = a.samples;
        res.status = RT_OK;
        s->done.push_back(res);
        s->st.completed++;
        if (s->cfg.autotune && !s->tuned && !s->cfg.dry_run && s->tune_jobs.size() < RT_SERVER_TUNE_REQUESTS) {
            rt_job_desc j{};
            j.scene = p.handle;
            j.cam = q.r.cam;
            j.width = q.r.width;
            j.height = q.r.height;
            j.mode = q.r.mode;
            j.spp = q.r.spp;
            j.max_bounces = q.r.max_bounces;
            j.albedo = q.r.albedo;
            j.sky = q.r.sky;
            j.ambient = q.r.ambient;
            j.nlights = (uint32_t)(q.lights.size() / 3);
            j.seed = q.r.seed;
            j.frame = q.frame;
            j.accum = a.dev;  // the search renders into its own scratch accumulators
            j.tile_order = s->cfg.tile_order;
            s->tune_jobs.push_back(j);
            s->tune_keys.push_back(q.r.scene_key);
            s->tune_lights.push_back(q.lights);
        }
    }
    s->inflight.swap(keep);
    if (s->cfg.autotune && !s->tuned && s->tune_jobs.size() >= RT_SERVER_TUNE_REQUESTS) {
        // the collected requests whose scenes are still resident, whole images
        std::vector<rt_job_desc> jobs_t;
        for (size_t k = 0; k < s->tune_jobs.size(); ++k) {
            const PoolScene &p = s->pool.at(s->tune_keys[k]);
            if (!p.resident || p.handle != s->tune_jobs[k].scene) continue;
            rt_job_desc j = s->tune_jobs[k];
            j.lights = s->tune_lights[k].empty() ? nullptr : s->tune_lights[k].data();
            jobs_t.push_back(j);
        }
        const Clock::time_point t1 = Clock::now();
        rt_tuning best = s->ctx->tune;
        const int rc = jobs_t.empty() ? RT_OK
                                      : rt::search_jobs_shape_block(s->ctx, jobs_t.data(), (uint32_t)jobs_t.size(),
                                                                    s->ctx->tune.block, &best);
        if (rc != RT_OK) return fail(s, rc, rt_last_error(s->ctx));
        const double tune_s = std::chrono::duration<double>(Clock::now() - t1).count();
        s->st.tune_s += tune_s;
        dt += tune_s;  // the search holds the device: the clock moves on
        s->tune = best;
        s->tuned = !jobs_t.empty();
        s->tune_jobs.clear();
        s->tune_keys.clear();
        s->tune_lights.clear();
    }
    if (elapsed) *elapsed = dt;
    return RT_OK;
}

int rt_server_poll(rt_server *s, rt_request_result *out, uint32_t cap, uint32_t *n)
{
    if (!s || !n || (cap && !out)) return RT_ERR_ARG;
    const uint32_t k = (uint32_t)std::min<size_t>(cap, s->done.size());
    std::copy(s->done.begin(), s->done.begin() + k, out);
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *n = k;
    return RT_OK;
}

int rt_server_status(rt_server *s, rt_server_stats *out)
{
    if (!s || !out) return RT_ERR_ARG;
    *out = s->st;
    if (s->ctx) out->tuning = s->tuned ? s->tune : s->ctx->tune;
    out->bytes_used = s->used;
    out->pending = (uint32_t)s->pending.size();
    out->inflight = (uint32_t)s->inflight.size();
    uint32_t res = 0;
    for (auto &kv : s->pool) res += kv.second.resident;
    out->resident_scenes = res;
    out->accumulators = (uint32_t)s->accums.size();
    return RT_OK;
}

int rt_server_read_stream(rt_server *s, uint32_t scene_key, uint32_t view, uint32_t width, uint32_t height,
                          uint32_t mode, float *rgba_sum, uint32_t *samples)
{
    if (!s || !rgba_sum) return RT_ERR_ARG;
    auto it = s->accums.find(StreamKey(scene_key, view, width, height, mode));
    if (it == s->accums.end()) return fail(s, RT_ERR_STATE, "rt_server_read_stream: stream not resident");
    if (samples) *samples = it->second.samples;
    if (s->cfg.dry_run) {
        memset(rgba_sum, 0, it->second.bytes);
        return RT_OK;
    }
    return rt_device_read(s->ctx, rgba_sum, it->second.dev, it->second.bytes);
}

}  // extern "C"
