// C ABI of rtcore: contexts, scenes, build, trace, frames.
// Replaces the DeviceCore role of the reference (Device / Buffer / Launcher /
// Timer, DeviceCore/Device.hpp:68-87, Buffer.hpp:49-77, Kernel.hpp:47-75,
// Timer.hpp:27-66) and the resource plumbing of RayTracerCS::Init
// (RayTracerCS.cpp:313-514) with one HIP stream per context.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "rt_internal.h"

using rt::Scene;

namespace {

int fail(rt_ctx *ctx, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

int hip_fail(rt_ctx *ctx, hipError_t e, const char *what)
{
    return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_NOMEM : RT_ERR_HIP, "%s: %s", what,
                hipGetErrorString(e));
}

#define RT_HIP(ctx, call)                                        \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #call); \
    } while (0)

Scene *get_scene(rt_ctx *ctx, int scene)
{
    if (!ctx || scene < 0 || scene >= (int)ctx->scenes.size()) return nullptr;
    return ctx->scenes[scene];
}

template <typename T>
hipError_t scene_alloc(Scene &s, T **p, size_t count)
{
    const size_t bytes = sizeof(T) * (count ? count : 1);
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, bytes);
    if (e != hipSuccess) return e;
    s.allocs.push_back(v);
    s.bytes += bytes;
    *p = static_cast<T *>(v);
    return hipSuccess;
}

void scene_free(Scene *s)
{
    if (!s) return;
    for (void *p : s->allocs) hipFree(p);
    delete s;
}

int ensure_staging(rt_ctx *ctx, uint64_t bytes)
{
    if (ctx->staging_bytes >= bytes) return RT_OK;
    if (ctx->staging) {
        hipStreamSynchronize(ctx->stream);
        hipFree(ctx->staging);
        ctx->bytes -= ctx->staging_bytes;
        ctx->staging = nullptr;
        ctx->staging_bytes = 0;
    }
    RT_HIP(ctx, hipMalloc(&ctx->staging, bytes));
    ctx->staging_bytes = bytes;
    ctx->bytes += bytes;
    return RT_OK;
}

int ensure_frame(rt_ctx *ctx, uint32_t w, uint32_t h, uint32_t nl)
{
    rt::FrameBuffers &f = ctx->frame;
    const uint64_t px = (uint64_t)w * h;
    if (f.rgba && f.width == w && f.height == h && f.nlights >= nl) return RT_OK;
    hipStreamSynchronize(ctx->stream);
    hipFree(f.rgba);
    hipFree(f.hits);
    hipFree(f.occ);
    ctx->bytes -= f.bytes;
    f = rt::FrameBuffers();
    const uint32_t nlc = nl > 0 ? nl : 1;
    RT_HIP(ctx, hipMalloc(&f.rgba, sizeof(float4) * (px ? px : 1)));
    RT_HIP(ctx, hipMalloc(&f.hits, sizeof(rt_hit24) * (px ? px : 1)));
    RT_HIP(ctx, hipMalloc(&f.occ, (px ? px : 1) * nlc));
    f.width = w;
    f.height = h;
    f.nlights = nlc;
    f.bytes = (sizeof(float4) + sizeof(rt_hit24) + nlc) * (px ? px : 1);
    ctx->bytes += f.bytes;
    return RT_OK;
}

__global__ void k_max_index(const uint32_t *idx, uint64_t n, uint32_t *out)
{
    uint32_t m = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        m = idx[i] > m ? idx[i] : m;
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t y = __shfl_xor(m, o);
        m = y > m ? y : m;
    }
    if (__lane_id() == 0) atomicMax(out, m);
}

// Returns n events from the profiling pool and records which call they time.
hipEvent_t *prof_events(rt_ctx *ctx, int kind, size_t n)
{
    while (ctx->pool.size() < ctx->pool_used + n) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->pool.push_back(e);
    }
    rt_ctx::ProfRec r{kind, ctx->pool_used};
    ctx->recs.push_back(r);
    ctx->pool_used += n;
    return &ctx->pool[r.first];
}

}  // namespace

extern "C" {

int rt_create(rt_ctx **out, int device)
{
    if (!out) return RT_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RT_ERR_NODEV;
    if (device < 0 || device >= ndev) return RT_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return RT_ERR_HIP;
    rt_ctx *ctx = new rt_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return RT_ERR_HIP;
    }
    for (auto &e : ctx->ev) hipEventCreate(&e);
    if (hipMalloc(&ctx->stats, 8 * sizeof(unsigned long long)) != hipSuccess) {
        rt_destroy(ctx);
        return RT_ERR_NOMEM;
    }
    ctx->bytes += 8 * sizeof(unsigned long long);
    *out = ctx;
    return RT_OK;
}

void rt_destroy(rt_ctx *ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (Scene *s : ctx->scenes) scene_free(s);
    hipFree(ctx->frame.rgba);
    hipFree(ctx->frame.hits);
    hipFree(ctx->frame.occ);
    hipFree(ctx->stats);
    hipFree(ctx->staging);
    for (auto &e : ctx->ev)
        if (e) hipEventDestroy(e);
    for (auto &e : ctx->pool) hipEventDestroy(e);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *rt_last_error(const rt_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rt_sync(rt_ctx *ctx)
{
    if (!ctx) return RT_ERR_ARG;
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RT_OK;
}

int rt_get_stream(rt_ctx *ctx, void **stream)
{
    if (!ctx || !stream) return RT_ERR_ARG;
    *stream = (void *)ctx->stream;
    return RT_OK;
}

int rt_memory_in_use(rt_ctx *ctx, uint64_t *bytes)
{
    if (!ctx || !bytes) return RT_ERR_ARG;
    uint64_t b = ctx->bytes;
    for (Scene *s : ctx->scenes)
        if (s) b += s->bytes;
    *bytes = b;
    return RT_OK;
}

int rt_scene_create(rt_ctx *ctx, uint32_t max_tris, uint32_t max_verts, uint32_t depth, int *scene)
{
    if (!ctx || !scene) return fail(ctx, RT_ERR_ARG, "rt_scene_create: null argument");
    if (depth < 2 || depth > 26)
        return fail(ctx, RT_ERR_ARG, "rt_scene_create: depth %u outside [2, 26]", depth);
    if (max_tris > (1u << 25))
        return fail(ctx, RT_ERR_CAPACITY, "rt_scene_create: max_tris %u above 2^25", max_tris);
    hipSetDevice(ctx->device);
    Scene *s = new Scene();
    s->max_tris = max_tris;
    s->max_verts = max_verts;
    s->depth = depth;
    const uint32_t nt = max_tris ? max_tris : 1;
    const uint32_t tiles = rt::sort_tiles(nt);
    const uint32_t nb = rt::bounds_blocks(nt);
    hipError_t e = hipSuccess;
    uint32_t *words = nullptr;
    const size_t nwords = rt::kSortPasses * rt::kRadix + rt::kCounters +
                          (size_t)rt::kSortPasses * tiles * rt::kRadix;
#define ALLOC(p, n) \
    if (e == hipSuccess) e = scene_alloc(*s, &(p), (n))
    ALLOC(s->xyz, 3 * (size_t)(max_verts ? max_verts : 1));
    ALLOC(s->idx, 3 * (size_t)nt);
    ALLOC(s->pbox, nt);
    ALLOC(s->cent, 3 * (size_t)nt);
    ALLOC(s->partials, 6 * (size_t)nb);
    ALLOC(s->keys[0], nt);
    ALLOC(s->keys[1], nt);
    ALLOC(s->vals[0], nt);
    ALLOC(s->vals[1], nt);
    ALLOC(words, nwords);
    ALLOC(s->frontier[0], nt);
    ALLOC(s->frontier[1], nt);
    ALLOC(s->leaves, nt);
    ALLOC(s->heap, (size_t)1 << depth);
    ALLOC(s->trav, nt);
    ALLOC(s->trav_parent, nt);
    ALLOC(s->flags, nt);
    ALLOC(s->tris, nt);
    ALLOC(s->hdr, 1);
#undef ALLOC
    if (e != hipSuccess) {
        scene_free(s);
        return hip_fail(ctx, e, "rt_scene_create: hipMalloc");
    }
    s->hist = words;
    s->counters = words + rt::kSortPasses * rt::kRadix;
    s->status = s->counters + rt::kCounters;
    ctx->scenes.push_back(s);
    *scene = (int)ctx->scenes.size() - 1;
    return RT_OK;
}

int rt_scene_destroy(rt_ctx *ctx, int scene)
{
    Scene *s = get_scene(ctx, scene);
    if (!s) return fail(ctx, RT_ERR_ARG, "rt_scene_destroy: unknown scene %d", scene);
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    scene_free(s);
    ctx->scenes[scene] = nullptr;
    return RT_OK;
}

static int check_upload(rt_ctx *ctx, Scene *s, int scene, uint32_t nv, uint32_t nt)
{
    if (!s) return fail(ctx, RT_ERR_ARG, "unknown scene %d", scene);
    if (nt > s->max_tris || nv > s->max_verts)
        return fail(ctx, RT_ERR_CAPACITY, "scene %d holds %u tris / %u verts, got %u / %u", scene,
                    s->max_tris, s->max_verts, nt, nv);
    return RT_OK;
}

int rt_scene_upload(rt_ctx *ctx, int scene, const float *xyz, uint32_t nv, const uint32_t *idx,
                    uint32_t nt)
{
    Scene *s = get_scene(ctx, scene);
    int rc = check_upload(ctx, s, scene, nv, nt);
    if (rc) return rc;
    if ((nv && !xyz) || (nt && !idx)) return fail(ctx, RT_ERR_ARG, "rt_scene_upload: null array");
    for (uint64_t i = 0; i < 3ull * nt; ++i)
        if (idx[i] >= nv)
            return fail(ctx, RT_ERR_ARG, "rt_scene_upload: index %u at %llu >= %u vertices", idx[i],
                        (unsigned long long)i, nv);
    hipSetDevice(ctx->device);
    if (nv) RT_HIP(ctx, hipMemcpyAsync(s->xyz, xyz, 12ull * nv, hipMemcpyHostToDevice, ctx->stream));
    if (nt) RT_HIP(ctx, hipMemcpyAsync(s->idx, idx, 12ull * nt, hipMemcpyHostToDevice, ctx->stream));
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->nv = nv;
    s->nt = nt;
    s->built = false;
    s->have_rest = false;
    return RT_OK;
}

int rt_scene_upload_device(rt_ctx *ctx, int scene, const float *d_xyz, uint32_t nv,
                           const uint32_t *d_idx, uint32_t nt)
{
    Scene *s = get_scene(ctx, scene);
    int rc = check_upload(ctx, s, scene, nv, nt);
    if (rc) return rc;
    if ((nv && !d_xyz) || (nt && !d_idx))
        return fail(ctx, RT_ERR_ARG, "rt_scene_upload_device: null array");
    hipSetDevice(ctx->device);
    if (nv)
        RT_HIP(ctx, hipMemcpyAsync(s->xyz, d_xyz, 12ull * nv, hipMemcpyDeviceToDevice, ctx->stream));
    if (nt) {
        RT_HIP(ctx, hipMemcpyAsync(s->idx, d_idx, 12ull * nt, hipMemcpyDeviceToDevice, ctx->stream));
        // indices are validated on the device before any kernel gathers through them
        RT_HIP(ctx, hipMemsetAsync(s->counters, 0, sizeof(uint32_t), ctx->stream));
        hipLaunchKernelGGL(k_max_index, dim3(256), dim3(256), 0, ctx->stream, s->idx, 3ull * nt,
                           s->counters);
        uint32_t mx = 0;
        RT_HIP(ctx, hipMemcpyAsync(&mx, s->counters, 4, hipMemcpyDeviceToHost, ctx->stream));
        RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (mx >= nv)
            return fail(ctx, RT_ERR_ARG, "rt_scene_upload_device: index %u >= %u vertices", mx, nv);
    }
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->nv = nv;
    s->nt = nt;
    s->built = false;
    s->have_rest = false;
    return RT_OK;
}

int rt_scene_transform(rt_ctx *ctx, int scene, const float m[12])
{
    Scene *s = get_scene(ctx, scene);
    if (!s || !m) return fail(ctx, RT_ERR_ARG, "rt_scene_transform: bad argument");
    hipSetDevice(ctx->device);
    if (!s->have_rest) {
        if (!s->xyz_rest) {
            hipError_t e = scene_alloc(*s, &s->xyz_rest, 3 * (size_t)(s->max_verts ? s->max_verts : 1));
            if (e != hipSuccess) return hip_fail(ctx, e, "rt_scene_transform: hipMalloc");
        }
        if (s->nv)
            RT_HIP(ctx, hipMemcpyAsync(s->xyz_rest, s->xyz, 12ull * s->nv, hipMemcpyDeviceToDevice,
                                       ctx->stream));
        s->have_rest = true;
    }
    rt::Xform x;
    memcpy(x.m, m, sizeof(x.m));
    RT_HIP(ctx, rt::launch_transform(*s, x, ctx->stream));
    s->built = false;
    return RT_OK;
}

int rt_scene_info(rt_ctx *ctx, int scene, uint32_t *nt, uint32_t *nv, uint32_t *depth)
{
    Scene *s = get_scene(ctx, scene);
    if (!s) return fail(ctx, RT_ERR_ARG, "rt_scene_info: unknown scene %d", scene);
    if (nt) *nt = s->nt;
    if (nv) *nv = s->nv;
    if (depth) *depth = s->depth;
    return RT_OK;
}

int rt_lbvh_build(rt_ctx *ctx, int scene, rt_build_times *times)
{
    Scene *s = get_scene(ctx, scene);
    if (!s) return fail(ctx, RT_ERR_ARG, "rt_lbvh_build: unknown scene %d", scene);
    hipSetDevice(ctx->device);
    hipEvent_t *ev = times ? ctx->ev : nullptr;
    if (!times && ctx->profiling) ev = prof_events(ctx, 0, 6);
    RT_HIP(ctx, rt::launch_build(*s, ctx->stream, ev));
    s->built = true;
    if (times) {
        RT_HIP(ctx, hipEventSynchronize(ctx->ev[5]));
        float ms[5];
        for (int i = 0; i < 5; ++i) hipEventElapsedTime(&ms[i], ctx->ev[i], ctx->ev[i + 1]);
        times->bounds_ms = ms[0];
        times->morton_ms = ms[1];
        times->sort_ms = ms[2];
        times->emit_ms = ms[3];
        times->refit_ms = ms[4];
        hipEventElapsedTime(&times->total_ms, ctx->ev[0], ctx->ev[5]);
    }
    return RT_OK;
}

int rt_lbvh_read(rt_ctx *ctx, int scene, rt_node32 *nodes, uint32_t *sorted_codes,
                 int32_t *sorted_ids, float scene_box[6])
{
    Scene *s = get_scene(ctx, scene);
    if (!s) return fail(ctx, RT_ERR_ARG, "rt_lbvh_read: unknown scene %d", scene);
    if (!s->built) return fail(ctx, RT_ERR_STATE, "rt_lbvh_read: scene %d not built", scene);
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    if (nodes)
        RT_HIP(ctx, hipMemcpyAsync(nodes, s->heap, sizeof(rt_node32) << s->depth,
                                   hipMemcpyDeviceToHost, st));
    if (sorted_codes && s->nt)
        RT_HIP(ctx, hipMemcpyAsync(sorted_codes, s->keys[0], 4ull * s->nt, hipMemcpyDeviceToHost, st));
    if (sorted_ids && s->nt)
        RT_HIP(ctx, hipMemcpyAsync(sorted_ids, s->vals[0], 4ull * s->nt, hipMemcpyDeviceToHost, st));
    rt::SceneHdr h;
    RT_HIP(ctx, hipMemcpyAsync(&h, s->hdr, sizeof(h), hipMemcpyDeviceToHost, st));
    RT_HIP(ctx, hipStreamSynchronize(st));
    if (scene_box) memcpy(scene_box, h.scene_box, sizeof(float) * 6);
    return RT_OK;
}

static int trace_common(rt_ctx *ctx, Scene *s, const rt_ray48 *d_rays, uint32_t n, rt_hit24 *d_hits,
                        uint32_t flags, rt_frame_stats *stats, bool sync_out)
{
    hipStream_t st = ctx->stream;
    const bool count = (flags & RT_FLAG_COUNT) != 0;
    const bool timing = (flags & RT_FLAG_TIMING) != 0 || stats != nullptr;
    if (stats) RT_HIP(ctx, hipMemsetAsync(ctx->stats, 0, 8 * sizeof(unsigned long long), st));
    hipEvent_t *ev = timing ? &ctx->ev[6] : nullptr;
    if (!stats && ctx->profiling) ev = prof_events(ctx, 1, 2);
    if (ev) hipEventRecord(ev[0], st);
    RT_HIP(ctx, rt::launch_trace(*s, d_rays, n, d_hits, flags, ctx->stats, st));
    if (ev) hipEventRecord(ev[1], st);
    if (ctx->profiling) ctx->prof_primary += n;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        unsigned long long c[8] = {};
        if (count)
            RT_HIP(ctx, hipMemcpyAsync(c, ctx->stats, sizeof(c), hipMemcpyDeviceToHost, st));
        RT_HIP(ctx, hipStreamSynchronize(st));
        hipEventElapsedTime(&stats->trace_ms, ctx->ev[6], ctx->ev[7]);
        stats->primary_rays = n;
        stats->box_tests = c[0];
        stats->tri_tests = c[1];
    } else if (sync_out) {
        RT_HIP(ctx, hipStreamSynchronize(st));
    }
    return RT_OK;
}

int rt_trace(rt_ctx *ctx, int scene, const rt_ray48 *rays, uint32_t n, rt_hit24 *hits,
             uint32_t flags, rt_frame_stats *stats)
{
    Scene *s = get_scene(ctx, scene);
    if (!s) return fail(ctx, RT_ERR_ARG, "rt_trace: unknown scene %d", scene);
    if (!s->built) return fail(ctx, RT_ERR_STATE, "rt_trace: scene %d not built", scene);
    if (n && (!rays || !hits)) return fail(ctx, RT_ERR_ARG, "rt_trace: null array");
    hipSetDevice(ctx->device);
    const uint64_t rb = (sizeof(rt_ray48) * (uint64_t)n + 255) & ~255ull;
    int rc = ensure_staging(ctx, rb + sizeof(rt_hit24) * (uint64_t)n + 256);
    if (rc) return rc;
    rt_ray48 *d_rays = (rt_ray48 *)ctx->staging;
    rt_hit24 *d_hits = (rt_hit24 *)((char *)ctx->staging + rb);
    if (n)
        RT_HIP(ctx, hipMemcpyAsync(d_rays, rays, sizeof(rt_ray48) * (uint64_t)n,
                                   hipMemcpyHostToDevice, ctx->stream));
    rc = trace_common(ctx, s, d_rays, n, d_hits, flags, stats, false);
    if (rc) return rc;
    if (n)
        RT_HIP(ctx, hipMemcpyAsync(hits, d_hits, sizeof(rt_hit24) * (uint64_t)n,
                                   hipMemcpyDeviceToHost, ctx->stream));
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RT_OK;
}

int rt_trace_device(rt_ctx *ctx, int scene, const rt_ray48 *d_rays, uint32_t n, rt_hit24 *d_hits,
                    uint32_t flags, rt_frame_stats *stats)
{
    Scene *s = get_scene(ctx, scene);
    if (!s) return fail(ctx, RT_ERR_ARG, "rt_trace_device: unknown scene %d", scene);
    if (!s->built) return fail(ctx, RT_ERR_STATE, "rt_trace_device: scene %d not built", scene);
    if (n && (!d_rays || !d_hits)) return fail(ctx, RT_ERR_ARG, "rt_trace_device: null array");
    if (((uintptr_t)d_rays & 15) || ((uintptr_t)d_hits & 7))
        return fail(ctx, RT_ERR_ARG, "rt_trace_device: rays need 16-byte, hits 8-byte alignment");
    hipSetDevice(ctx->device);
    return trace_common(ctx, s, d_rays, n, d_hits, flags, stats, false);
}

int rt_render_direct(rt_ctx *ctx, int scene, const rt_render_desc *desc, uint32_t flags,
                     rt_frame_stats *stats)
{
    Scene *s = get_scene(ctx, scene);
    if (!s || !desc) return fail(ctx, RT_ERR_ARG, "rt_render_direct: bad argument");
    if (!s->built) return fail(ctx, RT_ERR_STATE, "rt_render_direct: scene %d not built", scene);
    if (desc->nlights > 8 || (desc->nlights && !desc->lights))
        return fail(ctx, RT_ERR_ARG, "rt_render_direct: 0..8 lights with a light array");
    if (desc->width == 0 || desc->height == 0 || desc->width > 16384 || desc->height > 16384)
        return fail(ctx, RT_ERR_ARG, "rt_render_direct: image %ux%u", desc->width, desc->height);
    hipSetDevice(ctx->device);
    int rc = ensure_frame(ctx, desc->width, desc->height, desc->nlights);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    const bool timing = (flags & RT_FLAG_TIMING) != 0 || stats != nullptr;
    if (stats || !ctx->profiling)
        RT_HIP(ctx, hipMemsetAsync(ctx->stats, 0, 8 * sizeof(unsigned long long), st));
    hipEvent_t *ev = timing ? &ctx->ev[6] : nullptr;
    if (!stats && ctx->profiling) ev = prof_events(ctx, 1, 2);
    if (ev) hipEventRecord(ev[0], st);
    RT_HIP(ctx, rt::launch_render_direct(*s, *desc, ctx->frame, flags, ctx->stats, st));
    if (ev) hipEventRecord(ev[1], st);
    if (ctx->profiling) ctx->prof_primary += (uint64_t)desc->width * desc->height;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        unsigned long long c[8] = {};
        RT_HIP(ctx, hipMemcpyAsync(c, ctx->stats, sizeof(c), hipMemcpyDeviceToHost, st));
        RT_HIP(ctx, hipStreamSynchronize(st));
        hipEventElapsedTime(&stats->trace_ms, ctx->ev[6], ctx->ev[7]);
        stats->primary_rays = (uint64_t)desc->width * desc->height;
        stats->shadow_rays = c[2];
        stats->box_tests = c[0];
        stats->tri_tests = c[1];
    }
    ctx->frame.nlights = desc->nlights > 0 ? desc->nlights : 1;
    return RT_OK;
}

int rt_frame_read(rt_ctx *ctx, float *rgba, rt_hit24 *hits, uint8_t *occluded)
{
    if (!ctx) return RT_ERR_ARG;
    rt::FrameBuffers &f = ctx->frame;
    if (!f.rgba) return fail(ctx, RT_ERR_STATE, "rt_frame_read: no frame rendered");
    hipSetDevice(ctx->device);
    const uint64_t px = (uint64_t)f.width * f.height;
    hipStream_t st = ctx->stream;
    if (rgba) RT_HIP(ctx, hipMemcpyAsync(rgba, f.rgba, 16 * px, hipMemcpyDeviceToHost, st));
    if (hits) RT_HIP(ctx, hipMemcpyAsync(hits, f.hits, sizeof(rt_hit24) * px, hipMemcpyDeviceToHost, st));
    if (occluded) RT_HIP(ctx, hipMemcpyAsync(occluded, f.occ, px * f.nlights, hipMemcpyDeviceToHost, st));
    RT_HIP(ctx, hipStreamSynchronize(st));
    return RT_OK;
}

int rt_profile_begin(rt_ctx *ctx)
{
    if (!ctx) return RT_ERR_ARG;
    hipSetDevice(ctx->device);
    RT_HIP(ctx, hipMemsetAsync(ctx->stats, 0, 8 * sizeof(unsigned long long), ctx->stream));
    ctx->recs.clear();
    ctx->pool_used = 0;
    ctx->prof_primary = 0;
    ctx->profiling = true;
    return RT_OK;
}

int rt_profile_end(rt_ctx *ctx, rt_profile *out)
{
    if (!ctx || !out) return RT_ERR_ARG;
    if (!ctx->profiling) return fail(ctx, RT_ERR_STATE, "rt_profile_end without rt_profile_begin");
    hipSetDevice(ctx->device);
    ctx->profiling = false;
    unsigned long long c[8] = {};
    RT_HIP(ctx, hipMemcpyAsync(c, ctx->stats, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    memset(out, 0, sizeof(*out));
    for (const rt_ctx::ProfRec &r : ctx->recs) {
        hipEvent_t *e = &ctx->pool[r.first];
        if (r.kind == 0) {
            float ms;
            for (int i = 0; i < 5; ++i) {
                hipEventElapsedTime(&ms, e[i], e[i + 1]);
                out->build_ms[i] += ms;
            }
            hipEventElapsedTime(&ms, e[0], e[5]);
            out->build_ms[5] += ms;
            out->builds++;
        } else {
            float ms;
            hipEventElapsedTime(&ms, e[0], e[1]);
            out->trace_ms += ms;
            out->launches++;
        }
    }
    out->primary_rays = ctx->prof_primary;
    out->box_tests = c[0];
    out->tri_tests = c[1];
    out->shadow_rays = c[2];
    out->bounce_rays = c[3];
    ctx->recs.clear();
    ctx->pool_used = 0;
    return RT_OK;
}

void rt_camera_look_at(rt_camera *cam, const float pos[3], const float target[3], const float up[3],
                       float fov_y_deg, float aspect)
{
    double f[3], r[3], u[3];
    for (int k = 0; k < 3; ++k) f[k] = (double)target[k] - (double)pos[k];
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (int k = 0; k < 3; ++k) f[k] /= fl;
    // right = f x up, up' = right x f
    r[0] = f[1] * up[2] - f[2] * up[1];
    r[1] = f[2] * up[0] - f[0] * up[2];
    r[2] = f[0] * up[1] - f[1] * up[0];
    double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (int k = 0; k < 3; ++k) r[k] /= rl;
    u[0] = r[1] * f[2] - r[2] * f[1];
    u[1] = r[2] * f[0] - r[0] * f[2];
    u[2] = r[0] * f[1] - r[1] * f[0];
    const double ty = std::tan(0.5 * (double)fov_y_deg * 3.14159265358979323846 / 180.0);
    for (int k = 0; k < 3; ++k) {
        cam->pos[k] = pos[k];
        cam->fwd[k] = (float)f[k];
        cam->right[k] = (float)r[k];
        cam->up[k] = (float)u[k];
    }
    cam->tan_y = (float)ty;
    cam->tan_x = (float)(ty * (double)aspect);
}

}  // extern "C"
