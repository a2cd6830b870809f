// C ABI of rtcore: contexts, scenes, build, trace, frames.
// Replaces the DeviceCore role of the reference (Device / Buffer / Launcher /
// Timer, DeviceCore/Device.hpp:68-87, Buffer.hpp:49-77, Kernel.hpp:47-75,
// Timer.hpp:27-66) and the resource plumbing of RayTracerCS::Init
// (RayTracerCS.cpp:313-514) with one HIP stream per context.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "rt_internal.h"

using rt::Batch;
using rt::SceneDesc;

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

// Resolves a scene handle to its batch and index within the batch.
bool get_scene(rt_ctx *ctx, int scene, Batch **b, uint32_t *index)
{
    if (!ctx || scene < 0 || scene >= (int)ctx->scenes.size() || !ctx->scenes[scene].batch)
        return false;
    *b = ctx->scenes[scene].batch;
    *index = ctx->scenes[scene].index;
    return true;
}

template <typename T>
hipError_t batch_alloc(Batch &b, T **p, size_t count)
{
    const size_t bytes = sizeof(T) * (count ? count : 1);
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, bytes);
    if (e != hipSuccess) return e;
    b.allocs.push_back(v);
    b.bytes += bytes;
    *p = static_cast<T *>(v);
    return hipSuccess;
}

void batch_free(Batch *b)
{
    if (!b) return;
    for (void *p : b->allocs) hipFree(p);
    for (Batch::Shading &s : b->shading) hipFree(s.mats);
    for (Batch::Bvh &v : b->bvh) {
        hipFree(v.trav);
        hipFree(v.tris);
        hipFree(v.hdr);
    }
    delete b;
}

// Recomputes the compact (current-count) offsets of the build arrays and
// copies the scene descriptors to the device.
int batch_sync_desc(rt_ctx *ctx, Batch &b)
{
    uint32_t prim = 0, blocks = 0;
    for (SceneDesc &d : b.desc) {
        d.prim_off = prim;
        d.block_off = blocks;
        d.nblocks = rt::bounds_blocks(d.nt);
        prim += d.nt;
        blocks += d.nblocks;
    }
    b.total_tris = prim;
    b.total_blocks = blocks;
    RT_HIP(ctx, hipMemcpyAsync(b.d_desc, b.desc.data(), sizeof(SceneDesc) * b.nscenes,
                               hipMemcpyHostToDevice, ctx->stream));
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RT_OK;
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
    hipFree(f.order_mem);
    ctx->bytes -= f.bytes;
    f = rt::FrameBuffers();
    const uint32_t nlc = nl > 0 ? nl : 1;
    const uint64_t tiles = (uint64_t)((w + 15) / 16) * ((h + 3) / 4);
    RT_HIP(ctx, hipMalloc(&f.rgba, sizeof(float4) * (px ? px : 1)));
    RT_HIP(ctx, hipMalloc(&f.hits, sizeof(rt_hit24) * (px ? px : 1)));
    RT_HIP(ctx, hipMalloc(&f.occ, (px ? px : 1) * nlc));
    const uint64_t cap = tiles ? tiles : 1, order_words = 2 * (uint64_t)rt::FrameBuffers::kOrderSlots * cap;
    RT_HIP(ctx, hipMalloc(&f.order_mem, sizeof(uint32_t) * order_words));
    for (int k = 0; k < rt::FrameBuffers::kOrderSlots; ++k) {
        f.slots[k].order = f.order_mem + 2 * k * cap;
        f.slots[k].cost = f.order_mem + (2 * k + 1) * cap;
    }
    f.width = w;
    f.height = h;
    f.nlights = nlc;
    f.bytes = (sizeof(float4) + sizeof(rt_hit24) + nlc) * (px ? px : 1) + sizeof(uint32_t) * order_words;
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

namespace rt {

// Device bytes create_batch allocates for one scene (serving admission).
uint64_t scene_bytes(uint32_t max_tris, uint32_t max_verts, uint32_t depth)
{
    const uint64_t nt = max_tris ? max_tris : 1, nv = max_verts ? max_verts : 1;
    const uint64_t tiles = sort_tiles((uint32_t)nt);
    const uint64_t words = kCodePasses * kRadix + kCounters + (uint64_t)kCodePasses * tiles * kRadix;
    const uint64_t table = ((uint64_t)1 << table_bits(max_tris)) + 1;
    return 12 * nv + 12 * nt + sizeof(SceneDesc) + sizeof(SceneHdr) + sizeof(TriRec) * nt + 12 * nt +
           24 * (uint64_t)bounds_blocks((uint32_t)nt) + 16 * nt + 4 * words + 4 * table + 16 * 4 * nt +
           4 * kMaxLevels + sizeof(rt_node32) * ((uint64_t)1 << depth) + sizeof(TravNode) * nt +
           sizeof(TravTri) * nt;
}

}  // namespace rt

namespace {

int create_batch(rt_ctx *ctx, uint32_t n, const uint32_t *max_tris, const uint32_t *max_verts,
                 const uint32_t *depths, int *first)
{
    if (!ctx || !first || n == 0 || !max_tris || !max_verts || !depths)
        return fail(ctx, RT_ERR_ARG, "scene creation: null argument or no scenes");
    if (n > (uint32_t)rt::kMaxBatchScenes)
        return fail(ctx, RT_ERR_ARG, "a batch holds at most %d scenes, got %u", rt::kMaxBatchScenes, n);
    uint64_t cap_tris = 0, cap_verts = 0, heap = 0, table = 0, blocks = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (depths[i] < 2 || depths[i] > 26)
            return fail(ctx, RT_ERR_ARG, "scene %u: depth %u outside [2, 26]", i, depths[i]);
        cap_tris += max_tris[i];
        cap_verts += max_verts[i];
        heap += (uint64_t)1 << depths[i];
        table += ((uint64_t)1 << rt::table_bits(max_tris[i])) + 1;
        blocks += rt::bounds_blocks(max_tris[i]);
    }
    if (cap_tris > (1u << 25) || cap_verts > (3u << 25))
        return fail(ctx, RT_ERR_CAPACITY, "batch capacity %llu triangles / %llu vertices too large",
                    (unsigned long long)cap_tris, (unsigned long long)cap_verts);
    hipSetDevice(ctx->device);
    Batch *b = new Batch();
    b->nscenes = n;
    b->desc.resize(n);
    b->max_tris.assign(max_tris, max_tris + n);
    b->max_verts.assign(max_verts, max_verts + n);
    b->rest.assign(n, 0);
    uint32_t in_prim = 0, vert = 0;
    uint64_t hoff = 0, toff = 0;
    for (uint32_t i = 0; i < n; ++i) {
        SceneDesc &d = b->desc[i];
        memset(&d, 0, sizeof(d));
        d.in_prim_off = in_prim;
        d.vert_off = vert;
        d.depth = depths[i];
        d.tbits = rt::table_bits(max_tris[i]);
        d.heap_off = hoff;
        d.table_off = toff;
        in_prim += max_tris[i];
        vert += max_verts[i];
        hoff += (uint64_t)1 << depths[i];
        toff += ((uint64_t)1 << d.tbits) + 1;
        if (depths[i] > b->max_depth) b->max_depth = depths[i];
    }
    b->total_heap = heap;
    b->total_table = table;
    const uint32_t nt = cap_tris ? (uint32_t)cap_tris : 1;
    const uint32_t tiles = rt::sort_tiles(nt);
    const uint32_t passes = rt::kCodePasses + (n > 1 ? 1 : 0);
    const size_t nwords = rt::kCodePasses * rt::kRadix + rt::kCounters + (size_t)passes * tiles * rt::kRadix;
    hipError_t e = hipSuccess;
    uint32_t *words = nullptr;
#define ALLOC(p, cnt) \
    if (e == hipSuccess) e = batch_alloc(*b, &(p), (cnt))
    ALLOC(b->xyz, 3 * (size_t)(cap_verts ? cap_verts : 1));
    ALLOC(b->idx, 3 * (size_t)nt);
    ALLOC(b->d_desc, n);
    ALLOC(b->hdr, n);
    ALLOC(b->trec, nt);
    ALLOC(b->cent, 3 * (size_t)nt);
    ALLOC(b->partials, 6 * (size_t)(blocks ? blocks : 1));
    ALLOC(b->keys[0], nt);
    ALLOC(b->keys[1], nt);
    ALLOC(b->vals[0], nt);
    ALLOC(b->vals[1], nt);
    ALLOC(words, nwords);
    ALLOC(b->table, table);
    ALLOC(b->big[0], nt);
    ALLOC(b->big[1], nt);
    ALLOC(b->small, nt);
    ALLOC(b->top_nodes, nt);
    ALLOC(b->top_counts, (size_t)n * rt::kMaxLevels);
    ALLOC(b->heap, heap);
    ALLOC(b->trav, nt);
    ALLOC(b->tris, nt);
#undef ALLOC
    if (e != hipSuccess) {
        batch_free(b);
        return hip_fail(ctx, e, "scene creation: hipMalloc");
    }
    b->hist = words;
    b->counters = words + rt::kCodePasses * rt::kRadix;
    b->status = b->counters + rt::kCounters;
    int rc = batch_sync_desc(ctx, *b);
    if (rc) {
        batch_free(b);
        return rc;
    }
    *first = (int)ctx->scenes.size();
    for (uint32_t i = 0; i < n; ++i) ctx->scenes.push_back(rt::SceneRef{b, i});
    b->refs = (int)n;
    return RT_OK;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Scratch of rt_render_jobs_noalloc: one static device array per GPU holding
// the job table (RT_NOALLOC_MAX_JOBS records), the persistent launch's tile
// counter and the ray counters.  In host code the array's name is not a
// device address: rt_create looks the address up once per device
// (hipGetSymbolAddress, with the device current) and every host use goes
// through that pointer.  All calls on one GPU share the array.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocTableBytes = (size_t)RT_NOALLOC_MAX_JOBS * rt::kRenderJobBytesMax;
constexpr size_t kNoAllocCounterOff = kNoAllocTableBytes;       // 256 B: tile counter
constexpr size_t kNoAllocStatsOff = kNoAllocTableBytes + 256;   // 8 x u64 ray counters
constexpr size_t kNoAllocScratchBytes = kNoAllocTableBytes + 512;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<unsigned char *> g_scratch_ptr;  // per device; null until looked up

static int noalloc_setup(rt_ctx *ctx)
{
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if ((int)g_scratch_ptr.size() <= ctx->device) g_scratch_ptr.resize(ctx->device + 1, nullptr);
        if (!g_scratch_ptr[ctx->device]) {
            void *p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // device is current
            if (e != hipSuccess || !p) return hip_fail(ctx, e, "rt_create: hipGetSymbolAddress(scratch)");
            g_scratch_ptr[ctx->device] = static_cast<unsigned char *>(p);
        }
        ctx->noalloc_scratch = g_scratch_ptr[ctx->device];
    }
    ctx->noalloc_host.assign((size_t)RT_NOALLOC_MAX_JOBS * rt::render_job_bytes(), 0);
    ctx->noalloc_offs.assign(RT_NOALLOC_MAX_JOBS + 1, 0u);
    RT_HIP(ctx, hipEventCreateWithFlags(&ctx->noalloc_ev, hipEventDisableTiming));
    return RT_OK;
}

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
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cus > 0)
        ctx->cu_count = cus;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return RT_ERR_HIP;
    }
    for (auto &e : ctx->ev) hipEventCreate(&e);
    hipEventCreateWithFlags(&ctx->jobs_ev, hipEventDisableTiming);
    if (hipMalloc(&ctx->stats, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&ctx->tile_counter, 256) != hipSuccess || hipMalloc(&ctx->wcount, 128) != hipSuccess) {
        rt_destroy(ctx);
        return RT_ERR_NOMEM;
    }
    ctx->bytes += 8 * sizeof(unsigned long long) + 256 + 128;
    if (noalloc_setup(ctx) != RT_OK) {
        rt_destroy(ctx);
        return RT_ERR_HIP;
    }
    *out = ctx;
    return RT_OK;
}

void rt_destroy(rt_ctx *ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (rt::SceneRef &r : ctx->scenes) {
        if (r.batch && --r.batch->refs == 0) batch_free(r.batch);
        r.batch = nullptr;
    }
    hipFree(ctx->frame.rgba);
    hipFree(ctx->frame.hits);
    hipFree(ctx->frame.occ);
    hipFree(ctx->frame.order_mem);
    for (rt_ctx::Graph &g : ctx->graphs) {
        if (g.exec) hipGraphExecDestroy(g.exec);
        if (g.graph) hipGraphDestroy(g.graph);
        hipFree(g.jobs);
        hipFree(g.counter);
        if (g.batch && --g.batch->refs == 0) batch_free(g.batch);
    }
    hipFree(ctx->stats);
    hipFree(ctx->staging);
    hipFree(ctx->tile_counter);
    hipFree(ctx->wcount);
    hipFree(ctx->wqueue);
    hipFree(ctx->jobs_dev);
    if (ctx->jobs_ev) hipEventDestroy(ctx->jobs_ev);
    if (ctx->noalloc_ev) hipEventDestroy(ctx->noalloc_ev);
    for (auto &kv : ctx->tile_orders) hipFree(kv.second);
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
    std::vector<Batch *> seen;
    auto count = [&](Batch *x) {
        if (!x) return;
        for (Batch *y : seen)
            if (y == x) return;
        seen.push_back(x);
        b += x->bytes;
    };
    for (const rt::SceneRef &r : ctx->scenes) count(r.batch);
    for (const rt_ctx::Graph &g : ctx->graphs) count(g.batch);  // batches only a graph still holds
    *bytes = b;
    return RT_OK;
}

int rt_scene_create(rt_ctx *ctx, uint32_t max_tris, uint32_t max_verts, uint32_t depth, int *scene)
{
    return create_batch(ctx, 1, &max_tris, &max_verts, &depth, scene);
}

int rt_batch_create(rt_ctx *ctx, uint32_t nscenes, const uint32_t *max_tris,
                    const uint32_t *max_verts, const uint32_t *depths, int *first_scene)
{
    return create_batch(ctx, nscenes, max_tris, max_verts, depths, first_scene);
}

int rt_scene_destroy(rt_ctx *ctx, int scene)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i)) return fail(ctx, RT_ERR_ARG, "rt_scene_destroy: unknown scene %d", scene);
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    ctx->scenes[scene].batch = nullptr;
    b->generation++;
    if (--b->refs == 0) batch_free(b);
    return RT_OK;
}

static int check_upload(rt_ctx *ctx, Batch *b, uint32_t i, int scene, uint32_t nv, uint32_t nt)
{
    if (nt > b->max_tris[i] || nv > b->max_verts[i])
        return fail(ctx, RT_ERR_CAPACITY, "scene %d holds %u tris / %u verts, got %u / %u", scene,
                    b->max_tris[i], b->max_verts[i], nt, nv);
    return RT_OK;
}

// The BVH is static (built once from the vertices of its build): new
// vertices drop it and the scene goes back to the LBVH.
static void drop_bvh(Batch *b, uint32_t i)
{
    if ((i < b->use_bvh.size() && b->use_bvh[i]) || (i < b->bvh.size() && b->bvh[i].valid)) b->generation++;
    if (i < b->use_bvh.size()) b->use_bvh[i] = 0;
    if (i < b->bvh.size()) b->bvh[i].valid = false;
}

static int finish_upload(rt_ctx *ctx, Batch *b, uint32_t i, uint32_t nv, uint32_t nt)
{
    drop_bvh(b, i);
    b->desc[i].nv = nv;
    b->desc[i].nt = nt;
    b->built = false;
    b->rest[i] = 0;
    return batch_sync_desc(ctx, *b);
}

int rt_scene_upload(rt_ctx *ctx, int scene, const float *xyz, uint32_t nv, const uint32_t *idx,
                    uint32_t nt)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i)) return fail(ctx, RT_ERR_ARG, "rt_scene_upload: unknown scene %d", scene);
    int rc = check_upload(ctx, b, i, scene, nv, nt);
    if (rc) return rc;
    if ((nv && !xyz) || (nt && !idx)) return fail(ctx, RT_ERR_ARG, "rt_scene_upload: null array");
    for (uint64_t k = 0; k < 3ull * nt; ++k)
        if (idx[k] >= nv)
            return fail(ctx, RT_ERR_ARG, "rt_scene_upload: index %u at %llu >= %u vertices", idx[k],
                        (unsigned long long)k, nv);
    hipSetDevice(ctx->device);
    const SceneDesc &d = b->desc[i];
    if (nv)
        RT_HIP(ctx, hipMemcpyAsync(b->xyz + 3 * (size_t)d.vert_off, xyz, 12ull * nv,
                                   hipMemcpyHostToDevice, ctx->stream));
    if (nt)
        RT_HIP(ctx, hipMemcpyAsync(b->idx + 3 * (size_t)d.in_prim_off, idx, 12ull * nt,
                                   hipMemcpyHostToDevice, ctx->stream));
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return finish_upload(ctx, b, i, nv, nt);
}

int rt_scene_upload_device(rt_ctx *ctx, int scene, const float *d_xyz, uint32_t nv,
                           const uint32_t *d_idx, uint32_t nt)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i)) return fail(ctx, RT_ERR_ARG, "rt_scene_upload_device: unknown scene %d", scene);
    int rc = check_upload(ctx, b, i, scene, nv, nt);
    if (rc) return rc;
    if ((nv && !d_xyz) || (nt && !d_idx))
        return fail(ctx, RT_ERR_ARG, "rt_scene_upload_device: null array");
    hipSetDevice(ctx->device);
    const SceneDesc &d = b->desc[i];
    if (nv)
        RT_HIP(ctx, hipMemcpyAsync(b->xyz + 3 * (size_t)d.vert_off, d_xyz, 12ull * nv,
                                   hipMemcpyDeviceToDevice, ctx->stream));
    if (nt) {
        uint32_t *didx = b->idx + 3 * (size_t)d.in_prim_off;
        RT_HIP(ctx, hipMemcpyAsync(didx, d_idx, 12ull * nt, hipMemcpyDeviceToDevice, ctx->stream));
        // indices are validated on the device before any kernel gathers through them
        RT_HIP(ctx, hipMemsetAsync(b->counters, 0, sizeof(uint32_t), ctx->stream));
        hipLaunchKernelGGL(k_max_index, dim3(256), dim3(256), 0, ctx->stream, didx, 3ull * nt,
                           b->counters);
        uint32_t mx = 0;
        RT_HIP(ctx, hipMemcpyAsync(&mx, b->counters, 4, hipMemcpyDeviceToHost, ctx->stream));
        RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (mx >= nv) {
            b->desc[i].nt = 0;  // the rejected indices stay out of every build
            batch_sync_desc(ctx, *b);
            return fail(ctx, RT_ERR_ARG, "rt_scene_upload_device: index %u >= %u vertices", mx, nv);
        }
    }
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return finish_upload(ctx, b, i, nv, nt);
}

int rt_scene_shading(rt_ctx *ctx, int scene, const float *normals, uint32_t nv, const uint32_t *mat_ids,
                     uint32_t nt, const rt_material *materials, uint32_t nmat)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i)) return fail(ctx, RT_ERR_ARG, "rt_scene_shading: unknown scene %d", scene);
    const SceneDesc &d = b->desc[i];
    if (nv != d.nv || nt != d.nt)
        return fail(ctx, RT_ERR_ARG, "rt_scene_shading: scene %d has %u vertices / %u triangles, got %u / %u", scene,
                    d.nv, d.nt, nv, nt);
    if (materials ? (nmat == 0 || nmat > 65536) : nmat != 0)
        return fail(ctx, RT_ERR_ARG, "rt_scene_shading: 1..65536 materials with a table, or none (got %u)", nmat);
    const uint32_t ntab = materials ? nmat : 1u;
    if (mat_ids)
        for (uint32_t t = 0; t < nt; ++t)
            if (mat_ids[t] >= ntab)
                return fail(ctx, RT_ERR_ARG, "rt_scene_shading: triangle %u has material %u of %u", t, mat_ids[t],
                            ntab);
    hipSetDevice(ctx->device);
    if ((normals && !b->nrm) || (mat_ids && !b->mat)) {
        uint64_t cap_v = 0, cap_t = 0;
        for (uint32_t k = 0; k < b->nscenes; ++k) {
            cap_v += b->max_verts[k];
            cap_t += b->max_tris[k];
        }
        if (normals && !b->nrm) RT_HIP(ctx, batch_alloc(*b, &b->nrm, 3 * (size_t)(cap_v ? cap_v : 1)));
        if (mat_ids && !b->mat) RT_HIP(ctx, batch_alloc(*b, &b->mat, (size_t)(cap_t ? cap_t : 1)));
    }
    if (b->shading.size() < b->nscenes) b->shading.resize(b->nscenes);
    Batch::Shading &S = b->shading[i];
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));  // the old table may still be read
    if (S.mats) {
        hipFree(S.mats);
        b->bytes -= sizeof(rt_material) * S.nmat;
        S.mats = nullptr;
        S.nmat = 0;
    }
    if (materials) {
        RT_HIP(ctx, hipMalloc(&S.mats, sizeof(rt_material) * nmat));
        S.nmat = nmat;
        b->bytes += sizeof(rt_material) * nmat;
        RT_HIP(ctx, hipMemcpyAsync(S.mats, materials, sizeof(rt_material) * nmat, hipMemcpyHostToDevice,
                                   ctx->stream));
    }
    if (normals && nv)
        RT_HIP(ctx, hipMemcpyAsync(b->nrm + 3 * (size_t)d.vert_off, normals, 12ull * nv, hipMemcpyHostToDevice,
                                   ctx->stream));
    if (mat_ids && nt)
        RT_HIP(ctx, hipMemcpyAsync(b->mat + d.in_prim_off, mat_ids, 4ull * nt, hipMemcpyHostToDevice, ctx->stream));
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    S.normals = normals != nullptr;
    S.ids = mat_ids != nullptr;
    S.nv = nv;
    S.nt = nt;
    return RT_OK;
}

int rt_scene_transform(rt_ctx *ctx, int scene, const float m[12])
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i) || !m) return fail(ctx, RT_ERR_ARG, "rt_scene_transform: bad argument");
    hipSetDevice(ctx->device);
    const SceneDesc &d = b->desc[i];
    if (!b->rest[i]) {
        if (!b->xyz_rest) {
            uint64_t cap = 0;
            for (uint32_t v : b->max_verts) cap += v;
            hipError_t e = batch_alloc(*b, &b->xyz_rest, 3 * (size_t)(cap ? cap : 1));
            if (e != hipSuccess) return hip_fail(ctx, e, "rt_scene_transform: hipMalloc");
        }
        if (d.nv)
            RT_HIP(ctx, hipMemcpyAsync(b->xyz_rest + 3 * (size_t)d.vert_off, b->xyz + 3 * (size_t)d.vert_off,
                                       12ull * d.nv, hipMemcpyDeviceToDevice, ctx->stream));
        b->rest[i] = 1;
    }
    rt::Xform x;
    memcpy(x.m, m, sizeof(x.m));
    RT_HIP(ctx, rt::launch_transform(*b, i, x, ctx->stream));
    b->built = false;
    drop_bvh(b, i);
    return RT_OK;
}

int rt_bvh_build_host(const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt, uint32_t max_prims,
                      uint32_t split, rt_node32 *nodes, int32_t *ordered_ids, rt_bvh_info *info)
{
    if (!xyz || !idx || nt == 0 || nt > (1u << 25) || max_prims == 0 || split > RT_BVH_EQUAL) return RT_ERR_ARG;
    for (uint64_t k = 0; k < 3ull * nt; ++k)
        if (idx[k] >= nv) return RT_ERR_ARG;
    rt::BvhHost h;
    rt::bvh_build_host(xyz, idx, nt, max_prims, split, h, info);
    if (nodes) memcpy(nodes, h.nodes.data(), sizeof(rt_node32) * h.nodes.size());
    if (ordered_ids) memcpy(ordered_ids, h.ordered.data(), sizeof(int32_t) * h.ordered.size());
    return RT_OK;
}

int rt_bvh_build(rt_ctx *ctx, int scene, uint32_t max_prims, uint32_t split, rt_bvh_info *info)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i)) return fail(ctx, RT_ERR_ARG, "rt_bvh_build: unknown scene %d", scene);
    if (max_prims == 0 || split > RT_BVH_EQUAL)
        return fail(ctx, RT_ERR_ARG, "rt_bvh_build: max_prims >= 1 and a split method 0..2");
    const SceneDesc &d = b->desc[i];
    if (d.nt == 0) return fail(ctx, RT_ERR_STATE, "rt_bvh_build: scene %d has no triangles", scene);
    hipSetDevice(ctx->device);
    // the scene's current vertices and indices, as the kernels see them
    std::vector<float> xyz(3 * (size_t)d.nv);
    std::vector<uint32_t> idx(3 * (size_t)d.nt);
    RT_HIP(ctx, hipMemcpyAsync(xyz.data(), b->xyz + 3 * (size_t)d.vert_off, 12ull * d.nv, hipMemcpyDeviceToHost,
                               ctx->stream));
    RT_HIP(ctx, hipMemcpyAsync(idx.data(), b->idx + 3 * (size_t)d.in_prim_off, 12ull * d.nt,
                               hipMemcpyDeviceToHost, ctx->stream));
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (b->bvh.size() < b->nscenes) b->bvh.resize(b->nscenes);
    if (b->use_bvh.size() < b->nscenes) b->use_bvh.assign(b->nscenes, 0);
    Batch::Bvh &v = b->bvh[i];
    v.valid = false;
    b->use_bvh[i] = 0;
    b->generation++;  // the arrays below are replaced
    rt::bvh_build_host(xyz.data(), idx.data(), d.nt, max_prims, split, v.host, info);
    std::vector<rt::TravNode> trav;
    std::vector<rt::TravTri> tris;
    rt::SceneHdr hdr;
    rt::bvh_device_layout(xyz.data(), idx.data(), v.host, trav, tris, hdr);
    hipFree(v.trav);
    hipFree(v.tris);
    hipFree(v.hdr);
    v.trav = nullptr;
    v.tris = nullptr;
    v.hdr = nullptr;
    b->bytes -= v.bytes;
    v.bytes = 0;
    RT_HIP(ctx, hipMalloc(&v.trav, sizeof(rt::TravNode) * trav.size()));
    RT_HIP(ctx, hipMalloc(&v.tris, sizeof(rt::TravTri) * tris.size()));
    RT_HIP(ctx, hipMalloc(&v.hdr, sizeof(rt::SceneHdr)));
    v.bytes = sizeof(rt::TravNode) * trav.size() + sizeof(rt::TravTri) * tris.size() + sizeof(rt::SceneHdr);
    b->bytes += v.bytes;
    RT_HIP(ctx, hipMemcpyAsync(v.trav, trav.data(), sizeof(rt::TravNode) * trav.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    RT_HIP(ctx, hipMemcpyAsync(v.tris, tris.data(), sizeof(rt::TravTri) * tris.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    RT_HIP(ctx, hipMemcpyAsync(v.hdr, &hdr, sizeof(hdr), hipMemcpyHostToDevice, ctx->stream));
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    v.valid = true;
    b->use_bvh[i] = 1;
    return RT_OK;
}

int rt_scene_select(rt_ctx *ctx, int scene, int structure)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i)) return fail(ctx, RT_ERR_ARG, "rt_scene_select: unknown scene %d", scene);
    if (structure == RT_STRUCTURE_BVH) {
        if (i >= b->bvh.size() || !b->bvh[i].valid)
            return fail(ctx, RT_ERR_STATE, "rt_scene_select: scene %d has no BVH (rt_bvh_build)", scene);
        if (!b->use_bvh[i]) b->generation++;
        b->use_bvh[i] = 1;
        return RT_OK;
    }
    if (structure == RT_STRUCTURE_LBVH) {
        if (i < b->use_bvh.size() && b->use_bvh[i]) {
            b->use_bvh[i] = 0;
            b->generation++;
        }
        return RT_OK;
    }
    return fail(ctx, RT_ERR_ARG, "rt_scene_select: structure %d (0 BVH, 9 LBVH)", structure);
}

int rt_bvh_read(rt_ctx *ctx, int scene, rt_node32 *nodes, int32_t *ordered_ids)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i)) return fail(ctx, RT_ERR_ARG, "rt_bvh_read: unknown scene %d", scene);
    if (i >= b->bvh.size() || !b->bvh[i].valid)
        return fail(ctx, RT_ERR_STATE, "rt_bvh_read: scene %d has no BVH", scene);
    const rt::BvhHost &h = b->bvh[i].host;
    if (nodes) memcpy(nodes, h.nodes.data(), sizeof(rt_node32) * h.nodes.size());
    if (ordered_ids) memcpy(ordered_ids, h.ordered.data(), sizeof(int32_t) * h.ordered.size());
    return RT_OK;
}

int rt_scene_info(rt_ctx *ctx, int scene, uint32_t *nt, uint32_t *nv, uint32_t *depth)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i)) return fail(ctx, RT_ERR_ARG, "rt_scene_info: unknown scene %d", scene);
    if (nt) *nt = b->desc[i].nt;
    if (nv) *nv = b->desc[i].nv;
    if (depth) *depth = b->desc[i].depth;
    return RT_OK;
}

int rt_lbvh_build(rt_ctx *ctx, int scene, rt_build_times *times)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i)) return fail(ctx, RT_ERR_ARG, "rt_lbvh_build: unknown scene %d", scene);
    hipSetDevice(ctx->device);
    // rt_build_times: an event at every stage boundary.  A profiling window
    // times the build end to end with two events only: each event recorded
    // between two kernels stalls the stream (about 5 us on gfx950), which
    // would make the window's builds slower than unobserved ones.
    hipEvent_t *ev = times ? ctx->ev : nullptr;
    hipEvent_t ends[6] = {};
    if (!times && ctx->profiling) {
        hipEvent_t *p = prof_events(ctx, 0, 2);
        if (!p) return fail(ctx, RT_ERR_HIP, "rt_lbvh_build: hipEventCreate failed");
        ends[0] = p[0];
        ends[5] = p[1];
        ev = ends;
    }
    RT_HIP(ctx, rt::launch_build(*b, ctx->stream, ev, ctx->cu_count));
    b->built = true;
    if (times) {
        RT_HIP(ctx, hipEventSynchronize(ctx->ev[5]));
        float ms[5];
        for (int k = 0; k < 5; ++k) hipEventElapsedTime(&ms[k], ctx->ev[k], ctx->ev[k + 1]);
        times->bounds_ms = ms[0];
        times->morton_ms = ms[1];
        times->sort_ms = ms[2];
        times->emit_ms = ms[3];
        times->refit_ms = ms[4];
        hipEventElapsedTime(&times->total_ms, ctx->ev[0], ctx->ev[5]);
    }
    return RT_OK;
}

__global__ void k_local_ids(const int32_t *vals, uint32_t off, uint32_t n, int32_t *out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = vals[off + i] - (int32_t)off;
}

int rt_lbvh_read(rt_ctx *ctx, int scene, rt_node32 *nodes, uint32_t *sorted_codes,
                 int32_t *sorted_ids, float scene_box[6])
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i)) return fail(ctx, RT_ERR_ARG, "rt_lbvh_read: unknown scene %d", scene);
    if (!b->built) return fail(ctx, RT_ERR_STATE, "rt_lbvh_read: scene %d not built", scene);
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const SceneDesc &d = b->desc[i];
    if (nodes)
        RT_HIP(ctx, hipMemcpyAsync(nodes, b->heap + d.heap_off, sizeof(rt_node32) << d.depth,
                                   hipMemcpyDeviceToHost, st));
    if (sorted_codes && d.nt)
        RT_HIP(ctx, hipMemcpyAsync(sorted_codes, b->keys[0] + d.prim_off, 4ull * d.nt,
                                   hipMemcpyDeviceToHost, st));
    if (sorted_ids && d.nt) {
        // ids are batch-global on the device; shift them to the scene's own
        int32_t *tmp = b->vals[1];
        hipLaunchKernelGGL(k_local_ids, dim3((d.nt + 255) / 256), dim3(256), 0, st, b->vals[0],
                           d.prim_off, d.nt, tmp);
        RT_HIP(ctx, hipMemcpyAsync(sorted_ids, tmp, 4ull * d.nt, hipMemcpyDeviceToHost, st));
    }
    rt::SceneHdr h;
    RT_HIP(ctx, hipMemcpyAsync(&h, b->hdr + i, sizeof(h), hipMemcpyDeviceToHost, st));
    RT_HIP(ctx, hipStreamSynchronize(st));
    if (scene_box) memcpy(scene_box, h.scene_box, sizeof(float) * 6);
    return RT_OK;
}

static int trace_common(rt_ctx *ctx, Batch *b, uint32_t i, const rt_ray48 *d_rays, uint32_t n,
                        rt_hit24 *d_hits, uint32_t flags, rt_frame_stats *stats)
{
    hipStream_t st = ctx->stream;
    const bool count = (flags & RT_FLAG_COUNT) != 0;
    const bool timing = (flags & RT_FLAG_TIMING) != 0 || stats != nullptr;
    if (stats) RT_HIP(ctx, hipMemsetAsync(ctx->stats, 0, 8 * sizeof(unsigned long long), st));
    hipEvent_t *ev = timing ? &ctx->ev[6] : nullptr;
    if (!stats && ctx->profiling) ev = prof_events(ctx, 1, 2);
    if (ev) hipEventRecord(ev[0], st);
    RT_HIP(ctx, rt::launch_trace(*b, i, d_rays, n, d_hits, flags, ctx->stats, ctx->tune, st));
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
    }
    return RT_OK;
}

int rt_trace(rt_ctx *ctx, int scene, const rt_ray48 *rays, uint32_t n, rt_hit24 *hits,
             uint32_t flags, rt_frame_stats *stats)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i)) return fail(ctx, RT_ERR_ARG, "rt_trace: unknown scene %d", scene);
    if (!traversable(*b, i)) return fail(ctx, RT_ERR_STATE, "rt_trace: scene %d not built", scene);
    if (n && (!rays || !hits)) return fail(ctx, RT_ERR_ARG, "rt_trace: null array");
    if ((flags & RT_FLAG_START_AT_IROOT) && uses_bvh(*b, i))
        return fail(ctx, RT_ERR_STATE, "rt_trace: RT_FLAG_START_AT_IROOT needs the LBVH's heap node numbers");
    hipSetDevice(ctx->device);
    const uint64_t rb = (sizeof(rt_ray48) * (uint64_t)n + 255) & ~255ull;
    int rc = ensure_staging(ctx, rb + sizeof(rt_hit24) * (uint64_t)n + 256);
    if (rc) return rc;
    rt_ray48 *d_rays = (rt_ray48 *)ctx->staging;
    rt_hit24 *d_hits = (rt_hit24 *)((char *)ctx->staging + rb);
    if (n)
        RT_HIP(ctx, hipMemcpyAsync(d_rays, rays, sizeof(rt_ray48) * (uint64_t)n,
                                   hipMemcpyHostToDevice, ctx->stream));
    if (n && (flags & RT_FLAG_START_AT_IROOT))  // the start nodes travel in the hits
        RT_HIP(ctx, hipMemcpyAsync(d_hits, hits, sizeof(rt_hit24) * (uint64_t)n, hipMemcpyHostToDevice,
                                   ctx->stream));
    rc = trace_common(ctx, b, i, d_rays, n, d_hits, flags, stats);
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
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i)) return fail(ctx, RT_ERR_ARG, "rt_trace_device: unknown scene %d", scene);
    if (!traversable(*b, i)) return fail(ctx, RT_ERR_STATE, "rt_trace_device: scene %d not built", scene);
    if (n && (!d_rays || !d_hits)) return fail(ctx, RT_ERR_ARG, "rt_trace_device: null array");
    if (((uintptr_t)d_rays & 15) || ((uintptr_t)d_hits & 7))
        return fail(ctx, RT_ERR_ARG, "rt_trace_device: rays need 16-byte, hits 8-byte alignment");
    if ((flags & RT_FLAG_START_AT_IROOT) && uses_bvh(*b, i))
        return fail(ctx, RT_ERR_STATE, "rt_trace_device: RT_FLAG_START_AT_IROOT needs the LBVH's heap node numbers");
    hipSetDevice(ctx->device);
    return trace_common(ctx, b, i, d_rays, n, d_hits, flags, stats);
}

int rt_render_direct(rt_ctx *ctx, int scene, const rt_render_desc *desc, uint32_t flags,
                     rt_frame_stats *stats)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i) || !desc) return fail(ctx, RT_ERR_ARG, "rt_render_direct: bad argument");
    if (!traversable(*b, i)) return fail(ctx, RT_ERR_STATE, "rt_render_direct: scene %d not built", scene);
    if (desc->nlights > 8 || (desc->nlights && !desc->lights))
        return fail(ctx, RT_ERR_ARG, "rt_render_direct: 0..8 lights with a light array");
    if (desc->width == 0 || desc->height == 0 || desc->width > 16384 || desc->height > 16384)
        return fail(ctx, RT_ERR_ARG, "rt_render_direct: image %ux%u", desc->width, desc->height);
    if (!rt::camera_ok(desc->cam)) return fail(ctx, RT_ERR_ARG, "rt_render_direct: camera kind %u", desc->cam.kind);
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
    RT_HIP(ctx, rt::launch_render_direct(*b, i, *desc, ctx->frame, flags, ctx->stats, ctx->tune, st));
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

int rt_render_whitted(rt_ctx *ctx, int scene, const rt_whitted_desc *desc, uint32_t flags, rt_frame_stats *stats)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i) || !desc) return fail(ctx, RT_ERR_ARG, "rt_render_whitted: bad argument");
    if (!traversable(*b, i)) return fail(ctx, RT_ERR_STATE, "rt_render_whitted: scene %d not built", scene);
    if (desc->width == 0 || desc->height == 0 || desc->width > 16384 || desc->height > 16384)
        return fail(ctx, RT_ERR_ARG, "rt_render_whitted: image %ux%u", desc->width, desc->height);
    if (!rt::camera_ok(desc->cam)) return fail(ctx, RT_ERR_ARG, "rt_render_whitted: camera kind %u", desc->cam.kind);
    if (desc->bounces > 16) return fail(ctx, RT_ERR_ARG, "rt_render_whitted: %u bounces (0..16)", desc->bounces);
    if (desc->env_flag != -1 && desc->env_flag != 0)
        return fail(ctx, RT_ERR_ARG, "rt_render_whitted: env_flag %d (-1 or 0)", desc->env_flag);
    if (desc->switches & ~(uint32_t)(RT_WHITTED_SHADOWS | RT_WHITTED_PHONG | RT_WHITTED_GLOSS | RT_WHITTED_ONE_PASS |
                                     RT_WHITTED_SPLIT | RT_WHITTED_QUADS))
        return fail(ctx, RT_ERR_ARG, "rt_render_whitted: unknown switches 0x%x", desc->switches);
    if ((desc->switches & RT_WHITTED_ONE_PASS) && (desc->switches & RT_WHITTED_SPLIT))
        return fail(ctx, RT_ERR_ARG, "rt_render_whitted: RT_WHITTED_ONE_PASS and RT_WHITTED_SPLIT together");
    if ((desc->switches & RT_WHITTED_QUADS) && !(desc->switches & RT_WHITTED_SPLIT))
        return fail(ctx, RT_ERR_ARG, "rt_render_whitted: RT_WHITTED_QUADS without RT_WHITTED_SPLIT");
    const SceneDesc &d = b->desc[i];
    const bool have = i < b->shading.size();
    if (have && (b->shading[i].nv != d.nv || b->shading[i].nt != d.nt) &&
        (b->shading[i].normals || b->shading[i].ids || b->shading[i].mats))
        return fail(ctx, RT_ERR_STATE, "rt_render_whitted: scene %d was re-uploaded after rt_scene_shading", scene);
    if ((desc->switches & RT_WHITTED_PHONG) && !(have && b->shading[i].normals))
        return fail(ctx, RT_ERR_STATE, "rt_render_whitted: Phong shading needs rt_scene_shading normals");
    hipSetDevice(ctx->device);
    int rc = ensure_frame(ctx, desc->width, desc->height, 1);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    // wavefront form: two ray queues of one 64-byte record per pixel
    const uint64_t qbytes = rt::whitted_queue_bytes((uint64_t)desc->width * desc->height);
    if (!(desc->switches & RT_WHITTED_ONE_PASS) && desc->bounces > 0 && ctx->wqueue_bytes < qbytes) {
        hipStreamSynchronize(st);
        hipFree(ctx->wqueue);
        ctx->bytes -= ctx->wqueue_bytes;
        ctx->wqueue = nullptr;
        ctx->wqueue_bytes = 0;
        RT_HIP(ctx, hipMalloc(&ctx->wqueue, qbytes));
        ctx->wqueue_bytes = qbytes;
        ctx->bytes += qbytes;
    }
    const bool timing = (flags & RT_FLAG_TIMING) != 0 || stats != nullptr;
    if (stats || !ctx->profiling)
        RT_HIP(ctx, hipMemsetAsync(ctx->stats, 0, 8 * sizeof(unsigned long long), st));
    hipEvent_t *ev = timing ? &ctx->ev[6] : nullptr;
    if (!stats && ctx->profiling) ev = prof_events(ctx, 1, 2);
    if (ev) hipEventRecord(ev[0], st);
    RT_HIP(ctx, rt::launch_render_whitted(*b, i, *desc, ctx->frame, flags, ctx->stats, ctx->tune, ctx->wqueue,
                                          ctx->wcount, ctx->cu_count, st));
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
        stats->bounce_rays = c[3];
        stats->box_tests = c[0];
        stats->tri_tests = c[1];
    }
    ctx->frame.nlights = 1;
    return RT_OK;
}

// ---------------------------------------------------------------------------
// Render jobs, tuning, graphs
// ---------------------------------------------------------------------------
static const int kForms = 3;  // traversal forms: loop, step, loop with the shared fetch

static int check_tuning(rt_ctx *ctx, const rt_tuning &t)
{
    if (t.block != 64 && t.block != 128 && t.block != 256)
        return fail(ctx, RT_ERR_ARG, "tuning: block must be 64, 128 or 256");
    if (t.lds_stack < 1 || t.lds_stack > 26) return fail(ctx, RT_ERR_ARG, "tuning: lds_stack in 1..26");
    if (t.blocks_per_cu < 1 || t.blocks_per_cu > 32)
        return fail(ctx, RT_ERR_ARG, "tuning: blocks_per_cu in 1..32");
    if (t.traversal < 0 || t.traversal >= kForms) return fail(ctx, RT_ERR_ARG, "tuning: traversal 0, 1 or 2");
    if (t.quad_tiles < 0 || t.quad_tiles > (1 << 24)) return fail(ctx, RT_ERR_ARG, "tuning: quad_tiles in 0..2^24");
    return RT_OK;
}

// The instantiated launch shapes, in the order of rt::tune_index.
static const rt_tuning kShapes[] = {{64, 16, 4, 0},  {64, 20, 4, 0},  {128, 12, 4, 0}, {128, 16, 4, 0},
                                    {128, 20, 4, 0}, {128, 22, 4, 0}, {128, 26, 4, 0}, {256, 12, 4, 0}};

int rt_tuning_get(rt_ctx *ctx, rt_tuning *t)
{
    if (!ctx || !t) return RT_ERR_ARG;
    *t = ctx->tune;
    return RT_OK;
}

int rt_tuning_set(rt_ctx *ctx, const rt_tuning *t)
{
    if (!ctx || !t) return RT_ERR_ARG;
    int rc = check_tuning(ctx, *t);
    if (rc) return rc;
    // snap to the nearest instantiated shape; keep blocks_per_cu as given
    rt_tuning s = kShapes[rt::tune_index(*t)];
    s.blocks_per_cu = t->blocks_per_cu;
    s.traversal = t->traversal;
    s.quad_tiles = t->quad_tiles;
    ctx->tune = s;
    return RT_OK;
}

// Builds the device job table for jobs; returns the total tile count.
// The device copy of the RT_TILES_CENTRE order of a width x height image in
// tiles of 16 x tile_h, made on first use and kept until rt_destroy; null with
// may_alloc false when it is not there yet.
static int centre_order(rt_ctx *ctx, uint32_t width, uint32_t height, uint32_t tile_h, bool may_alloc,
                        const uint32_t **out)
{
    *out = nullptr;
    const auto key = std::make_tuple(width, height, tile_h);
    auto it = ctx->tile_orders.find(key);
    if (it != ctx->tile_orders.end()) {
        *out = it->second;
        return RT_OK;
    }
    if (!may_alloc) return RT_OK;
    std::vector<uint32_t> order;
    rt::centre_tile_order(width, height, tile_h, order);
    uint32_t *d = nullptr;
    RT_HIP(ctx, hipMalloc(&d, sizeof(uint32_t) * order.size()));
    const hipError_t e = hipMemcpy(d, order.data(), sizeof(uint32_t) * order.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(d);
        return hip_fail(ctx, e, "tile order upload");
    }
    ctx->bytes += sizeof(uint32_t) * order.size();
    ctx->tile_orders.emplace(key, d);
    *out = d;
    return RT_OK;
}

// Checks the jobs and writes their device records into host (render_job_bytes()
// each) and the first tile of each job into offs[0..njobs] (optional).
// may_alloc false (rt_render_jobs_noalloc): a RT_TILES_CENTRE job whose order
// is not on the device yet renders in row order, which gives the same image
// when it covers the whole image; a tile range over such an order is refused.
static int fill_jobs(rt_ctx *ctx, const rt_job_desc *jobs, uint32_t njobs, const rt_tuning &t,
                     unsigned char *host, uint32_t *offs, uint32_t *total, uint32_t *max_depth,
                     bool may_alloc = true)
{
    const size_t jb = rt::render_job_bytes();
    uint32_t off = 0;
    *max_depth = 0;
    for (uint32_t k = 0; k < njobs; ++k) {
        const rt_job_desc &d = jobs[k];
        Batch *b;
        uint32_t i;
        if (!get_scene(ctx, d.scene, &b, &i)) return fail(ctx, RT_ERR_ARG, "job %u: unknown scene %d", k, d.scene);
        if (!traversable(*b, i)) return fail(ctx, RT_ERR_STATE, "job %u: scene %d not built", k, d.scene);
        if (!d.accum || ((uintptr_t)d.accum & 15))
            return fail(ctx, RT_ERR_ARG, "job %u: accum must be a 16-byte aligned device pointer", k);
        if (d.width == 0 || d.height == 0 || d.width > 16384 || d.height > 16384 || d.spp == 0 ||
            d.spp > 65536 || d.max_bounces > 64 || d.nlights > 8 || (d.nlights && !d.lights) ||
            (d.mode != RT_MODE_DIRECT && d.mode != RT_MODE_GI) || !rt::camera_ok(d.cam) ||
            (d.tile_order != RT_TILES_ROWS && d.tile_order != RT_TILES_CENTRE))
            return fail(ctx, RT_ERR_ARG, "job %u: bad size, spp, bounces, lights, mode, camera or tile order", k);
        const uint32_t *order = nullptr;
        if (d.tile_order == RT_TILES_CENTRE) {
            int rc = centre_order(ctx, d.width, d.height, (uint32_t)t.block / 16, may_alloc, &order);
            if (rc) return rc;
            if (!order && d.tile_count)
                return fail(ctx, RT_ERR_STATE, "job %u: a tile range over a centre order not made yet", k);
        }
        uint32_t tiles = 0;
        rt::fill_render_job(host + jb * k, *b, i, d, t, off, &tiles, order);
        if (trav_depth(*b, i) > *max_depth) *max_depth = trav_depth(*b, i);
        if (offs) offs[k] = off;
        off += tiles;
    }
    if (offs) offs[njobs] = off;
    *total = off;
    return RT_OK;
}

static int stage_jobs(rt_ctx *ctx, const rt_job_desc *jobs, uint32_t njobs, const rt_tuning &t,
                      void *dst_dev, std::vector<uint32_t> *tile_offs, uint32_t *total, uint32_t *max_depth)
{
    const size_t jb = rt::render_job_bytes();
    std::vector<unsigned char> host(jb * njobs);
    if (tile_offs) tile_offs->assign(njobs + 1, 0u);
    int rc = fill_jobs(ctx, jobs, njobs, t, host.data(), tile_offs ? tile_offs->data() : nullptr, total, max_depth);
    if (rc) return rc;
    // the previous launch may still read the table: wait for it before overwriting
    RT_HIP(ctx, hipEventSynchronize(ctx->jobs_ev));
    RT_HIP(ctx, hipMemcpy(dst_dev, host.data(), jb * njobs, hipMemcpyHostToDevice));
    return RT_OK;
}

static int ensure_jobs(rt_ctx *ctx, uint32_t njobs)
{
    const uint64_t need = (uint64_t)rt::render_job_bytes() * (njobs ? njobs : 1);
    if (ctx->jobs_cap >= need) return RT_OK;
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(ctx->jobs_dev);
    ctx->bytes -= ctx->jobs_cap;
    ctx->jobs_dev = nullptr;
    ctx->jobs_cap = 0;
    RT_HIP(ctx, hipMalloc(&ctx->jobs_dev, need));
    ctx->jobs_cap = need;
    ctx->bytes += need;
    return RT_OK;
}

static int run_jobs(rt_ctx *ctx, const rt_job_desc *jobs, uint32_t njobs, uint32_t flags,
                    const rt_tuning &t, rt_frame_stats *stats, float *ms_out)
{
    int rc = ensure_jobs(ctx, njobs);
    if (rc) return rc;
    std::vector<uint32_t> offs;
    uint32_t total = 0, depth = 0;
    rc = stage_jobs(ctx, jobs, njobs, t, ctx->jobs_dev, &offs, &total, &depth);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    const bool persistent = (flags & RT_FLAG_PERSISTENT) != 0;
    const bool timing = stats != nullptr || ms_out != nullptr || (flags & RT_FLAG_TIMING);
    if (stats || !ctx->profiling)
        RT_HIP(ctx, hipMemsetAsync(ctx->stats, 0, 8 * sizeof(unsigned long long), st));
    hipEvent_t *ev = timing ? &ctx->ev[6] : nullptr;
    if (!timing && ctx->profiling) ev = prof_events(ctx, 1, 2);
    if (ev) hipEventRecord(ev[0], st);
    const bool count = (flags & RT_FLAG_COUNT) != 0;  // one lane per ray, tests counted
    if (persistent) {
        RT_HIP(ctx, rt::launch_render_jobs(ctx->jobs_dev, njobs, total, 0, total, true, ctx->tile_counter,
                                           nullptr, ctx->stats, t, depth, ctx->cu_count, st, count));
    } else {
        for (uint32_t k = 0; k < njobs; ++k)
            RT_HIP(ctx, rt::launch_render_jobs(ctx->jobs_dev, njobs, total, offs[k],
                                               offs[k + 1] - offs[k], false, ctx->tile_counter,
                                               nullptr, ctx->stats, t, depth, ctx->cu_count, st, count));
    }
    if (ev) hipEventRecord(ev[1], st);
    RT_HIP(ctx, hipEventRecord(ctx->jobs_ev, st));
    if (ctx->profiling) {
        for (uint32_t k = 0; k < njobs; ++k)
            ctx->prof_primary += (uint64_t)jobs[k].width * jobs[k].height * jobs[k].spp;
    }
    if (timing) {
        unsigned long long c[8] = {};
        RT_HIP(ctx, hipMemcpyAsync(c, ctx->stats, sizeof(c), hipMemcpyDeviceToHost, st));
        RT_HIP(ctx, hipStreamSynchronize(st));
        float ms = 0.0f;
        hipEventElapsedTime(&ms, ev[0], ev[1]);
        if (ms_out) *ms_out = ms;
        if (stats) {
            memset(stats, 0, sizeof(*stats));
            stats->trace_ms = ms;
            stats->primary_rays = c[4];
            stats->shadow_rays = c[2];
            stats->bounce_rays = c[3];
            stats->box_tests = count ? c[0] : 0u;
            stats->tri_tests = count ? c[1] : 0u;
        }
    }
    return RT_OK;
}

int rt_render_jobs(rt_ctx *ctx, const rt_job_desc *jobs, uint32_t njobs, uint32_t flags,
                   rt_frame_stats *stats)
{
    if (!ctx || (njobs && !jobs)) return fail(ctx, RT_ERR_ARG, "rt_render_jobs: bad argument");
    if (njobs == 0) return RT_OK;
    hipSetDevice(ctx->device);
    return run_jobs(ctx, jobs, njobs, flags, ctx->tune, stats, nullptr);
}

int rt_render_jobs_noalloc(rt_ctx *ctx, const rt_job_desc *jobs, uint32_t njobs, uint32_t flags,
                           rt_frame_stats *stats)
{
    if (!ctx || (njobs && !jobs)) return fail(ctx, RT_ERR_ARG, "rt_render_jobs_noalloc: bad argument");
    if (njobs > RT_NOALLOC_MAX_JOBS)
        return fail(ctx, RT_ERR_CAPACITY, "rt_render_jobs_noalloc: %u jobs, at most %d", njobs,
                    RT_NOALLOC_MAX_JOBS);
    if (njobs == 0) return RT_OK;
    if (!ctx->noalloc_scratch) return fail(ctx, RT_ERR_STATE, "rt_render_jobs_noalloc: no scratch on this device");
    if (ctx->profiling) return fail(ctx, RT_ERR_STATE, "rt_render_jobs_noalloc: inside a profiling window");
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    // the previous call's copy may still read the host staging: wait for it
    RT_HIP(ctx, hipEventSynchronize(ctx->noalloc_ev));
    uint32_t total = 0, depth = 0;
    int rc = fill_jobs(ctx, jobs, njobs, ctx->tune, ctx->noalloc_host.data(), ctx->noalloc_offs.data(), &total,
                       &depth, false);
    if (rc) return rc;
    unsigned char *base = ctx->noalloc_scratch;
    void *d_jobs = base;
    uint32_t *d_counter = reinterpret_cast<uint32_t *>(base + kNoAllocCounterOff);
    unsigned long long *d_stats = reinterpret_cast<unsigned long long *>(base + kNoAllocStatsOff);
    // ordered on the stream after earlier launches that read the table
    RT_HIP(ctx, hipMemcpyAsync(d_jobs, ctx->noalloc_host.data(), rt::render_job_bytes() * njobs,
                               hipMemcpyHostToDevice, st));
    RT_HIP(ctx, hipEventRecord(ctx->noalloc_ev, st));
    RT_HIP(ctx, rt::launch_zero_words(reinterpret_cast<uint32_t *>(d_stats), 16, st));
    if (stats) RT_HIP(ctx, hipEventRecord(ctx->ev[6], st));
    const uint32_t *offs = ctx->noalloc_offs.data();
    const bool count = (flags & RT_FLAG_COUNT) != 0;  // one lane per ray, tests counted
    if (flags & RT_FLAG_PERSISTENT) {
        RT_HIP(ctx, rt::launch_render_jobs(d_jobs, njobs, total, 0, total, true, d_counter, nullptr, d_stats,
                                           ctx->tune, depth, ctx->cu_count, st, count));
    } else {
        for (uint32_t k = 0; k < njobs; ++k)
            RT_HIP(ctx, rt::launch_render_jobs(d_jobs, njobs, total, offs[k], offs[k + 1] - offs[k], false,
                                               d_counter, nullptr, d_stats, ctx->tune, depth, ctx->cu_count, st,
                                               count));
    }
    if (stats) {
        RT_HIP(ctx, hipEventRecord(ctx->ev[7], st));
        unsigned long long c[8] = {};
        RT_HIP(ctx, hipMemcpyAsync(c, d_stats, sizeof(c), hipMemcpyDeviceToHost, st));
        RT_HIP(ctx, hipStreamSynchronize(st));
        memset(stats, 0, sizeof(*stats));
        hipEventElapsedTime(&stats->trace_ms, ctx->ev[6], ctx->ev[7]);
        stats->primary_rays = c[4];
        stats->shadow_rays = c[2];
        stats->bounce_rays = c[3];
        stats->box_tests = count ? c[0] : 0u;
        stats->tri_tests = count ? c[1] : 0u;
    }
    return RT_OK;
}

static std::string tune_key(rt_ctx *ctx, const rt_job_desc *jobs, uint32_t njobs)
{
    uint64_t px = 0, tris = 0, bounces = 0, spp = 0, modes = 0;
    for (uint32_t k = 0; k < njobs; ++k) {
        px += (uint64_t)jobs[k].width * jobs[k].height;
        Batch *b;
        uint32_t i;
        if (get_scene(ctx, jobs[k].scene, &b, &i)) tris += b->desc[i].nt;
        bounces += jobs[k].max_bounces;
        spp += jobs[k].spp;
        modes |= 1ull << jobs[k].mode;
    }
    char buf[256];
    snprintf(buf, sizeof(buf), "jobs%u_px%llu_tris%llu_spp%llu_b%llu_m%llu_cu%d", njobs,
             (unsigned long long)px, (unsigned long long)tris, (unsigned long long)spp,
             (unsigned long long)bounces, (unsigned long long)modes, ctx->cu_count);
    return buf;
}

// Autotune cache: a text file, one record per line,
//   <key> block lds_stack blocks_per_cu traversal quad_tiles extra
// (extra: the Whitted pass form, 0 for the other searches).  Lines of any
// other shape are skipped; the last valid record of a key wins.
static bool cache_find(rt_ctx *ctx, const char *path, const char *key, rt_tuning *t, int *extra)
{
    FILE *f = path ? fopen(path, "r") : nullptr;
    if (!f) return false;
    bool found = false;
    char line[512];
    while (fgets(line, sizeof(line), f)) {
        char k[256];
        rt_tuning x{};
        int e = 0, n = 0;
        if (sscanf(line, "%255s %d %d %d %d %d %d %n", k, &x.block, &x.lds_stack, &x.blocks_per_cu, &x.traversal,
                   &x.quad_tiles, &e, &n) == 7 &&
            line[n] == '\0' && strcmp(k, key) == 0 && check_tuning(ctx, x) == RT_OK) {
            *t = x;
            *extra = e;
            found = true;
        }
    }
    fclose(f);
    return found;
}

static void cache_add(const char *path, const char *key, const rt_tuning &t, int extra)
{
    if (FILE *f = path ? fopen(path, "a") : nullptr) {
        fprintf(f, "%s %d %d %d %d %d %d\n", key, t.block, t.lds_stack, t.blocks_per_cu, t.traversal, t.quad_tiles,
                extra);
        fclose(f);
    }
}

// Times every launch shape (block size fixed_block only, when non-zero) on
// the jobs in persistent launches into scratch accumulators; *best = the
// fastest (quad_tiles kept from the context).  The context's tuning is left as is.
static int search_jobs_shape(rt_ctx *ctx, const rt_job_desc *jobs, uint32_t njobs, int fixed_block,
                             rt_tuning *best)
{
    // scratch accumulators so the search leaves the callers' buffers alone
    std::vector<rt_job_desc> tj(jobs, jobs + njobs);
    uint64_t need = 0;
    for (const rt_job_desc &d : tj) need += 16ull * d.width * d.height;
    void *scratch = nullptr;
    RT_HIP(ctx, hipMalloc(&scratch, need ? need : 16));
    uint64_t off = 0;
    for (rt_job_desc &d : tj) {
        d.accum = reinterpret_cast<float *>(static_cast<char *>(scratch) + off);
        off += 16ull * d.width * d.height;
    }
    rt_tuning bt = ctx->tune;
    float bms = 1e30f;
    bool any = false;
    const int bpcs[] = {2, 4, 8};
    int rc = RT_OK;
    for (int form = 0; form < kForms && rc == RT_OK; ++form) {
        for (const rt_tuning &shape : kShapes) {
            if (fixed_block && shape.block != fixed_block) continue;
            for (int bpc : bpcs) {
                rt_tuning t = shape;
                t.blocks_per_cu = bpc;
                t.traversal = form;
                t.quad_tiles = ctx->tune.quad_tiles;  // not used by jobs; kept for later direct frames
                float ms = 0.0f;
                rc = run_jobs(ctx, tj.data(), njobs, RT_FLAG_PERSISTENT, t, nullptr, &ms);  // warm
                if (rc) break;
                rc = run_jobs(ctx, tj.data(), njobs, RT_FLAG_PERSISTENT, t, nullptr, &ms);
                if (rc) break;
                if (ms < bms) {
                    bms = ms;
                    bt = t;
                    any = true;
                }
            }
            if (rc) break;
        }
    }
    hipStreamSynchronize(ctx->stream);
    hipFree(scratch);
    if (rc) return rc;
    if (!any) return fail(ctx, RT_ERR_ARG, "launch shape search: no shape of block %d", fixed_block);
    *best = bt;
    return RT_OK;
}

}  // extern "C"

namespace rt {
int search_jobs_shape_block(rt_ctx *ctx, const rt_job_desc *jobs, uint32_t njobs, int block, rt_tuning *best)
{
    return search_jobs_shape(ctx, jobs, njobs, block, best);
}
}  // namespace rt

extern "C" {

int rt_autotune_jobs(rt_ctx *ctx, const rt_job_desc *jobs, uint32_t njobs, const char *cache_path,
                     rt_tuning *best)
{
    if (!ctx || !jobs || njobs == 0) return fail(ctx, RT_ERR_ARG, "rt_autotune_jobs: bad argument");
    hipSetDevice(ctx->device);
    const std::string key = tune_key(ctx, jobs, njobs);
    {
        rt_tuning t{};
        int extra = 0;
        if (cache_find(ctx, cache_path, key.c_str(), &t, &extra)) {
            t.quad_tiles = ctx->tune.quad_tiles;  // quads are a direct-frame choice: kept as set
            ctx->tune = t;
            if (best) *best = t;
            return RT_OK;
        }
    }
    rt_tuning bt = ctx->tune;
    const int rc = search_jobs_shape(ctx, jobs, njobs, 0, &bt);
    if (rc) return rc;
    ctx->tune = bt;
    if (best) *best = bt;
    cache_add(cache_path, key.c_str(), bt, 0);
    return RT_OK;
}

int rt_autotune_direct(rt_ctx *ctx, int scene, const rt_render_desc *desc, const char *cache_path,
                       rt_tuning *best)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i) || !desc) return fail(ctx, RT_ERR_ARG, "rt_autotune_direct: bad argument");
    if (ctx->profiling) return fail(ctx, RT_ERR_STATE, "rt_autotune_direct: inside a profiling window");
    char key[160];
    snprintf(key, sizeof(key), "direct:cu%d:%s:t%u:d%u:%ux%u:l%u", ctx->cu_count, uses_bvh(*b, i) ? "bvh" : "lbvh",
             b->desc[i].nt, trav_depth(*b, i), desc->width, desc->height, desc->nlights);
    {
        rt_tuning t{};
        int extra = 0;
        if (cache_find(ctx, cache_path, key, &t, &extra)) {
            ctx->tune = t;
            if (best) *best = t;
            return RT_OK;
        }
    }
    const rt_tuning saved = ctx->tune;
    rt_tuning bt = saved;
    float bms = 1e30f;
    int rc = RT_OK;
    for (int form = 0; form < kForms && rc == RT_OK; ++form) {
        for (const rt_tuning &shape : kShapes) {
            rt_tuning t = shape;
            t.blocks_per_cu = saved.blocks_per_cu;
            t.traversal = form;
            ctx->tune = t;
            float ms = 1e30f;
            for (int rep = 0; rep < 3 && rc == RT_OK; ++rep) {  // first run warms caches
                rt_frame_stats st;
                rc = rt_render_direct(ctx, scene, desc, 0, &st);
                if (rep > 0 && st.trace_ms < ms) ms = st.trace_ms;
            }
            if (rc) break;
            if (ms < bms) {
                bms = ms;
                bt = t;
            }
        }
    }
    // quads for the heaviest tiles, with the fastest shape and form (the
    // frames of the loop above leave a tile order for this view); kept only
    // when at least 3 % faster, so that timing noise does not pick them
    static const int32_t kQuadTiles[] = {64, 256, 1024};
    const float one_lane_ms = bms;
    if (rc == RT_OK && !uses_bvh(*b, i)) {
        for (int32_t q : kQuadTiles) {
            rt_tuning t = bt;
            t.quad_tiles = q;
            ctx->tune = t;
            float ms = 1e30f;
            for (int rep = 0; rep < 3 && rc == RT_OK; ++rep) {
                rt_frame_stats st;
                rc = rt_render_direct(ctx, scene, desc, 0, &st);
                if (rep > 0 && st.trace_ms < ms) ms = st.trace_ms;
            }
            if (rc) break;
            if (ms < bms && ms < 0.97f * one_lane_ms) {
                bms = ms;
                bt = t;
            }
        }
    }
    ctx->tune = rc ? saved : bt;
    if (rc) return rc;
    if (best) *best = bt;
    cache_add(cache_path, key, bt, 0);
    return RT_OK;
}

// The Whitted pass forms the autotuner tries: 0 per-pass wavefront,
// 1 RT_WHITTED_ONE_PASS, 2 RT_WHITTED_SPLIT, 3 RT_WHITTED_SPLIT | RT_WHITTED_QUADS.
constexpr int kWhittedForms = 4;
static uint32_t whitted_pass_form(uint32_t switches, int form)
{
    switches &= ~(uint32_t)(RT_WHITTED_ONE_PASS | RT_WHITTED_SPLIT | RT_WHITTED_QUADS);
    if (form == 1) switches |= RT_WHITTED_ONE_PASS;
    if (form == 2) switches |= RT_WHITTED_SPLIT;
    if (form == 3) switches |= RT_WHITTED_SPLIT | RT_WHITTED_QUADS;
    return switches;
}

int rt_autotune_whitted(rt_ctx *ctx, int scene, rt_whitted_desc *desc, const char *cache_path, rt_tuning *best)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i) || !desc) return fail(ctx, RT_ERR_ARG, "rt_autotune_whitted: bad argument");
    if (ctx->profiling) return fail(ctx, RT_ERR_STATE, "rt_autotune_whitted: inside a profiling window");
    char key[192];
    snprintf(key, sizeof(key), "whitted:cu%d:%s:t%u:d%u:%ux%u:b%u:s%u", ctx->cu_count,
             uses_bvh(*b, i) ? "bvh" : "lbvh", b->desc[i].nt, trav_depth(*b, i), desc->width, desc->height,
             desc->bounces, whitted_pass_form(desc->switches, 0));
    {
        rt_tuning t{};
        int pf = 0;
        if (cache_find(ctx, cache_path, key, &t, &pf) && pf >= 0 && pf < kWhittedForms) {
            t.quad_tiles = ctx->tune.quad_tiles;  // a direct-frame choice: kept as set
            ctx->tune = t;
            desc->switches = whitted_pass_form(desc->switches, pf);
            if (best) *best = t;
            return RT_OK;
        }
    }
    const rt_tuning saved = ctx->tune;
    rt_tuning bt = saved;
    int bpf = 0;
    float bms = 1e30f;
    int rc = RT_OK;
    rt_whitted_desc d = *desc;
    for (int pf = 0; pf < kWhittedForms && rc == RT_OK; ++pf) {
        if (pf == 3 && uses_bvh(*b, i)) break;  // quads need the LBVH's heap indices
        d.switches = whitted_pass_form(desc->switches, pf);
        for (int form = 0; form < kForms && rc == RT_OK; ++form) {
            for (const rt_tuning &shape : kShapes) {
                rt_tuning t = shape;
                t.blocks_per_cu = saved.blocks_per_cu;
                t.traversal = form;
                t.quad_tiles = saved.quad_tiles;  // not used by Whitted frames; kept for direct ones
                ctx->tune = t;
                float ms = 1e30f;
                for (int rep = 0; rep < 3 && rc == RT_OK; ++rep) {  // first run warms caches
                    rt_frame_stats st;
                    rc = rt_render_whitted(ctx, scene, &d, 0, &st);
                    if (rep > 0 && st.trace_ms < ms) ms = st.trace_ms;
                }
                if (rc) break;
                if (ms < bms) {
                    bms = ms;
                    bt = t;
                    bpf = pf;
                }
            }
        }
    }
    ctx->tune = rc ? saved : bt;
    if (rc) return rc;
    desc->switches = whitted_pass_form(desc->switches, bpf);
    if (best) *best = bt;
    cache_add(cache_path, key, bt, bpf);
    return RT_OK;
}

int rt_graph_capture_frame(rt_ctx *ctx, int scene, int rebuild, const rt_job_desc *jobs,
                           uint32_t njobs, int *graph)
{
    Batch *b;
    uint32_t i;
    if (!get_scene(ctx, scene, &b, &i) || !graph || (njobs && !jobs))
        return fail(ctx, RT_ERR_ARG, "rt_graph_capture_frame: bad argument");
    for (uint32_t k = 0; k < njobs; ++k) {
        Batch *bk;
        uint32_t ik;
        if (!get_scene(ctx, jobs[k].scene, &bk, &ik) || bk != b)
            return fail(ctx, RT_ERR_ARG, "rt_graph_capture_frame: job %u is not in the batch of scene %d", k, scene);
    }
    hipSetDevice(ctx->device);
    if (!b->built) {  // the job table needs built scenes; build once outside the graph
        RT_HIP(ctx, rt::launch_build(*b, ctx->stream, nullptr, ctx->cu_count));
        b->built = true;
    }
    rt_ctx::Graph g;
    const size_t jb = rt::render_job_bytes();
    RT_HIP(ctx, hipMalloc(&g.jobs, jb * (njobs ? njobs : 1)));
    RT_HIP(ctx, hipMalloc(&g.counter, 256));
    RT_HIP(ctx, hipMemset(g.counter, 0, 256));
    uint32_t total = 0, depth = 0;
    int rc = stage_jobs(ctx, jobs, njobs, ctx->tune, g.jobs, nullptr, &total, &depth);
    if (rc) {
        hipFree(g.jobs);
        hipFree(g.counter);
        return rc;
    }
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hipStream_t st = ctx->stream;
    RT_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    if (rebuild) e = rt::launch_build(*b, st, nullptr, ctx->cu_count);
    // g.counter[0] = replay index (read by the kernel), g.counter[16] = tile queue
    if (e == hipSuccess)
        e = rt::launch_render_jobs(g.jobs, njobs, total, 0, total, true, g.counter + 16, g.counter,
                                   ctx->stats, ctx->tune, depth, ctx->cu_count, st);
    if (e == hipSuccess) e = rt::launch_bump_counter(g.counter, st);
    hipError_t e2 = hipStreamEndCapture(st, &g.graph);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (g.graph) hipGraphDestroy(g.graph);
        hipFree(g.jobs);
        hipFree(g.counter);
        return hip_fail(ctx, e, "rt_graph_capture_frame");
    }
    g.batch = b;
    g.generation = b->generation;
    b->refs++;
    *graph = (int)ctx->graphs.size();
    ctx->graphs.push_back(g);
    return RT_OK;
}

int rt_graph_launch(rt_ctx *ctx, int graph)
{
    if (!ctx || graph < 0 || graph >= (int)ctx->graphs.size() || !ctx->graphs[graph].exec)
        return fail(ctx, RT_ERR_ARG, "rt_graph_launch: unknown graph %d", graph);
    const rt_ctx::Graph &gr = ctx->graphs[graph];
    if (gr.batch->generation != gr.generation)
        return fail(ctx, RT_ERR_STATE,
                    "rt_graph_launch: graph %d is stale (a scene of its batch was destroyed, a BVH built or "
                    "dropped, or the structure switched since capture); capture it again", graph);
    hipSetDevice(ctx->device);
    hipEvent_t *ev = ctx->profiling ? prof_events(ctx, 1, 2) : nullptr;
    if (ev) hipEventRecord(ev[0], ctx->stream);
    RT_HIP(ctx, hipGraphLaunch(ctx->graphs[graph].exec, ctx->stream));
    if (ev) hipEventRecord(ev[1], ctx->stream);
    return RT_OK;
}

int rt_graph_destroy(rt_ctx *ctx, int graph)
{
    if (!ctx || graph < 0 || graph >= (int)ctx->graphs.size() || !ctx->graphs[graph].exec)
        return fail(ctx, RT_ERR_ARG, "rt_graph_destroy: unknown graph %d", graph);
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    rt_ctx::Graph &g = ctx->graphs[graph];
    hipGraphExecDestroy(g.exec);
    hipGraphDestroy(g.graph);
    hipFree(g.jobs);
    hipFree(g.counter);
    if (g.batch && --g.batch->refs == 0) batch_free(g.batch);
    g = rt_ctx::Graph();
    return RT_OK;
}

int rt_device_alloc(rt_ctx *ctx, uint64_t bytes, void **ptr)
{
    if (!ctx || !ptr) return RT_ERR_ARG;
    hipSetDevice(ctx->device);
    RT_HIP(ctx, hipMalloc(ptr, bytes ? bytes : 16));
    return RT_OK;
}

int rt_device_free(rt_ctx *ctx, void *ptr)
{
    if (!ctx) return RT_ERR_ARG;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    RT_HIP(ctx, hipFree(ptr));
    return RT_OK;
}

int rt_device_zero(rt_ctx *ctx, void *ptr, uint64_t bytes)
{
    if (!ctx || (bytes && !ptr)) return RT_ERR_ARG;
    hipSetDevice(ctx->device);
    RT_HIP(ctx, hipMemsetAsync(ptr, 0, bytes, ctx->stream));
    return RT_OK;
}

int rt_device_read(rt_ctx *ctx, void *host, const void *ptr, uint64_t bytes)
{
    if (!ctx || (bytes && (!host || !ptr))) return RT_ERR_ARG;
    hipSetDevice(ctx->device);
    RT_HIP(ctx, hipMemcpyAsync(host, ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream));
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
        if (r.kind == 0) {  // a build: its first and last events (stages are not recorded)
            float ms;
            hipEventElapsedTime(&ms, e[0], e[1]);
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

void rt_camera_from_world(rt_camera *cam, const float world[16])
{
    memset(cam, 0, sizeof(*cam));
    cam->kind = RT_CAMERA_WORLD;
    memcpy(cam->world, world, sizeof(cam->world));
    *cam = rt::device_camera(*cam);
}

void rt_camera_look_at(rt_camera *cam, const float pos[3], const float target[3], const float up[3],
                       float fov_y_deg, float aspect)
{
    memset(cam, 0, sizeof(*cam));
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
