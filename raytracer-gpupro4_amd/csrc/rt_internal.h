// Internal types shared by the HIP kernels and the host runtime of rtcore.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "rtcore.h"

namespace rt {

// Sort geometry: 8-bit digits, four passes over the 30-bit Morton codes, plus
// one pass over the scene index when a batch holds more than one scene.
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kCodePasses = 4;
// Sort tiles: below kSortLargeN keys, look-back passes over tiles of 512
// threads x 8 keys (a tile's wait is a chain of look-back round trips over
// the earlier tiles, so fewer, larger tiles; each wave ranks 8 keys per lane
// either way: 256 threads measured 51.4 against 49.7 us for four passes at
// 262,144 keys, 768 threads 53.8, 1,024 threads 59.5); from kSortLargeN, count / scan /
// scatter passes over tiles of 256 threads x 32 keys (long per-digit runs,
// so the scatter writes coalesce).  Scratch is sized for the smaller tile.
constexpr int kSortThreads = 256;
constexpr int kSortItemsSmall = 8;
constexpr int kSortItemsLarge = 32;
constexpr uint32_t kSortLargeN = 1u << 20;
constexpr int kSortThreadsSmall = 512;
constexpr int kSortTile = kSortThreadsSmall * kSortItemsSmall;
constexpr int kMaxBatchScenes = kRadix;  // the scene pass is one 8-bit digit

// Bounds / Morton kernels: 256 threads x 4 triangles per block.
constexpr int kBoundsThreads = 256;
constexpr int kBoundsItems = 4;
constexpr int kBoundsTile = kBoundsThreads * kBoundsItems;

// Emit: nodes with more than kSubtreeMax primitives are emitted by one block
// per scene (k_top); every smaller subtree is emitted and refitted inside one
// block's LDS (k_subtree).
constexpr int kSubtreeMax = 512;
constexpr int kSubtreeThreads = 256;
constexpr int kTopThreads = 1024;
constexpr int kTopLds = 1024;  // k_top: frontier entries of a level kept in LDS
// A batch of one scene from this many primitives emits and refits its top
// levels with one launch per level over the whole chip (k_top_level,
// k_top_refit_level) instead of one block.
constexpr uint32_t kTopWideMinN = 1u << 22;
constexpr int kTopLevelThreads = 256;
constexpr int kMaxLevels = 32;

// Leaf references in traversal nodes carry bit 31; bits 28-30 hold the leaf's
// triangle count capped at 7 (so a traversal can fetch a leaf's triangles
// without first reading the count); the low 28 bits are the first sorted
// primitive of the leaf (a position in the batch's sorted arrays, below 2^25).
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafPosMask = 0x0FFFFFFFu;
constexpr int kLeafCountShift = 28;
constexpr uint32_t kLeafCountCap = 7u;
__host__ __device__ constexpr uint32_t leaf_ref(uint32_t first, uint32_t count)
{
    return kLeafBit | ((count < kLeafCountCap ? count : kLeafCountCap) << kLeafCountShift) | first;
}

constexpr int32_t kInvalid = -2147483647;

// Traversal node: one per internal node of the LBVH, indexed by its split
// position k (the last sorted primitive of its left child), which is unique
// among internal nodes.  It holds both children's boxes so that one 64-byte
// load tests both; refs are internal indices or kLeafBit | first primitive.
struct alignas(64) TravNode {
    float box[12];  // child 0: lo xyz, hi xyz; child 1: lo xyz, hi xyz
    int32_t ref[2];
    int32_t axis;   // split axis: the child on side dirIsNeg[axis] is visited first
    int32_t heap;   // heap index of this node
};
static_assert(sizeof(TravNode) == 64, "TravNode must be 64 bytes");

// Triangle in sorted order: A.xyz + id, E1.xyz + (leaf prim count at the
// leaf's first triangle), E2.xyz + (leaf heap index at the leaf's first
// triangle).  E1 = B - A and E2 = C - A, exactly as the intersection test
// computes them.
struct alignas(16) TravTri {
    float4 a;
    float4 e1;
    float4 e2;
};
static_assert(sizeof(TravTri) == 48, "TravTri must be 48 bytes");

// Triangle in input order, written by the bounds pass where it reads the
// vertices anyway: A.xyz + scene-local id, B.xyz, C.xyz.  The emit pass
// gathers one record per sorted triangle for its box and traversal record.
struct alignas(16) TriRec {
    float4 a;
    float4 b;
    float4 c;
};
static_assert(sizeof(TriRec) == 48, "TriRec must be 48 bytes");

// Per scene of a batch, device side.
// Input arrays (vertices, indices) are laid out by capacity; build arrays
// (boxes, keys, nodes, triangles) by the current triangle counts.
struct SceneDesc {
    uint32_t prim_off;     // first triangle / sorted position in the build arrays
    uint32_t nt;
    uint32_t in_prim_off;  // first triangle in the index array
    uint32_t vert_off;     // first vertex in the vertex array
    uint32_t nv;
    uint32_t depth;        // heap depth
    uint32_t tbits;        // bits of the split table
    uint32_t block_off;    // first bounds block
    uint32_t nblocks;
    uint32_t pad;
    uint64_t heap_off;     // first heap node
    uint64_t table_off;    // first split-table entry
};

// Per scene, written by the build.
struct SceneHdr {
    int32_t root_ref;  // internal index or kLeafBit | first; -1 when empty
    uint32_t levels;   // number of big (k_top) levels
    float root_box[6];
    float scene_box[6];
    uint32_t pad[2];
};

// Counters region (uint32): [0] small-subtree count, [1] subtree work
// counter, [2..2+kCodePasses] sort tile counters, [8..8+kMaxLevels*256) big
// internal nodes per level per scene are kept in top_counts instead.
constexpr int kCntSmall = 0;
constexpr int kCntWork = 1;
constexpr int kCntTiles = 2;
constexpr int kCntFront = 8;  // frontier size of each top level (k_top_level)
constexpr int kCounters = kCntFront + 40;

// SAH BVH of one scene as built on the host (LinearBVHNode order) plus its
// depth; see bvh.cpp.
struct BvhHost {
    std::vector<rt_node32> nodes;
    std::vector<int32_t> ordered;  // orderedPrimitivesIds
    uint32_t max_level = 0;        // deepest leaf, root = 0
};

// A batch: scenes sharing one set of concatenated device arrays, built by one
// sequence of launches.  A plain scene is a batch of one.
struct Batch {
    uint32_t nscenes = 0;
    std::vector<SceneDesc> desc;      // host copy
    std::vector<uint32_t> max_tris, max_verts;
    uint32_t total_tris = 0, total_verts = 0, total_blocks = 0;
    uint64_t total_heap = 0, total_table = 0;
    uint32_t max_depth = 0;
    bool built = false;
    std::vector<uint8_t> rest;       // per scene: rest positions captured
    float *xyz = nullptr;       // current positions (3 floats per vertex)
    float *xyz_rest = nullptr;  // rest positions for rt_scene_transform
    uint32_t *idx = nullptr;    // per-scene vertex indices (3 per triangle)
    SceneDesc *d_desc = nullptr;
    SceneHdr *hdr = nullptr;
    TriRec *trec = nullptr;
    float *cent = nullptr;      // 3 floats per triangle (3x centroid)
    float *partials = nullptr;  // 6 floats per bounds block
    uint32_t *keys[2] = {nullptr, nullptr};
    int32_t *vals[2] = {nullptr, nullptr};
    uint32_t *hist = nullptr;      // kCodePasses * kRadix
    uint32_t *counters = nullptr;  // kCounters
    uint32_t *status = nullptr;    // passes * tiles * kRadix
    uint32_t *table = nullptr;     // split tables
    int4 *big[2] = {nullptr, nullptr};
    int4 *small = nullptr;
    int4 *top_nodes = nullptr;     // big internal nodes, per scene region, by level
    uint32_t *top_counts = nullptr;  // per scene, kMaxLevels level sizes
    rt_node32 *heap = nullptr;
    TravNode *trav = nullptr;
    TravTri *tris = nullptr;
    uint64_t bytes = 0;
    std::vector<void *> allocs;
    int refs = 0;                  // scene handles and captured graphs pointing here
    // Bumped whenever what a captured graph baked in goes stale: a scene
    // handle destroyed, a BVH (re)built or dropped, the selected structure
    // switched.  rt_graph_launch refuses a graph captured under another value.
    uint64_t generation = 0;
    // Shading attributes (rt_scene_shading), allocated at their first upload:
    // normals in the vertex layout of xyz, material ids in the triangle
    // layout of idx, and one material table per scene.
    float *nrm = nullptr;
    uint32_t *mat = nullptr;
    struct Shading {
        rt_material *mats = nullptr;  // device; null: the default MaterialData
        uint32_t nmat = 0, nv = 0, nt = 0;
        bool normals = false, ids = false;
    };
    std::vector<Shading> shading;  // per scene
    // Second structure (rt_bvh_build): per scene, the host tree and its
    // device layout; use_bvh selects it for every traversal of the scene.
    struct Bvh {
        BvhHost host;
        TravNode *trav = nullptr;
        TravTri *tris = nullptr;
        SceneHdr *hdr = nullptr;
        uint64_t bytes = 0;
        bool valid = false;
    };
    std::vector<Bvh> bvh;
    std::vector<uint8_t> use_bvh;
};

inline bool uses_bvh(const Batch &b, uint32_t s) { return s < b.use_bvh.size() && b.use_bvh[s]; }
// The scene can be traversed: its selected structure has been built.
inline bool traversable(const Batch &b, uint32_t s) { return uses_bvh(b, s) || b.built; }
// Depth of the selected structure (levels of the deepest leaf + 1): a
// traversal stack holds at most depth - 1 entries.
inline uint32_t trav_depth(const Batch &b, uint32_t s)
{
    return uses_bvh(b, s) ? b.bvh[s].host.max_level + 1 : b.desc[s].depth;
}

struct SceneRef {
    Batch *batch = nullptr;
    uint32_t index = 0;
};

struct FrameBuffers {
    uint32_t width = 0, height = 0, nlights = 0;
    float4 *rgba = nullptr;
    rt_hit24 *hits = nullptr;
    uint8_t *occ = nullptr;
    // Tile orders of the frame kernels (the heaviest tiles of the last frame
    // of a view first) and the per-tile costs they are made from, one slot
    // per view: a slot is sized for the smallest tile (16 x 4, block 64) and
    // holds one (scene, camera, kernel, block size, tile count); a new view
    // takes the least recently used slot.
    static constexpr int kOrderSlots = 8;
    struct OrderSlot {
        uint32_t *order = nullptr, *cost = nullptr;
        bool used = false, valid = false;  // valid: order written by a frame of this view
        uint64_t key = 0, tick = 0;
        uint32_t block = 0, tiles = 0;
    };
    OrderSlot slots[kOrderSlots];
    uint32_t *order_mem = nullptr;  // the slots' orders and costs
    uint64_t order_tick = 0;
    uint64_t bytes = 0;
};

}  // namespace rt

struct rt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<rt::SceneRef> scenes;
    rt::FrameBuffers frame;
    unsigned long long *stats = nullptr;  // device counters: 8 x u64
    void *staging = nullptr;              // rays + hits for host-pointer rt_trace
    uint64_t staging_bytes = 0;
    hipEvent_t ev[8] = {};
    uint64_t bytes = 0;
    int cu_count = 256;
    rt_tuning tune{128, 26, 8, 0};
    uint32_t *tile_counter = nullptr;     // persistent job launches
    void *wqueue = nullptr;               // Whitted wavefront ray queues
    uint64_t wqueue_bytes = 0;
    uint32_t *wcount = nullptr;           // Whitted queue lengths, one per pass (17)
    void *jobs_dev = nullptr;             // render-job table
    uint64_t jobs_cap = 0;
    hipEvent_t jobs_ev = nullptr;         // last launch that read jobs_dev
    // rt_render_jobs_noalloc: this device's static scratch array (job table,
    // tile counter, ray counters; set up by rt_create), and host staging of
    // RT_NOALLOC_MAX_JOBS job records and tile offsets, sized at rt_create
    unsigned char *noalloc_scratch = nullptr;
    std::vector<unsigned char> noalloc_host;
    std::vector<uint32_t> noalloc_offs;
    hipEvent_t noalloc_ev = nullptr;      // the last copy out of noalloc_host
    // RT_TILES_CENTRE orders on the device, by (width, height, tile rows)
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, uint32_t *> tile_orders;
    struct Graph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        void *jobs = nullptr;
        uint32_t *counter = nullptr;      // frame counter read by the job kernel
        rt::Batch *batch = nullptr;       // held (refs) until rt_graph_destroy
        uint64_t generation = 0;          // batch->generation at capture
    };
    std::vector<Graph> graphs;
    // profiling window: events come from a pool and are read back at the end
    bool profiling = false;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    struct ProfRec {
        int kind;  // 0 = build (6 events), 1 = trace/render (2 events)
        size_t first;
    };
    std::vector<ProfRec> recs;
    uint64_t prof_primary = 0;
};

namespace rt {

struct Xform {
    float m[12];
};

// Host-side launchers (build.hip / trace.hip).
hipError_t launch_build(Batch &b, hipStream_t st, hipEvent_t *ev /* 6 or null */, int cu_count);
hipError_t launch_transform(Batch &b, uint32_t scene, const Xform &x, hipStream_t st);
hipError_t launch_trace(const Batch &b, uint32_t scene, const rt_ray48 *rays, uint32_t n,
                        rt_hit24 *hits, uint32_t flags, unsigned long long *stats,
                        const rt_tuning &tune, hipStream_t st);
// The camera as the kernels read it: kind checked by camera_ok; for
// RT_CAMERA_WORLD, pos becomes the matrix's origin (M (0,0,0,1)).xyz / w,
// each row summed left to right (the kernels' primary_ray subtracts it).
bool camera_ok(const rt_camera &c);
rt_camera device_camera(const rt_camera &c);
hipError_t launch_render_direct(const Batch &b, uint32_t scene, const rt_render_desc &d,
                                FrameBuffers &fb, uint32_t flags, unsigned long long *stats,
                                const rt_tuning &tune, hipStream_t st);
// queue: whitted_queue_bytes(width * height) of device memory (two ray
// queues); qcount: bounces + 1 words.  Unused by the one-pass form.
hipError_t launch_render_whitted(const Batch &b, uint32_t scene, const rt_whitted_desc &d,
                                 FrameBuffers &fb, uint32_t flags, unsigned long long *stats,
                                 const rt_tuning &tune, void *queue, uint32_t *qcount, int cu_count,
                                 hipStream_t st);
size_t whitted_queue_bytes(uint64_t pixels);
size_t render_job_bytes();
// order (device, optional): the tile at each queue position of the job.
void fill_render_job(void *dst, const Batch &b, uint32_t scene, const rt_job_desc &d,
                     const rt_tuning &t, uint32_t tile_off, uint32_t *tiles, const uint32_t *order);
// count: one lane per ray, box and triangle tests counted into stats; without
// it, LBVH jobs run with lane groups and only ray counts are kept.
hipError_t launch_render_jobs(const void *d_jobs, uint32_t njobs, uint32_t total_tiles,
                              uint32_t first_tile, uint32_t tile_count, bool persistent,
                              uint32_t *tile_counter, const uint32_t *frame_counter,
                              unsigned long long *stats, const rt_tuning &tune, uint32_t max_depth,
                              int cu_count,
                              hipStream_t st, bool count = false);
hipError_t launch_bump_counter(uint32_t *c, hipStream_t st);
hipError_t launch_zero_words(uint32_t *p, uint32_t n, hipStream_t st);
// Times the launch shapes of one block size on the jobs (persistent
// launches into scratch accumulators); *best = the fastest.  The context's
// tuning is left as it is.
int search_jobs_shape_block(rt_ctx *ctx, const rt_job_desc *jobs, uint32_t njobs, int block, rt_tuning *best);
// RT_TILES_CENTRE: the tiles (16 x tile_h, row-major indices) of a width x
// height image by distance of their centre from the image's centre, ties by index.
void centre_tile_order(uint32_t width, uint32_t height, uint32_t tile_h, std::vector<uint32_t> &order);
// Upper bound of render_job_bytes() (checked where the record is defined).
constexpr size_t kRenderJobBytesMax = 512;

// SAH BVH (bvh.cpp).
int bvh_build_host(const float *xyz, const uint32_t *idx, uint32_t nt, uint32_t max_prims, uint32_t split,
                   BvhHost &out, rt_bvh_info *info);
void bvh_device_layout(const float *xyz, const uint32_t *idx, const BvhHost &h, std::vector<TravNode> &trav,
                       std::vector<TravTri> &tris, SceneHdr &hdr);
int tune_index(const rt_tuning &t);
uint32_t sort_tiles(uint32_t n);
uint32_t bounds_blocks(uint32_t n);
uint32_t table_bits(uint32_t max_tris);
uint64_t scene_bytes(uint32_t max_tris, uint32_t max_verts, uint32_t depth);

}  // namespace rt
