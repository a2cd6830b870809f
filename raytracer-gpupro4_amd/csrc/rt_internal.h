// Internal types shared by the HIP kernels and the host runtime of rtcore.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "rtcore.h"

namespace rt {

// Sort geometry: 8-bit digits, four passes over the 30-bit Morton codes.
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortPasses = 4;
constexpr int kSortThreads = 256;
constexpr int kSortItems = 4;
constexpr int kSortTile = kSortThreads * kSortItems;

// Bounds / Morton kernels: 256 threads x 4 triangles per block.
constexpr int kBoundsThreads = 256;
constexpr int kBoundsItems = 4;
constexpr int kBoundsTile = kBoundsThreads * kBoundsItems;

// Leaf references in traversal nodes carry bit 31; the low bits are the first
// sorted primitive of the leaf.
constexpr uint32_t kLeafBit = 0x80000000u;

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

struct PrimBox {
    float mn[3];
    float mx[3];
};

// Device-side scene header written by the build.
struct SceneHdr {
    int32_t root_ref;     // internal index or kLeafBit | 0; -1 when empty
    int32_t pad0;
    float root_box[6];
    float scene_box[6];
    uint32_t nleaves;
    uint32_t pad1[3];
};

// Counters region (uint32): [0..31] frontier counts per level,
// [32] leaf count, [33..36] sort tile counters, [37..63] spare.
constexpr int kCntLeaves = 32;
constexpr int kCntTiles = 33;
constexpr int kCounters = 64;

struct Scene {
    uint32_t max_tris = 0, max_verts = 0, depth = 0;
    uint32_t nt = 0, nv = 0;
    bool built = false;
    bool have_rest = false;
    float *xyz = nullptr;       // current positions (3 floats per vertex)
    float *xyz_rest = nullptr;  // rest positions for rt_scene_transform
    uint32_t *idx = nullptr;
    PrimBox *pbox = nullptr;
    float *cent = nullptr;      // 3 floats per triangle (3x centroid)
    float *partials = nullptr;  // 6 floats per bounds block
    uint32_t *keys[2] = {nullptr, nullptr};
    int32_t *vals[2] = {nullptr, nullptr};
    uint32_t *hist = nullptr;      // kSortPasses * kRadix
    uint32_t *status = nullptr;    // kSortPasses * tiles * kRadix
    uint32_t *counters = nullptr;  // kCounters
    int4 *frontier[2] = {nullptr, nullptr};
    int4 *leaves = nullptr;
    rt_node32 *heap = nullptr;
    TravNode *trav = nullptr;
    int32_t *trav_parent = nullptr;
    uint32_t *flags = nullptr;
    TravTri *tris = nullptr;
    SceneHdr *hdr = nullptr;
    uint64_t bytes = 0;
    std::vector<void *> allocs;
};

struct FrameBuffers {
    uint32_t width = 0, height = 0, nlights = 0;
    float4 *rgba = nullptr;
    rt_hit24 *hits = nullptr;
    uint8_t *occ = nullptr;
    uint64_t bytes = 0;
};

}  // namespace rt

struct rt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<rt::Scene *> scenes;
    rt::FrameBuffers frame;
    unsigned long long *stats = nullptr;  // device counters: 8 x u64
    void *staging = nullptr;              // rays + hits for host-pointer rt_trace
    uint64_t staging_bytes = 0;
    hipEvent_t ev[8] = {};
    uint64_t bytes = 0;
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

// Host-side launchers (build.hip / trace.hip).
hipError_t launch_build(Scene &s, hipStream_t st, hipEvent_t *ev /* 6 or null */);
struct Xform {
    float m[12];
};
hipError_t launch_transform(Scene &s, const Xform &x, hipStream_t st);
hipError_t launch_trace(const Scene &s, const rt_ray48 *rays, uint32_t n, rt_hit24 *hits,
                        uint32_t flags, unsigned long long *stats, hipStream_t st);
hipError_t launch_render_direct(const Scene &s, const rt_render_desc &d, FrameBuffers &fb,
                                uint32_t flags, unsigned long long *stats, hipStream_t st);
uint32_t sort_tiles(uint32_t n);
uint32_t bounds_blocks(uint32_t n);

}  // namespace rt
