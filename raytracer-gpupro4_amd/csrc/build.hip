// LBVH build on gfx950 for a batch of scenes: primitive bounds, 30-bit Morton
// codes, onesweep radix sort, heap emission and refit.
//
// Reference stages (StructureConstruction::GPUBuildStructure,
// RayTracerCS/StructureConstruction.cpp:86-131):
//   LBVH_PreBuild   (LBVH.hlsl:15-132)       -> k_bounds + k_morton
//   BitonicSort     (BitonicSort.cpp:56-117) -> k_sort_scatter x 4 (+1 scene pass),
//                                               with k_sort_count + k_sort_scan from 1M keys
//   LBVH_Build      (LBVH.hlsl:134-226)      -> k_split_table + k_top + k_subtree
//   LBVH_BuildAABB  (LBVH.hlsl:228-270)      -> k_subtree + k_top_refit
// Node words follow the CPU builder (LBVH.cpp:190-314), including its axis
// rule; float arithmetic is IEEE single with no contraction (-ffp-contract=off).
//
// Emission splits the tree by size.  k_top (one block per scene) emits the
// nodes holding more than kSubtreeMax primitives level by level, finding each
// split through a bucket table over the sorted codes.  Every smaller subtree
// is emitted and refitted by one block of k_subtree entirely in LDS, with
// only block barriers.  k_top_refit (one block per scene) then merges the big
// nodes bottom-up.  No kernel needs device-scope fences.
#include "rt_internal.h"

#include <cmath>

namespace rt {

// ---------------------------------------------------------------------------
// Wave / block helpers (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Reduces lo[3]/hi[3] over a block of NW waves; the result is valid in every
// thread.  red must hold NW x 6 floats.
template <int NW>
__device__ __forceinline__ void block_box_reduce(float lo[3], float hi[3], float *red)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = wave_min(lo[k]);
        hi[k] = wave_max(hi[k]);
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            red[wave * 6 + k] = lo[k];
            red[wave * 6 + 3 + k] = hi[k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float a = red[k], b = red[3 + k];
        for (int w = 1; w < NW; ++w) {
            a = fminf(a, red[w * 6 + k]);
            b = fmaxf(b, red[w * 6 + 3 + k]);
        }
        lo[k] = a;
        hi[k] = b;
    }
    __syncthreads();
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *tmp, uint32_t *total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) tmp[wave] = x;
    __syncthreads();
    uint32_t wbase = 0, all = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint32_t t = tmp[w];
        if (w < wave) wbase += t;
        all += t;
    }
    __syncthreads();
    if (total) *total = all;
    return wbase + x - v;
}

// block_exclusive_scan over a block of NW waves.
template <int NW>
__device__ __forceinline__ uint32_t block_exclusive_scan_w(uint32_t v, uint32_t *tmp, uint32_t *total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) tmp[wave] = x;
    __syncthreads();
    uint32_t wbase = 0, all = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        uint32_t t = tmp[w];
        if (w < wave) wbase += t;
        all += t;
    }
    __syncthreads();
    if (total) *total = all;
    return wbase + x - v;
}

__device__ __forceinline__ uint32_t part1by2(uint32_t n)
{
    n &= 1023u;
    n = (n ^ (n << 16)) & 0xff0000ffu;
    n = (n ^ (n << 8)) & 0x0300f00fu;
    n = (n ^ (n << 4)) & 0x030c30c3u;
    n = (n ^ (n << 2)) & 0x09249249u;
    return n;
}

__device__ __forceinline__ uint32_t quantise(float c, float m, float f)
{
    const float s = (c - m) * f;
    if (!(s == s)) return 0u;  // flat axis: 0 * inf
    return (uint32_t)(int)s;
}

// Last scene s with key(s) <= x over S ascending keys read through get().
template <typename F>
__device__ __forceinline__ uint32_t upper_scene(uint32_t S, uint32_t x, F get)
{
    uint32_t lo = 0, hi = S;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (get(mid) <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t scene_of_prim(const SceneDesc *d, uint32_t S, uint32_t g)
{
    return upper_scene(S, g, [&](uint32_t s) { return d[s].prim_off; });
}

// Leaf rule and axis bits of build_recursive (LBVH.cpp:194-198, 231-266,
// 304-310) for a node of cnt primitives at heap index h entered under a split
// at bit pp (31 for the root): returns 1 for a leaf and sets *axis; returns 0
// for an internal node and sets *split_bit.
__device__ __forceinline__ int classify(uint32_t h, int32_t cnt, int pp, uint32_t leaf_level,
                                        uint32_t code_first, uint32_t code_last, int *axis,
                                        int *split_bit)
{
    if (cnt <= 1 || h >= leaf_level || pp == 0) {
        *axis = (31 - pp) % 3;
        return 1;
    }
    const uint32_t x = code_first ^ code_last;
    if (x == 0u) {
        *axis = 1;
        return 1;
    }
    *split_bit = 31 - __clz(x);
    *axis = (30 - *split_bit) % 3;
    return 0;
}

__device__ __forceinline__ void write_heap(rt_node32 *n, const float lo[3], const float hi[3],
                                           int32_t prim_count, int32_t prim_pos)
{
    float4 *p = reinterpret_cast<float4 *>(n);
    p[0] = make_float4(lo[0], lo[1], lo[2], hi[0]);
    p[1] = make_float4(hi[1], hi[2], __int_as_float(prim_count), __int_as_float(prim_pos));
}

__device__ __forceinline__ void write_heap_word(rt_node32 *n, int32_t prim_count, int32_t prim_pos)
{
    *reinterpret_cast<int2 *>(&n->prim_count) = make_int2(prim_count, prim_pos);
}

// Box and sorted-order triangle record of one leaf primitive, from its
// triangle record T (trec[vals[pos]]).
__device__ __forceinline__ void leaf_prim_rec(const TriRec &T, TravTri *__restrict__ tris, uint32_t pos, int32_t j,
                                              int32_t cnt, uint32_t h, float lo[3], float hi[3])
{
    const float A[3] = {T.a.x, T.a.y, T.a.z}, B[3] = {T.b.x, T.b.y, T.b.z}, C[3] = {T.c.x, T.c.y, T.c.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {  // the triangle's box, as the bounds pass computed it
        lo[k] = fminf(lo[k], fminf(A[k], fminf(B[k], C[k])));
        hi[k] = fmaxf(hi[k], fmaxf(A[k], fmaxf(B[k], C[k])));
    }
    TravTri t;
    t.a = T.a;
    t.e1 = make_float4(B[0] - A[0], B[1] - A[1], B[2] - A[2], __int_as_float(j == 0 ? cnt : 0));
    t.e2 = make_float4(C[0] - A[0], C[1] - A[1], C[2] - A[2], __int_as_float(j == 0 ? (int32_t)h : 0));
    tris[pos] = t;
}

__device__ __forceinline__ void leaf_prim(const int32_t *__restrict__ vals, const TriRec *__restrict__ trec,
                                          TravTri *__restrict__ tris, uint32_t pos, int32_t j, int32_t cnt,
                                          uint32_t h, float lo[3], float hi[3])
{
    leaf_prim_rec(trec[vals[pos]], tris, pos, j, cnt, h, lo, hi);
}

// ---------------------------------------------------------------------------
// Stage 1: per-triangle box and 3x centroid (LBVH.hlsl:26-36, LBVH.cpp:501-516);
// per-block partial scene boxes; clears the sort/emit scratch words.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBoundsThreads) void k_bounds(
    const SceneDesc *__restrict__ desc, uint32_t S, uint32_t total_blocks,
    const float *__restrict__ xyz, const uint32_t *__restrict__ idx, TriRec *__restrict__ trec,
    float *__restrict__ cent, float *__restrict__ partials, uint32_t *__restrict__ zero_words,
    uint32_t nzero)
{
    __shared__ float red[24];
    for (uint32_t j = blockIdx.x * kBoundsThreads + threadIdx.x; j < nzero;
         j += gridDim.x * kBoundsThreads)
        zero_words[j] = 0u;
    if (blockIdx.x >= total_blocks) return;
    const uint32_t s = upper_scene(S, blockIdx.x, [&](uint32_t i) { return desc[i].block_off; });
    const SceneDesc D = desc[s];
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const uint32_t base = (blockIdx.x - D.block_off) * (uint32_t)kBoundsTile;
#pragma unroll
    for (int r = 0; r < kBoundsItems; ++r) {
        const uint32_t i = base + r * kBoundsThreads + threadIdx.x;
        if (i < D.nt) {
            const uint32_t *tri = idx + 3 * ((size_t)D.in_prim_off + i);
            const float *va = xyz + 3 * ((size_t)D.vert_off + tri[0]);
            const float *vb = xyz + 3 * ((size_t)D.vert_off + tri[1]);
            const float *vc = xyz + 3 * ((size_t)D.vert_off + tri[2]);
            const size_t g = (size_t)D.prim_off + i;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float a = va[k], bb = vb[k], c = vc[k];
                const float sum = a + bb;
                cent[3 * g + k] = sum + c;
                lo[k] = fminf(lo[k], fminf(a, fminf(bb, c)));
                hi[k] = fmaxf(hi[k], fmaxf(a, fmaxf(bb, c)));
            }
            TriRec t;
            t.a = make_float4(va[0], va[1], va[2], __int_as_float((int32_t)i));
            t.b = make_float4(vb[0], vb[1], vb[2], 0.0f);
            t.c = make_float4(vc[0], vc[1], vc[2], 0.0f);
            trec[g] = t;
        }
    }
    block_box_reduce<4>(lo, hi, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            partials[blockIdx.x * 6 + k] = lo[k];
            partials[blockIdx.x * 6 + 3 + k] = hi[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Stage 2: scene box from the scene's partials, Morton codes (LBVH.hlsl:92-114,
// LBVH.cpp:520-553), identity ids (batch-global triangle index), and the digit
// histograms of the four code passes.
// ---------------------------------------------------------------------------
// A scene's box from its k_bounds block partials.  Scenes of at most
// kBoxInline blocks let every k_morton block reduce the partials itself (a
// few KB each, no extra launch); larger ones would make that O(blocks^2), so
// one block of this kernel per scene reduces them once into the scene header.
constexpr uint32_t kBoxInline = 512;
// k_morton blocks per CU (each walks several bounds tiles)
constexpr uint32_t kMortonBlocksPerCu = 4;

__global__ __launch_bounds__(1024) void k_scene_box(const SceneDesc *__restrict__ desc,
                                                    const float *__restrict__ partials,
                                                    SceneHdr *__restrict__ hdr)
{
    __shared__ float red[16 * 6];
    const SceneDesc D = desc[blockIdx.x];
    if (D.nblocks <= kBoxInline) return;  // reduced inline by k_morton
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    constexpr uint32_t kUnroll = 8;
    for (uint32_t j0 = 0; j0 < D.nblocks; j0 += 1024 * kUnroll) {
        float v[kUnroll][6];
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {  // all loads of the batch before any use
            const uint32_t j = j0 + u * 1024 + threadIdx.x;
            const float *p = partials + 6 * ((size_t)D.block_off + (j < D.nblocks ? j : 0));
#pragma unroll
            for (int k = 0; k < 6; ++k) v[u][k] = p[k];
        }
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
            if (j0 + u * 1024 + threadIdx.x < D.nblocks) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    lo[k] = fminf(lo[k], v[u][k]);
                    hi[k] = fmaxf(hi[k], v[u][3 + k]);
                }
            }
        }
    }
    block_box_reduce<16>(lo, hi, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            hdr[blockIdx.x].scene_box[k] = lo[k];
            hdr[blockIdx.x].scene_box[3 + k] = hi[k];
        }
    }
}

// Each block walks the bounds tiles blockIdx.x, blockIdx.x + gridDim.x, ...
// and keeps its digit histograms in LDS across them, so the batch-wide
// histograms take one round of global atomics per block, not per tile (at
// 33.5M triangles: about 1M atomics on 1,024 words instead of 27M).
__global__ __launch_bounds__(kBoundsThreads) void k_morton(
    const SceneDesc *__restrict__ desc, uint32_t S, uint32_t total_blocks, const float *__restrict__ cent,
    const float *__restrict__ partials, uint32_t *__restrict__ keys, int32_t *__restrict__ vals,
    uint32_t *__restrict__ hist, SceneHdr *__restrict__ hdr)
{
    __shared__ float red[24];
    __shared__ uint32_t lhist[kCodePasses * kRadix];
    for (int j = threadIdx.x; j < kCodePasses * kRadix; j += kBoundsThreads) lhist[j] = 0;
    __syncthreads();  // lhist cleared before the atomics below
    uint32_t cur = 0xFFFFFFFFu;  // scene whose box m, f hold (block-uniform)
    float m[3] = {0.0f, 0.0f, 0.0f}, f[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t blk = blockIdx.x; blk < total_blocks; blk += gridDim.x) {
        const uint32_t s = upper_scene(S, blk, [&](uint32_t i) { return desc[i].block_off; });
        const SceneDesc D = desc[s];
        const uint32_t lb = blk - D.block_off;
        if (s != cur) {
            cur = s;
            float lo[3] = {INFINITY, INFINITY, INFINITY};
            float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            if (D.nblocks <= kBoxInline) {
                for (uint32_t j = threadIdx.x; j < D.nblocks; j += kBoundsThreads) {
                    const float *p = partials + 6 * ((size_t)D.block_off + j);
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        lo[k] = fminf(lo[k], p[k]);
                        hi[k] = fmaxf(hi[k], p[3 + k]);
                    }
                }
                block_box_reduce<4>(lo, hi, red);
                if (lb == 0 && threadIdx.x == 0) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        hdr[s].scene_box[k] = lo[k];
                        hdr[s].scene_box[3 + k] = hi[k];
                    }
                }
            } else {  // k_scene_box reduced them
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    lo[k] = hdr[s].scene_box[k];
                    hi[k] = hdr[s].scene_box[3 + k];
                }
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float hk = 3.0f * hi[k];
                const float lk = 3.0f * lo[k];
                const float ext = hk - lk;
                m[k] = lk;
                f[k] = 1024.0f / ext;
            }
        }
        const uint32_t base = lb * (uint32_t)kBoundsTile;
#pragma unroll
        for (int r = 0; r < kBoundsItems; ++r) {
            const uint32_t i = base + r * kBoundsThreads + threadIdx.x;
            if (i < D.nt) {
                const size_t g = (size_t)D.prim_off + i;
                const uint32_t qx = quantise(cent[3 * g + 0], m[0], f[0]);
                const uint32_t qy = quantise(cent[3 * g + 1], m[1], f[1]);
                const uint32_t qz = quantise(cent[3 * g + 2], m[2], f[2]);
                const uint32_t code = (part1by2(qx) << 2) | (part1by2(qy) << 1) | part1by2(qz);
                keys[g] = code;
                vals[g] = (int32_t)g;
#pragma unroll
                for (int p = 0; p < kCodePasses; ++p)
                    atomicAdd(&lhist[p * kRadix + ((code >> (p * kRadixBits)) & (kRadix - 1))], 1u);
            }
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kCodePasses * kRadix; j += kBoundsThreads)
        if (lhist[j]) atomicAdd(&hist[j], lhist[j]);
}

// ---------------------------------------------------------------------------
// Stage 3: stable LSD passes of 8-bit digits.  Keys inside a tile are ranked
// in load order (wave, item, lane), which keeps each pass stable; ids start as
// the batch-global index, so each scene ends in the (code, id) order of
// cmpMorton (LBVH.cpp:21-38).  The scene pass (digit = scene of the id, global
// base = the scene's first sorted position) groups the scenes of a batch after
// the four code passes.
//
// Two forms of a pass:
//  - below kSortLargeN keys, onesweep: one launch; tiles are claimed in launch
//    order through an atomic counter and take their global offsets by a
//    decoupled look-back over earlier tiles (so a tile only waits on tiles
//    that are already running);
//  - from kSortLargeN keys, count / scan / scatter: k_sort_count writes every
//    tile's digit counts, k_sort_scan turns them into each tile's global
//    offsets, and k_sort_scatter places the keys.  4 more bytes read per key,
//    but no tile waits on another: at that size the look-back chain, not the
//    bytes, bounded the onesweep form.
// ---------------------------------------------------------------------------
constexpr int kLookBack = 4;  // status words a look-back step loads at once
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1u;

// The digit of a key of one pass: 8 bits of the code, or the scene of the id.
template <bool SCENE_PASS>
__device__ __forceinline__ uint32_t sort_digit(uint32_t key, int32_t val, uint32_t shift, uint32_t S,
                                               const uint32_t *s_soff)
{
    if (SCENE_PASS) return upper_scene(S, (uint32_t)val, [&](uint32_t i) { return s_soff[i]; });
    return (key >> shift) & (kRadix - 1);
}

template <bool SCENE_PASS, int ITEMS, bool LOOKBACK, int THREADS = kSortThreads>
__global__ __launch_bounds__(THREADS) void k_sort_scatter(
    const uint32_t *__restrict__ kin, const int32_t *__restrict__ vin, uint32_t *__restrict__ kout,
    int32_t *__restrict__ vout, uint32_t n, uint32_t shift, const uint32_t *__restrict__ hist,
    const SceneDesc *__restrict__ desc, uint32_t S, uint32_t *__restrict__ status,
    uint32_t *__restrict__ tile_counter, const uint32_t *__restrict__ offs, uint32_t ntiles)
{
    static_assert(THREADS >= kRadix, "a thread per digit");
    constexpr int kWaves = THREADS / 64;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave[kWaves][kRadix];
    __shared__ uint32_t s_excl[kRadix];
    __shared__ uint32_t s_gofs[kRadix];
    __shared__ uint32_t s_scan[kWaves];
    __shared__ uint32_t s_keys[THREADS * ITEMS];
    __shared__ int32_t s_vals[THREADS * ITEMS];
    __shared__ uint32_t s_soff[SCENE_PASS ? kRadix : 1];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (LOOKBACK && tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (int j = tid; j < kWaves * kRadix; j += THREADS) (&s_wave[0][0])[j] = 0u;
    if (SCENE_PASS)
        for (int j = tid; j < kRadix; j += THREADS) s_soff[j] = j < (int)S ? desc[j].prim_off : 0xFFFFFFFFu;
    __syncthreads();
    const uint32_t tile = LOOKBACK ? s_tile : blockIdx.x;
    const uint32_t base = tile * (uint32_t)(THREADS * ITEMS);

    uint32_t key[ITEMS];
    int32_t val[ITEMS];
    uint32_t rank[ITEMS];
    uint32_t dig[SCENE_PASS ? ITEMS : 1];  // the scene pass keeps its digits (a search each)
    bool valid[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const uint32_t i = base + wave * (64 * ITEMS) + r * 64 + lane;
        valid[r] = i < n;
        const uint32_t ic = valid[r] ? i : n - 1;  // past n: the last key again, then masked off
        key[r] = kin[ic];
        val[r] = vin[ic];
        if (SCENE_PASS) dig[SCENE_PASS ? r : 0] = valid[r] ? sort_digit<SCENE_PASS>(key[r], val[r], shift, S, s_soff) : 0u;
    }
    auto digit = [&](int r) -> uint32_t {
        if (SCENE_PASS) return dig[SCENE_PASS ? r : 0];
        return (key[r] >> shift) & (kRadix - 1);
    };
    // Rank: the lanes of a wave holding item r's digit are matched by eight
    // ballots; the lowest of them adds their number to the wave's counter of
    // that digit.  (Issuing the items' counter updates as LDS atomics without
    // this read-modify-write chain, the old values broadcast afterwards,
    // measured slower: 179 against 143 us per pass at 33.5M keys.)
    const uint64_t lt_mask = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const uint32_t d = digit(r);
        uint64_t peers = __ballot(valid[r]);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint64_t before_me = peers & lt_mask;
        const uint32_t prior = s_wave[wave][d];
        if (valid[r] && before_me == 0) s_wave[wave][d] = prior + (uint32_t)__popcll(peers);
        rank[r] = prior + (uint32_t)__popcll(before_me);
    }
    __syncthreads();

    // Thread t < 256 owns digit t: per-wave exclusive offsets, tile count,
    // tile scan (the scans run on every thread; the others add nothing).
    const uint32_t dg = (uint32_t)tid;
    const bool owner = tid < kRadix;
    uint32_t cnt = 0;
    if (owner) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = s_wave[w][dg];
            s_wave[w][dg] = cnt;
            cnt += c;
        }
    }
    uint32_t tile_total;
    const uint32_t excl = block_exclusive_scan_w<kWaves>(cnt, s_scan, &tile_total);
    if (owner) s_excl[dg] = excl;

    if (LOOKBACK) {
        // Decoupled look-back over earlier tiles for this digit.
        uint32_t *my = status + (size_t)tile * kRadix + dg;
        uint32_t prefix = 0;
        if (!owner) {
        } else if (tile == 0) {
            __hip_atomic_store(my, kFlagInc | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(my, kFlagAgg | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // Look back kLookBack tiles per round trip: their status words are
            // loaded together, then taken newest first until an inclusive
            // prefix (tile 0 always publishes one) or a tile not ready yet,
            // which is loaded again from there.
            int32_t j = (int32_t)tile - 1;
            while (j >= 0) {
                uint32_t sv[kLookBack];
#pragma unroll
                for (int u = 0; u < kLookBack; ++u)
                    sv[u] = j - u >= 0 ? __hip_atomic_load(status + (size_t)(j - u) * kRadix + dg, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT)
                                       : 0u;
                int taken = 0;
                bool done = false;
#pragma unroll
                for (int u = 0; u < kLookBack; ++u) {
                    if (j - u < 0) break;
                    const uint32_t flag = sv[u] & ~kValMask;
                    if (flag == 0u) break;
                    prefix += sv[u] & kValMask;
                    taken = u + 1;
                    if (flag == kFlagInc) {
                        done = true;
                        break;
                    }
                }
                if (done) break;
                if (taken == 0) __builtin_amdgcn_s_sleep(1);
                j -= taken;
            }
            __hip_atomic_store(my, kFlagInc | (prefix + cnt), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        uint32_t gbase;
        if (SCENE_PASS) {
            gbase = owner && dg < S ? s_soff[dg] : 0u;
        } else {
            gbase = block_exclusive_scan_w<kWaves>(owner ? hist[dg] : 0u, s_scan, nullptr);
        }
        if (owner) s_gofs[dg] = gbase + prefix;
    } else if (owner) {
        s_gofs[dg] = offs[(size_t)dg * ntiles + tile];  // k_sort_scan: global base + earlier tiles
    }
    __syncthreads();

#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        if (valid[r]) {
            const uint32_t d = digit(r);
            const uint32_t lp = s_excl[d] + s_wave[wave][d] + rank[r];
            s_keys[lp] = key[r];
            s_vals[lp] = val[r];
        }
    }
    __syncthreads();
    const uint32_t tile_n = (n - base) < (uint32_t)(THREADS * ITEMS) ? (n - base) : (uint32_t)(THREADS * ITEMS);
    for (uint32_t i = tid; i < tile_n; i += THREADS) {
        const uint32_t k = s_keys[i];
        const int32_t v = s_vals[i];
        const uint32_t d = sort_digit<SCENE_PASS>(k, v, shift, S, s_soff);
        const uint32_t dst = s_gofs[d] + (i - s_excl[d]);
        kout[dst] = k;
        vout[dst] = v;
    }
}

// Digit counts of every tile of a count / scan / scatter pass (the same tiles
// and digits as k_sort_scatter): counts[d * ntiles + tile].
template <bool SCENE_PASS, int ITEMS>
__global__ __launch_bounds__(kSortThreads) void k_sort_count(
    const uint32_t *__restrict__ kin, const int32_t *__restrict__ vin, uint32_t n, uint32_t shift,
    const SceneDesc *__restrict__ desc, uint32_t S, uint32_t *__restrict__ counts, uint32_t ntiles)
{
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t s_wave[kWaves][kRadix];
    __shared__ uint32_t s_soff[SCENE_PASS ? kRadix : 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int j = tid; j < kWaves * kRadix; j += kSortThreads) (&s_wave[0][0])[j] = 0u;
    if (SCENE_PASS)
        for (int j = tid; j < kRadix; j += kSortThreads) s_soff[j] = j < (int)S ? desc[j].prim_off : 0xFFFFFFFFu;
    __syncthreads();
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * (uint32_t)(kSortThreads * ITEMS);
    uint32_t key[ITEMS];
    int32_t val[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {  // all loads before any use; past n: the last key again, not counted
        const uint32_t i = base + wave * (64 * ITEMS) + r * 64 + lane;
        const uint32_t ic = i < n ? i : n - 1;
        key[r] = kin[ic];
        val[r] = SCENE_PASS ? vin[ic] : 0;
    }
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const uint32_t i = base + wave * (64 * ITEMS) + r * 64 + lane;
        if (i < n) atomicAdd(&s_wave[wave][sort_digit<SCENE_PASS>(key[r], val[r], shift, S, s_soff)], 1u);
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) c += s_wave[w][tid];
    counts[(size_t)tid * ntiles + tile] = c;
}

// Global offset of every (digit, tile) of a count / scan / scatter pass: the
// digit's base (the exclusive scan of the pass histogram, or the scene's first
// sorted position) plus the counts of the digit in earlier tiles.  One block
// per digit.
__global__ __launch_bounds__(kSortThreads) void k_sort_scan(const uint32_t *__restrict__ counts,
                                                            uint32_t *__restrict__ offs, uint32_t ntiles,
                                                            const uint32_t *__restrict__ hist,
                                                            const SceneDesc *__restrict__ desc, uint32_t S,
                                                            uint32_t scene_pass)
{
    constexpr int kPer = 16;  // consecutive tiles per thread and round
    // a round's counts, staged through LDS so that global loads and stores
    // are coalesced: tile j of the round at row j / kPer, column j % kPer,
    // rows padded to kPer + 1 words so that a thread's row reads without
    // bank conflicts
    __shared__ uint32_t s_c[kSortThreads * (kPer + 1)];
    __shared__ uint32_t s_scan[kSortThreads / 64];
    __shared__ uint32_t s_base;
    const uint32_t d = blockIdx.x;
    const int tid = threadIdx.x;
    if (scene_pass) {
        if (tid == 0) s_base = d < S ? desc[d].prim_off : 0u;
    } else {
        const uint32_t e = block_exclusive_scan(hist[tid], s_scan, nullptr);
        if ((uint32_t)tid == d) s_base = e;
    }
    __syncthreads();
    uint32_t run = s_base;
    const uint32_t *c = counts + (size_t)d * ntiles;
    uint32_t *o = offs + (size_t)d * ntiles;
    auto at = [](uint32_t j) { return (j / kPer) * (kPer + 1) + j % kPer; };
    for (uint32_t t0 = 0; t0 < ntiles; t0 += kSortThreads * kPer) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t j = (uint32_t)(k * kSortThreads + tid), t = t0 + j;
            s_c[at(j)] = t < ntiles ? c[t] : 0u;
        }
        __syncthreads();
        uint32_t v[kPer];
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            v[k] = s_c[tid * (kPer + 1) + k];
            sum += v[k];
        }
        uint32_t total;
        uint32_t x = run + block_exclusive_scan(sum, s_scan, &total);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            s_c[tid * (kPer + 1) + k] = x;
            x += v[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t j = (uint32_t)(k * kSortThreads + tid), t = t0 + j;
            if (t < ntiles) o[t] = s_c[at(j)];
        }
        run += total;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Stage 4a: split tables.  For scene s with t = tbits, T[w] = first sorted
// position whose code is >= w << (30 - t), for w in [0, 2^t]; a split search
// for a node then starts inside one bucket.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_split_table(const SceneDesc *__restrict__ desc, uint32_t S,
                                                     uint64_t total, const uint32_t *__restrict__ codes,
                                                     uint32_t *__restrict__ table)
{
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = upper_scene(S, (uint32_t)e, [&](uint32_t i) { return (uint32_t)desc[i].table_off; });
        const SceneDesc D = desc[s];
        const uint32_t w = (uint32_t)(e - D.table_off);
        const uint64_t target = (uint64_t)w << (30 - D.tbits);
        uint32_t lo = D.prim_off, hi = D.prim_off + D.nt;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if ((uint64_t)codes[mid] < target) lo = mid + 1;
            else hi = mid;
        }
        table[e] = lo;
    }
}

// First sorted position in (first, last] whose code is >= X (the split of a
// node spanning [first, last] at bit p, X = ((code_first >> p) | 1) << p).
__device__ __forceinline__ uint32_t split_search(const uint32_t *__restrict__ codes,
                                                 const uint32_t *__restrict__ table,
                                                 const SceneDesc &D, uint32_t first, uint32_t last,
                                                 uint32_t X)
{
    const uint32_t sh = 30 - D.tbits;
    const uint32_t w = X >> sh;
    uint32_t lo = table[D.table_off + w];
    lo = lo < first + 1 ? first + 1 : lo;
    // X on a bucket boundary (a split bit p >= 30 - tbits, as at most nodes
    // of the top levels): the bucket's first position is the answer, since
    // every code from there on is >= X; the search would only confirm it.
    if ((X & ((1u << sh) - 1u)) == 0u) return lo;
    uint32_t hi = table[D.table_off + w + 1];
    hi = hi > last ? last : hi;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (codes[mid] < X) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Stage 4b: big nodes (LBVH_Build, LBVH.hlsl:134-226, for nodes holding more
// than kSubtreeMax primitives).  One block per scene walks the big frontier
// level by level.  Entries are {heap index, first sorted position, count,
// parent} with parent = -1 for the root, else (k << 6) | (side << 5) | split
// bit of the parent.  Leaves and small nodes go to the subtree list.
// ---------------------------------------------------------------------------
// One big node (more than kSubtreeMax primitives) of the top levels: a leaf
// or a node too small for this stage goes to the subtree list; otherwise the
// node's heap word and traversal node are written, it goes to the tops (for
// the refit), and its children go to the next frontier or to the subtree
// list.  The caller appends the plan's entries to those lists.
struct TopPlan {
    int4 c0, c1, rec;    // leaf: c0 = the node itself; internal: the children and the tops record
    bool valid = false;  // a node was planned
    bool leaf = false;
    bool big0 = false, big1 = false;  // internal: child goes to the next frontier
    __device__ uint32_t nsmall() const { return !valid ? 0u : leaf ? 1u : (uint32_t)!big0 + (uint32_t)!big1; }
    __device__ uint32_t nnext() const { return valid && !leaf ? (uint32_t)big0 + (uint32_t)big1 : 0u; }
    __device__ uint32_t ntop() const { return valid && !leaf ? 1u : 0u; }
    // writes the plan's entries from slots small[bs..] and nxt[bn..], children in order
    __device__ void write_lists(int4 *small, uint32_t bs, int4 *nxt, uint32_t bn) const
    {
        if (!valid) return;
        if (leaf) {
            small[bs] = c0;
            return;
        }
        if (big0) nxt[bn++] = c0;
        else small[bs++] = c0;
        if (big1) nxt[bn] = c1;
        else small[bs] = c1;
    }
};

__device__ __forceinline__ void top_plan(const int4 e, const SceneDesc &D, uint32_t s, uint32_t leaf_level,
                                         const uint32_t *__restrict__ codes, const uint32_t *__restrict__ table,
                                         rt_node32 *__restrict__ sheap, TravNode *__restrict__ trav,
                                         SceneHdr *__restrict__ hdr, TopPlan &t)
{
    t.valid = true;
    const uint32_t h = (uint32_t)e.x, first = (uint32_t)e.y;
    const int32_t cnt = e.z, pr = e.w;
    const uint32_t last = first + cnt - 1;
    const int pp = pr < 0 ? 31 : (pr & 31);
    const uint32_t cf = codes[first], cl = codes[last];
    int axis = 0, p = 0;
    if (classify(h, cnt, pp, leaf_level, cf, cl, &axis, &p)) {
        t.leaf = true;
        t.c0 = e;
        return;
    }
    const uint32_t X = ((cf >> p) | 1u) << p;
    const uint32_t sp = split_search(codes, table, D, first, last, X);
    const int32_t k = (int32_t)sp - 1;
    write_heap_word(&sheap[h], (cnt << 3) | axis, (int32_t)(first - D.prim_off));
    TravNode *tn = &trav[k];
    tn->axis = axis;
    tn->heap = (int32_t)h;
    if (pr < 0) hdr[s].root_ref = k;
    else trav[pr >> 6].ref[(pr >> 5) & 1] = k;
    t.rec = make_int4(k, (int)h, pr, (int)first);
    const int32_t cl_cnt = (int32_t)(sp - first), cr_cnt = (int32_t)(last - sp + 1);
    t.c0 = make_int4((int)(2 * h), (int)first, cl_cnt, (k << 6) | p);
    t.c1 = make_int4((int)(2 * h + 1), (int)sp, cr_cnt, (k << 6) | (1 << 5) | p);
    t.big0 = cl_cnt > kSubtreeMax;
    t.big1 = cr_cnt > kSubtreeMax;
}

// Reserves cnt slots of a global list for each lane of the wave with one
// atomic: returns the lane's first slot.  Every lane of the wave calls it.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t *counter, uint32_t cnt)
{
    const int lane = __lane_id();
    uint32_t x = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    const uint32_t total = __shfl(x, 63);
    uint32_t base = 0;
    if (lane == 63 && total) base = atomicAdd(counter, total);
    base = __shfl(base, 63);
    return base + x - cnt;
}

__global__ __launch_bounds__(kTopThreads) void k_top(
    const SceneDesc *__restrict__ desc, SceneHdr *__restrict__ hdr, const uint32_t *__restrict__ codes,
    const uint32_t *__restrict__ table, int4 *__restrict__ big_a, int4 *__restrict__ big_b,
    int4 *__restrict__ small, uint32_t *__restrict__ counters, int4 *__restrict__ top_nodes,
    uint32_t *__restrict__ top_counts, rt_node32 *__restrict__ heap, TravNode *__restrict__ trav)
{
    // A level of at most kTopLds nodes is also kept in LDS with its first and
    // last codes, so that a level costs two global round trips (the split
    // table, then the codes on either side of the split), not five.
    __shared__ uint32_t s_n, s_next, s_ntop;
    __shared__ int4 s_fr[2][kTopLds];
    __shared__ uint2 s_cc[2][kTopLds];
    const uint32_t s = blockIdx.x;
    const SceneDesc D = desc[s];
    const int tid = threadIdx.x;
    if (D.nt == 0) {
        if (tid == 0) {
            hdr[s].root_ref = -1;
            hdr[s].levels = 0;
        }
        return;
    }
    const uint32_t leaf_level = 1u << (D.depth - 1);
    rt_node32 *sheap = heap + D.heap_off;
    int4 *cur = big_a + D.prim_off, *nxt = big_b + D.prim_off;
    int4 *tops = top_nodes + D.prim_off;
    uint32_t *tcounts = top_counts + (size_t)s * kMaxLevels;
    if (tid == 0) {
        const int4 root = make_int4(1, (int)D.prim_off, (int)D.nt, -1);
        if (D.nt > (uint32_t)kSubtreeMax) {
            cur[0] = root;
            s_fr[0][0] = root;
            s_cc[0][0] = make_uint2(codes[D.prim_off], codes[D.prim_off + D.nt - 1]);
            s_n = 1;
        } else {
            small[atomicAdd(&counters[kCntSmall], 1u)] = root;
            s_n = 0;
        }
        s_ntop = 0;
        s_next = 0;
    }
    __syncthreads();
    uint32_t level = 0;
    int c = 0;  // LDS frontier buffer of this level
    while (true) {
        const uint32_t n = s_n;
        if (n == 0) break;
        const bool in_lds = n <= (uint32_t)kTopLds;
        const uint32_t level_start = s_ntop;
        __syncthreads();
        for (uint32_t i = tid; i < n; i += kTopThreads) {
            const int4 e = in_lds ? s_fr[c][i] : cur[i];
            const uint32_t h = (uint32_t)e.x, first = (uint32_t)e.y;
            const int32_t cnt = e.z, pr = e.w;
            const uint32_t last = first + cnt - 1;
            const int pp = pr < 0 ? 31 : (pr & 31);
            uint32_t cf, cl;
            if (in_lds) {
                const uint2 cc = s_cc[c][i];
                cf = cc.x;
                cl = cc.y;
            } else {
                cf = codes[first];
                cl = codes[last];
            }
            int axis = 0, p = 0;
            if (classify(h, cnt, pp, leaf_level, cf, cl, &axis, &p)) {
                small[atomicAdd(&counters[kCntSmall], 1u)] = e;
                continue;
            }
            const uint32_t X = ((cf >> p) | 1u) << p;
            const uint32_t sp = split_search(codes, table, D, first, last, X);
            const uint32_t c_left = codes[sp - 1], c_right = codes[sp];  // the children's inner codes
            const int32_t k = (int32_t)sp - 1;
            write_heap_word(&sheap[h], (cnt << 3) | axis, (int32_t)(first - D.prim_off));
            TravNode *tn = &trav[k];
            tn->axis = axis;
            tn->heap = (int32_t)h;
            if (pr < 0) hdr[s].root_ref = k;
            else trav[pr >> 6].ref[(pr >> 5) & 1] = k;
            tops[atomicAdd(&s_ntop, 1u)] = make_int4(k, (int)h, pr, (int)first);
            const int32_t cl_cnt = (int32_t)(sp - first), cr_cnt = (int32_t)(last - sp + 1);
            const int4 c0 = make_int4((int)(2 * h), (int)first, cl_cnt, (k << 6) | p);
            const int4 c1 = make_int4((int)(2 * h + 1), (int)sp, cr_cnt, (k << 6) | (1 << 5) | p);
            if (cl_cnt > kSubtreeMax) {
                const uint32_t slot = atomicAdd(&s_next, 1u);
                nxt[slot] = c0;
                if (slot < (uint32_t)kTopLds) {
                    s_fr[c ^ 1][slot] = c0;
                    s_cc[c ^ 1][slot] = make_uint2(cf, c_left);
                }
            } else {
                small[atomicAdd(&counters[kCntSmall], 1u)] = c0;
            }
            if (cr_cnt > kSubtreeMax) {
                const uint32_t slot = atomicAdd(&s_next, 1u);
                nxt[slot] = c1;
                if (slot < (uint32_t)kTopLds) {
                    s_fr[c ^ 1][slot] = c1;
                    s_cc[c ^ 1][slot] = make_uint2(c_right, cl);
                }
            } else {
                small[atomicAdd(&counters[kCntSmall], 1u)] = c1;
            }
        }
        __syncthreads();
        if (tid == 0) {
            tcounts[level] = s_ntop - level_start;
            s_n = s_next;
            s_next = 0;
        }
        int4 *t = cur;
        cur = nxt;
        nxt = t;
        c ^= 1;
        ++level;
        __syncthreads();
    }
    if (tid == 0) hdr[s].levels = level;
}

// ---------------------------------------------------------------------------
// Stage 4b': the same emission for one big scene (a batch of one scene of at
// least kTopWideMinN primitives), one launch per level over the whole chip
// instead of one block: launch l processes the frontier of level l (level 0:
// the root), whose size the launch before counted in counters[kCntFront + l];
// the kernel boundaries order the levels.  Tops of level l land after those
// of the levels above (top_counts[l] counts them), as k_top lays them out.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kTopLevelThreads) void k_top_level(
    const SceneDesc *__restrict__ desc, SceneHdr *__restrict__ hdr, const uint32_t *__restrict__ codes,
    const uint32_t *__restrict__ table, int4 *__restrict__ big_a, int4 *__restrict__ big_b,
    int4 *__restrict__ small, uint32_t *__restrict__ counters, int4 *__restrict__ top_nodes,
    uint32_t *__restrict__ top_counts, rt_node32 *__restrict__ heap, TravNode *__restrict__ trav,
    uint32_t level)
{
    __shared__ uint32_t s_base;
    const SceneDesc D = desc[0];
    uint32_t *front = counters + kCntFront;
    const int tid = threadIdx.x;
    if (level == 0) {  // one block: clear the level counts, then the root
        for (int l = tid; l < kMaxLevels; l += kTopLevelThreads) top_counts[l] = 0u;
        __syncthreads();
        if (tid == 0) {
            hdr[0].levels = 1;
            TopPlan t;
            top_plan(make_int4(1, (int)D.prim_off, (int)D.nt, -1), D, 0, 1u << (D.depth - 1), codes, table,
                     heap + D.heap_off, trav, hdr, t);
            if (t.leaf) {
                small[atomicAdd(&counters[kCntSmall], 1u)] = t.c0;
            } else {
                if (t.big0) big_b[D.prim_off + atomicAdd(&front[1], 1u)] = t.c0;
                else small[atomicAdd(&counters[kCntSmall], 1u)] = t.c0;
                if (t.big1) big_b[D.prim_off + atomicAdd(&front[1], 1u)] = t.c1;
                else small[atomicAdd(&counters[kCntSmall], 1u)] = t.c1;
                top_nodes[D.prim_off + atomicAdd(&top_counts[0], 1u)] = t.rec;
            }
        }
        return;
    }
    const uint32_t n = front[level];
    if (n == 0) return;
    if (tid == 0) {
        uint32_t o = 0;
        for (uint32_t l = 0; l < level; ++l) o += top_counts[l];
        s_base = o;
        if (blockIdx.x == 0) hdr[0].levels = level + 1;
    }
    __syncthreads();
    const int4 *cur = ((level & 1) ? big_b : big_a) + D.prim_off;
    int4 *nxt = ((level & 1) ? big_a : big_b) + D.prim_off;
    int4 *tops = top_nodes + D.prim_off + s_base;
    // wave-uniform loop: every lane takes part in the list reservations
    const uint32_t lane = (uint32_t)__lane_id();
    for (uint32_t i0 = blockIdx.x * kTopLevelThreads + (uint32_t)(tid & ~63); i0 < n;
         i0 += gridDim.x * kTopLevelThreads) {
        TopPlan t;
        if (i0 + lane < n)
            top_plan(cur[i0 + lane], D, 0, 1u << (D.depth - 1), codes, table, heap + D.heap_off, trav, hdr, t);
        const uint32_t bs = wave_reserve(&counters[kCntSmall], t.nsmall());
        const uint32_t bn = wave_reserve(&front[level + 1], t.nnext());
        const uint32_t bt = wave_reserve(&top_counts[level], t.ntop());
        t.write_lists(small, bs, nxt, bn);
        if (t.ntop()) tops[bt] = t.rec;
    }
}

// ---------------------------------------------------------------------------
// Stage 4c: small subtrees.  A persistent block takes one entry of the
// subtree list at a time.  A leaf (of any size) is reduced by the whole block;
// an internal node (<= kSubtreeMax primitives) has its codes staged in LDS and
// its subtree emitted with the same rules as k_top, all nodes at once, in LDS
// only; its triangle records are fetched meanwhile and laid out once the
// leaves are known, and the subtree is refitted bottom-up (LBVH_BuildAABB,
// LBVH.hlsl:228-270) from LDS child boxes.
// ---------------------------------------------------------------------------
static_assert(kSubtreeMax % kSubtreeThreads == 0, "a subtree's positions split evenly over the threads");

struct SubtreeLds {
    uint32_t codes[kSubtreeMax];
    float pbox[kSubtreeMax][6];         // each primitive's box, by local position
    int4 inode[kSubtreeMax];            // k, heap index, parent, first (by depth)
    int32_t iword[kSubtreeMax];         // (count << 2) | axis
    float cbox[kSubtreeMax][12];        // child boxes by local internal index
    int32_t cref[kSubtreeMax][2];
    int4 leaf[kSubtreeMax];             // heap index, local first, (count << 2) | axis, parent
    uint16_t pos_leaf[kSubtreeMax];     // the leaf of each local position
    // emission, per split position i (between local positions i and i + 1)
    int8_t dl[kSubtreeMax];             // highest differing bit of codes i and i + 1; -1: equal codes
    uint16_t ra[kSubtreeMax], rb[kSubtreeMax];  // the node's positions [ra, rb]
    uint16_t par[kSubtreeMax];          // its parent split position (the root: itself)
    uint8_t side[kSubtreeMax];          // its side under the parent
    uint16_t anc[kSubtreeMax];          // pointer jumping: an ancestor,
    uint8_t dist[kSubtreeMax];          //   the edges up to it (kDeepPath: more than any heap index holds)
    uint32_t bits[kSubtreeMax];         //   and the side bits of that path, first edge highest
    uint8_t claim[2][kSubtreeMax];      // child side taken by an internal node
    uint32_t dcnt[kMaxLevels + 2];      // internal nodes per depth
    uint32_t lvl[kMaxLevels + 2];
    float red[kSubtreeThreads / 64 * 6];
    uint32_t item, nleaf, levels;
};
constexpr uint32_t kDeepPath = 31;

__global__ __launch_bounds__(kSubtreeThreads) void k_subtree(
    const SceneDesc *__restrict__ desc, uint32_t S, SceneHdr *__restrict__ hdr,
    const uint32_t *__restrict__ codes, const int32_t *__restrict__ vals,
    const TriRec *__restrict__ trec, const int4 *__restrict__ small, uint32_t *__restrict__ counters,
    rt_node32 *__restrict__ heap, TravNode *__restrict__ trav, TravTri *__restrict__ tris)
{
    __shared__ SubtreeLds L;
    const int tid = threadIdx.x;
    const uint32_t nsmall = counters[kCntSmall];
    while (true) {
        if (tid == 0) L.item = atomicAdd(&counters[kCntWork], 1u);
        __syncthreads();
        const uint32_t item = L.item;
        __syncthreads();
        if (item >= nsmall) break;
        const int4 root = small[item];
        const uint32_t rfirst = (uint32_t)root.y;
        const int32_t rcnt = root.z, rpr = root.w;
        const uint32_t s = scene_of_prim(desc, S, rfirst);
        const SceneDesc D = desc[s];
        const uint32_t leaf_level = 1u << (D.depth - 1);
        rt_node32 *sheap = heap + D.heap_off;
        const uint32_t rlast = rfirst + rcnt - 1;
        int raxis = 0, rp = 0;
        const int pp0 = rpr < 0 ? 31 : (rpr & 31);
        const bool root_leaf =
            classify((uint32_t)root.x, rcnt, pp0, leaf_level, codes[rfirst], codes[rlast], &raxis, &rp);
        if (root_leaf) {
            // one leaf, reduced by the whole block
            float lo[3] = {INFINITY, INFINITY, INFINITY};
            float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int32_t j = tid; j < rcnt; j += kSubtreeThreads)
                leaf_prim(vals, trec, tris, rfirst + j, j, rcnt, (uint32_t)root.x, lo, hi);
            block_box_reduce<kSubtreeThreads / 64>(lo, hi, L.red);
            if (tid == 0) {
                const uint32_t h = (uint32_t)root.x;
                write_heap(&sheap[h], lo, hi, (rcnt << 3) | 4 | raxis, (int32_t)(rfirst - D.prim_off));
                if (h < leaf_level) {
                    write_heap_word(&sheap[2 * h], kInvalid, kInvalid);
                    write_heap_word(&sheap[2 * h + 1], kInvalid, kInvalid);
                }
                const int32_t ref = (int32_t)leaf_ref(rfirst, (uint32_t)rcnt);
                if (rpr < 0) {
                    hdr[s].root_ref = ref;
                    for (int k = 0; k < 3; ++k) {
                        hdr[s].root_box[k] = lo[k];
                        hdr[s].root_box[3 + k] = hi[k];
                    }
                } else {
                    TravNode *pn = &trav[rpr >> 6];
                    const int side = (rpr >> 5) & 1;
                    pn->ref[side] = ref;
                    for (int k = 0; k < 3; ++k) {
                        pn->box[side * 6 + k] = lo[k];
                        pn->box[side * 6 + 3 + k] = hi[k];
                    }
                }
            }
            __syncthreads();
            continue;
        }
        // internal subtree root: rcnt <= kSubtreeMax.  The topology is emitted
        // from the codes in LDS; leaves are only listed here.
        // The subtree's triangle records are fetched now, by sorted id, so that
        // their latency (a random gather: the records are in input order)
        // overlaps the emission; the barriers below do not wait for them.
        constexpr int kPerThread = kSubtreeMax / kSubtreeThreads;
        TriRec pre[kPerThread];
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            const int32_t j = tid + q * kSubtreeThreads;
            if (j < rcnt) {
                L.codes[j] = codes[rfirst + j];
                pre[q] = trec[vals[rfirst + j]];
            }
        }
        const uint32_t C = (uint32_t)rcnt;
        const uint32_t hroot = (uint32_t)root.x;
        if (tid < kMaxLevels + 2) L.dcnt[tid] = 0u;
        if (tid == 0) L.nleaf = 0u;
        __syncthreads();
        // The topology of the subtree, with every node at once instead of
        // level by level (the radix tree of the sorted codes, as in Karras
        // 2012).  Split position i (between local positions i and i + 1) is
        // a node when the codes there differ; its split bit is their highest
        // differing bit d, it spans the positions sharing the code bits above
        // d, and its parent is whichever split bounding that span has the
        // lower split bit.  These are the nodes, ranges and splits that
        // build_recursive (LBVH.cpp:190-314) reaches level by level; the
        // leaves are the child sides no node takes, and the depth limit drops
        // every node whose heap index reaches 2^(depth-1).
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            const uint32_t i = (uint32_t)tid + q * kSubtreeThreads;
            if (i + 1 < C) {
                const uint32_t x = L.codes[i] ^ L.codes[i + 1];
                L.dl[i] = x ? (int8_t)(31 - __clz(x)) : (int8_t)-1;
            }
            if (i < C) L.claim[0][i] = L.claim[1][i] = 0;
        }
        __syncthreads();
        bool node[kPerThread];
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            const uint32_t i = (uint32_t)tid + q * kSubtreeThreads;
            node[q] = i + 1 < C && L.dl[i] >= 0;
            if (!node[q]) continue;
            const uint32_t sh = (uint32_t)L.dl[i] + 1u;  // <= 30
            const uint32_t lowX = (L.codes[i] >> sh) << sh, highX = lowX + (1u << sh);
            uint32_t lo = 0, hi = i;  // first position whose code is >= lowX
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (L.codes[mid] < lowX) lo = mid + 1;
                else hi = mid;
            }
            const uint32_t ra = lo;
            lo = i + 2;  // first position whose code is >= highX (i + 1 shares the prefix)
            hi = C;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (L.codes[mid] < highX) lo = mid + 1;
                else hi = mid;
            }
            const uint32_t rb = lo - 1;
            uint32_t parent = i, sd = 0;
            if (ra > 0 || rb + 1 < C) {
                if (ra == 0 || (rb + 1 < C && L.dl[rb] < L.dl[ra - 1])) {
                    parent = rb;  // the split after the span: this node is its left child
                } else {
                    parent = ra - 1;  // the split before the span: its right child
                    sd = 1;
                }
            }
            L.ra[i] = (uint16_t)ra;
            L.rb[i] = (uint16_t)rb;
            L.par[i] = (uint16_t)parent;
            L.side[i] = (uint8_t)sd;
            L.anc[i] = (uint16_t)parent;
            L.dist[i] = parent == i ? 0 : 1;
            L.bits[i] = sd;
        }
        // each node's heap index: its path to the subtree root by pointer
        // jumping, h = (h_root << depth) | side bits
        while (true) {
            uint32_t na[kPerThread], nd[kPerThread], nb[kPerThread];
            bool ch[kPerThread];
            bool any = false;
            __syncthreads();
#pragma unroll
            for (int q = 0; q < kPerThread; ++q) {
                const uint32_t i = (uint32_t)tid + q * kSubtreeThreads;
                ch[q] = false;
                if (!node[q]) continue;
                const uint32_t a0 = L.anc[i];
                const uint32_t a1 = L.anc[a0];
                if (a1 == a0) continue;  // a0 is the root
                const uint32_t d0 = L.dist[i], d1 = L.dist[a0];
                const uint32_t dsum = d0 + d1;
                nd[q] = dsum > kDeepPath - 1 ? kDeepPath : dsum;
                nb[q] = nd[q] == kDeepPath ? 0u : (L.bits[a0] << d0) | L.bits[i];
                na[q] = a1;
                ch[q] = any = true;
            }
            if (!__syncthreads_or(any)) break;
#pragma unroll
            for (int q = 0; q < kPerThread; ++q) {
                if (!ch[q]) continue;
                const uint32_t i = (uint32_t)tid + q * kSubtreeThreads;
                L.anc[i] = (uint16_t)na[q];
                L.dist[i] = (uint8_t)nd[q];
                L.bits[i] = nb[q];
            }
        }
        // nodes above the depth limit: depth counts, the child sides they take
        uint32_t rank[kPerThread], hh[kPerThread];
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            const uint32_t i = (uint32_t)tid + q * kSubtreeThreads;
            if (!node[q]) continue;
            const uint32_t d = L.dist[i];
            const uint64_t H = d < kDeepPath ? (((uint64_t)hroot << d) | L.bits[i]) : ~0ull;
            node[q] = H < (uint64_t)leaf_level;
            if (!node[q]) continue;
            hh[q] = (uint32_t)H;
            rank[q] = atomicAdd(&L.dcnt[d], 1u);
            if (L.par[i] != i) L.claim[L.side[i]][L.par[i]] = 1;
        }
        __syncthreads();
        if (tid < 64) {  // depth offsets: the depths in use are 0 .. levels - 1
            const uint32_t c = tid < kMaxLevels + 2 ? L.dcnt[tid] : 0u;
            uint32_t x = c;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o);
                if (tid >= o) x += y;
            }
            if (tid < kMaxLevels + 1) L.lvl[tid + 1] = x;
            if (tid == 0) L.lvl[0] = 0u;
            const uint64_t used = __ballot(c != 0u);
            if (tid == 0) L.levels = used ? 64u - (uint32_t)__clzll(used) : 0u;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            const uint32_t i = (uint32_t)tid + q * kSubtreeThreads;
            if (!node[q]) continue;
            const uint32_t d = L.dist[i], h = hh[q], pi = L.par[i];
            const int dd = L.dl[i];
            const uint32_t ra = L.ra[i], rb = L.rb[i];
            const int32_t k = (int32_t)(rfirst + i);
            const int32_t pr = pi == i ? rpr : (((int32_t)(rfirst + pi) << 6) | ((int32_t)L.side[i] << 5) | L.dl[pi]);
            const uint32_t slot = L.lvl[d] + rank[q];
            L.inode[slot] = make_int4(k, (int)h, pr, (int)(rfirst + ra));
            L.iword[slot] = ((int32_t)(rb - ra + 1) << 2) | ((30 - dd) % 3);
            if (pi == i) {  // the subtree root, a child of the top levels
                if (rpr < 0) hdr[s].root_ref = k;
                else trav[rpr >> 6].ref[(rpr >> 5) & 1] = k;
            } else {
                L.cref[pi][L.side[i]] = k;
            }
            // the child sides no node took are leaves (LBVH.cpp:194-209):
            // axis (31 - p) % 3 for one primitive, at the depth limit, or under
            // a split at bit 0; axis 1 for equal codes
            for (int sd = 0; sd < 2; ++sd) {
                if (L.claim[sd][i]) continue;
                const uint32_t lf = sd ? i + 1 : ra, ll = sd ? rb : i;
                const int32_t lc = (int32_t)(ll - lf + 1);
                const uint32_t hl = 2 * h + (uint32_t)sd;
                const int axis = (lc <= 1 || hl >= leaf_level || dd == 0) ? (31 - dd) % 3 : 1;
                const uint32_t ls = atomicAdd(&L.nleaf, 1u);
                L.leaf[ls] = make_int4((int)hl, (int)lf, (lc << 2) | axis, (k << 6) | (sd << 5) | dd);
                for (int32_t j = 0; j < lc; ++j) L.pos_leaf[lf + j] = (uint16_t)ls;
            }
        }
        __syncthreads();
        const uint32_t level = L.levels;
        // every leaf triangle at once: one thread per position gathers its
        // record, writes its sorted-order triangle and keeps its box in LDS
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            const int32_t p = tid + q * kSubtreeThreads;
            if (p < rcnt) {
                const int4 lf = L.leaf[L.pos_leaf[p]];
                float lo[3] = {INFINITY, INFINITY, INFINITY};
                float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
                leaf_prim_rec(pre[q], tris, rfirst + p, p - lf.y, lf.z >> 2, (uint32_t)lf.x, lo, hi);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    L.pbox[p][c] = lo[c];
                    L.pbox[p][3 + c] = hi[c];
                }
            }
        }
        __syncthreads();
        // leaf boxes (exact min / max, any order), heap words and parent slots
        for (uint32_t i = tid; i < L.nleaf; i += kSubtreeThreads) {
            const int4 lf = L.leaf[i];
            const uint32_t h = (uint32_t)lf.x, first = rfirst + (uint32_t)lf.y;
            const int32_t cnt = lf.z >> 2, axis = lf.z & 3, pr = lf.w;
            float lo[3] = {INFINITY, INFINITY, INFINITY};
            float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int32_t j = 0; j < cnt; ++j) {
                const float *b = L.pbox[lf.y + j];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    lo[c] = fminf(lo[c], b[c]);
                    hi[c] = fmaxf(hi[c], b[3 + c]);
                }
            }
            write_heap(&sheap[h], lo, hi, (cnt << 3) | 4 | axis, (int32_t)(first - D.prim_off));
            if (h < leaf_level) {
                write_heap_word(&sheap[2 * h], kInvalid, kInvalid);
                write_heap_word(&sheap[2 * h + 1], kInvalid, kInvalid);
            }
            // the root of this subtree is internal, so a leaf's parent is local
            const uint32_t ploc = (uint32_t)(pr >> 6) - rfirst;
            const int side = (pr >> 5) & 1;
            for (int c = 0; c < 3; ++c) {
                L.cbox[ploc][side * 6 + c] = lo[c];
                L.cbox[ploc][side * 6 + 3 + c] = hi[c];
            }
            L.cref[ploc][side] = (int32_t)leaf_ref(first, (uint32_t)cnt);
        }
        __syncthreads();
        // refit bottom-up; level l holds inode[lvl[l] .. lvl[l+1])
        for (int32_t l = (int32_t)level - 1; l >= 0; --l) {
            for (uint32_t i = L.lvl[l] + tid; i < L.lvl[l + 1]; i += kSubtreeThreads) {
                const int4 nd = L.inode[i];
                const int32_t k = nd.x, pr = nd.z;
                const uint32_t loc = (uint32_t)k - rfirst;
                const int32_t word = L.iword[i];
                const float *cb = L.cbox[loc];
                float lo[3], hi[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    lo[c] = fminf(cb[c], cb[6 + c]);
                    hi[c] = fmaxf(cb[3 + c], cb[9 + c]);
                }
                float4 *tp = reinterpret_cast<float4 *>(&trav[k]);
                tp[0] = make_float4(cb[0], cb[1], cb[2], cb[3]);
                tp[1] = make_float4(cb[4], cb[5], cb[6], cb[7]);
                tp[2] = make_float4(cb[8], cb[9], cb[10], cb[11]);
                reinterpret_cast<int4 *>(tp)[3] =
                    make_int4(L.cref[loc][0], L.cref[loc][1], word & 3, nd.y);
                write_heap(&sheap[nd.y], lo, hi, ((word >> 2) << 3) | (word & 3),
                           (int32_t)((uint32_t)nd.w - D.prim_off));
                if (l == 0) {
                    if (pr < 0) {
                        for (int c = 0; c < 3; ++c) {
                            hdr[s].root_box[c] = lo[c];
                            hdr[s].root_box[3 + c] = hi[c];
                        }
                    } else {
                        TravNode *pn = &trav[pr >> 6];
                        const int side = (pr >> 5) & 1;
                        for (int c = 0; c < 3; ++c) {
                            pn->box[side * 6 + c] = lo[c];
                            pn->box[side * 6 + 3 + c] = hi[c];
                        }
                    }
                } else {
                    const uint32_t ploc = (uint32_t)(pr >> 6) - rfirst;
                    const int side = (pr >> 5) & 1;
                    for (int c = 0; c < 3; ++c) {
                        L.cbox[ploc][side * 6 + c] = lo[c];
                        L.cbox[ploc][side * 6 + 3 + c] = hi[c];
                    }
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// Stage 5: refit of the big nodes, bottom-up, one block per scene.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kTopThreads) void k_top_refit(
    const SceneDesc *__restrict__ desc, SceneHdr *__restrict__ hdr,
    const int4 *__restrict__ top_nodes, const uint32_t *__restrict__ top_counts,
    rt_node32 *__restrict__ heap, TravNode *trav)
{
    __shared__ uint32_t s_off[kMaxLevels + 1];
    __shared__ int4 s_tops[kTopLds];
    const uint32_t s = blockIdx.x;
    const SceneDesc D = desc[s];
    if (D.nt == 0) return;
    const uint32_t levels = hdr[s].levels;
    if (threadIdx.x == 0) {
        uint32_t o = 0;
        for (uint32_t l = 0; l < levels; ++l) {
            s_off[l] = o;
            o += top_counts[(size_t)s * kMaxLevels + l];
        }
        s_off[levels] = o;
    }
    __syncthreads();
    const int4 *tops = top_nodes + D.prim_off;
    rt_node32 *sheap = heap + D.heap_off;
    // all of the scene's big nodes read at once when they fit in LDS: a level
    // then costs one global round trip (its children's boxes), not two
    const uint32_t total = s_off[levels];
    const bool in_lds = total <= (uint32_t)kTopLds;
    if (in_lds) {
        for (uint32_t i = threadIdx.x; i < total; i += kTopThreads) s_tops[i] = tops[i];
        __syncthreads();
    }
    for (int32_t l = (int32_t)levels - 1; l >= 0; --l) {
        for (uint32_t i = s_off[l] + threadIdx.x; i < s_off[l + 1]; i += kTopThreads) {
            const int4 nd = in_lds ? s_tops[i] : tops[i];
            const int32_t k = nd.x, pr = nd.z;
            const TravNode *tn = &trav[k];
            float lo[3], hi[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                lo[c] = fminf(tn->box[c], tn->box[6 + c]);
                hi[c] = fmaxf(tn->box[3 + c], tn->box[9 + c]);
            }
            rt_node32 *hn = &sheap[nd.y];
            for (int c = 0; c < 3; ++c) {
                hn->mn[c] = lo[c];
                hn->mx[c] = hi[c];
            }
            if (pr < 0) {
                for (int c = 0; c < 3; ++c) {
                    hdr[s].root_box[c] = lo[c];
                    hdr[s].root_box[3 + c] = hi[c];
                }
            } else {
                TravNode *pn = &trav[pr >> 6];
                const int side = (pr >> 5) & 1;
                for (int c = 0; c < 3; ++c) {
                    pn->box[side * 6 + c] = lo[c];
                    pn->box[side * 6 + 3 + c] = hi[c];
                }
            }
        }
        __syncthreads();
    }
}

// Stage 5': the refit of one big scene's top levels, one launch per level
// (deepest first) over the whole chip; launches past the scene's levels exit.
__global__ __launch_bounds__(kTopLevelThreads) void k_top_refit_level(
    SceneHdr *__restrict__ hdr, const SceneDesc *__restrict__ desc, const int4 *__restrict__ top_nodes,
    const uint32_t *__restrict__ top_counts, rt_node32 *__restrict__ heap, TravNode *trav, uint32_t level)
{
    __shared__ uint32_t s_base;
    if (level >= hdr[0].levels) return;
    const uint32_t n = top_counts[level];
    if (n == 0) return;
    const SceneDesc D = desc[0];
    if (threadIdx.x == 0) {
        uint32_t o = 0;
        for (uint32_t l = 0; l < level; ++l) o += top_counts[l];
        s_base = o;
    }
    __syncthreads();
    const int4 *tops = top_nodes + D.prim_off + s_base;
    rt_node32 *sheap = heap + D.heap_off;
    for (uint32_t i = blockIdx.x * kTopLevelThreads + threadIdx.x; i < n; i += gridDim.x * kTopLevelThreads) {
        const int4 nd = tops[i];
        const int32_t k = nd.x, pr = nd.z;
        const TravNode *tn = &trav[k];
        float lo[3], hi[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lo[c] = fminf(tn->box[c], tn->box[6 + c]);
            hi[c] = fmaxf(tn->box[3 + c], tn->box[9 + c]);
        }
        rt_node32 *hn = &sheap[nd.y];
        for (int c = 0; c < 3; ++c) {
            hn->mn[c] = lo[c];
            hn->mx[c] = hi[c];
        }
        if (pr < 0) {
            for (int c = 0; c < 3; ++c) {
                hdr[0].root_box[c] = lo[c];
                hdr[0].root_box[3 + c] = hi[c];
            }
        } else {
            TravNode *pn = &trav[pr >> 6];
            const int side = (pr >> 5) & 1;
            for (int c = 0; c < 3; ++c) {
                pn->box[side * 6 + c] = lo[c];
                pn->box[side * 6 + 3 + c] = hi[c];
            }
        }
    }
}

__global__ void k_transform(const float *__restrict__ rest, float *__restrict__ xyz, uint32_t nv,
                            Xform x)
{
    const float *m = x.m;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const float dx = rest[3 * (size_t)i + 0] - m[9];
    const float dy = rest[3 * (size_t)i + 1] - m[10];
    const float dz = rest[3 * (size_t)i + 2] - m[11];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const float a = m[3 * r + 0] * dx;
        const float b = m[3 * r + 1] * dy;
        const float c = m[3 * r + 2] * dz;
        const float sum = a + b;
        const float t = sum + c;
        xyz[3 * (size_t)i + r] = t + m[9 + r];
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
uint32_t sort_tiles(uint32_t n) { return (n + kSortTile - 1) / kSortTile; }
uint32_t bounds_blocks(uint32_t n) { return (n + kBoundsTile - 1) / kBoundsTile; }

uint32_t table_bits(uint32_t max_tris)
{
    uint32_t b = 0;
    while (b < 31 && (1u << b) < max_tris) ++b;
    const int t = (int)b - 1;
    return (uint32_t)(t < 4 ? 4 : (t > 18 ? 18 : t));
}

static inline uint32_t min_u32(uint32_t a, uint32_t b) { return a < b ? a : b; }

hipError_t launch_transform(Batch &b, uint32_t scene, const Xform &x, hipStream_t st)
{
    const SceneDesc &D = b.desc[scene];
    if (D.nv == 0) return hipSuccess;
    hipLaunchKernelGGL(k_transform, dim3((D.nv + 255) / 256), dim3(256), 0, st,
                       b.xyz_rest + 3 * (size_t)D.vert_off, b.xyz + 3 * (size_t)D.vert_off, D.nv, x);
    return hipGetLastError();
}

hipError_t launch_build(Batch &b, hipStream_t st, hipEvent_t *ev, int cu_count)
{
    const uint32_t S = b.nscenes;
    const uint32_t n = b.total_tris;
    const uint32_t tiles = sort_tiles(n);
    const uint32_t passes = kCodePasses + (S > 1 ? 1 : 0);
    const bool large = n >= kSortLargeN;  // count / scan / scatter passes, no look-back status to clear
    const uint32_t nzero = kCodePasses * kRadix + kCounters + (large ? 0u : passes * tiles * kRadix);
    if (ev && ev[0]) hipEventRecord(ev[0], st);
    const uint32_t nb = b.total_blocks > 0 ? b.total_blocks : 1;
    hipLaunchKernelGGL(k_bounds, dim3(nb), dim3(kBoundsThreads), 0, st, b.d_desc, S, b.total_blocks,
                       b.xyz, b.idx, b.trec, b.cent, b.partials, b.hist, nzero);
    bool big = false;
    for (uint32_t i = 0; i < S; ++i) big = big || b.desc[i].nblocks > kBoxInline;
    if (big) hipLaunchKernelGGL(k_scene_box, dim3(S), dim3(1024), 0, st, b.d_desc, b.partials, b.hdr);
    if (ev && ev[1]) hipEventRecord(ev[1], st);
    const int start = (int)(passes & 1);  // the sorted result ends in buffer 0
    if (b.total_blocks > 0)
        hipLaunchKernelGGL(k_morton, dim3(min_u32(b.total_blocks, kMortonBlocksPerCu * (uint32_t)cu_count)),
                           dim3(kBoundsThreads), 0, st, b.d_desc, S, b.total_blocks, b.cent, b.partials,
                           b.keys[start], b.vals[start], b.hist, b.hdr);
    if (ev && ev[2]) hipEventRecord(ev[2], st);
    if (n > 0) {
        // large: the tiles' digit counts and offsets in the status words (2 x lt x 256 of them)
        constexpr uint32_t kLargeTile = kSortThreads * kSortItemsLarge;
        const uint32_t lt = (n + kLargeTile - 1) / kLargeTile;
        uint32_t *counts = b.status, *offs = b.status + (size_t)lt * kRadix;
        for (uint32_t p = 0; p < passes; ++p) {
            const int src = (start + p) & 1, dst = src ^ 1;
            const bool scene_pass = p >= (uint32_t)kCodePasses;
            const uint32_t shift = scene_pass ? 0u : p * kRadixBits;
            const uint32_t *hist = b.hist + (scene_pass ? 0u : p * kRadix);
            if (large) {
                if (scene_pass)
                    hipLaunchKernelGGL((k_sort_count<true, kSortItemsLarge>), dim3(lt), dim3(kSortThreads), 0, st,
                                       b.keys[src], b.vals[src], n, shift, b.d_desc, S, counts, lt);
                else
                    hipLaunchKernelGGL((k_sort_count<false, kSortItemsLarge>), dim3(lt), dim3(kSortThreads), 0, st,
                                       b.keys[src], b.vals[src], n, shift, b.d_desc, S, counts, lt);
                hipLaunchKernelGGL(k_sort_scan, dim3(kRadix), dim3(kSortThreads), 0, st, counts, offs, lt, hist,
                                   b.d_desc, S, (uint32_t)scene_pass);
            }
            uint32_t *status = b.status + (size_t)p * tiles * kRadix;
            uint32_t *tc = b.counters + kCntTiles + p;
#define RT_SORT(SC, IT, LB, TH, GRID)                                                                      \
    hipLaunchKernelGGL((k_sort_scatter<SC, IT, LB, TH>), dim3(GRID), dim3(TH), 0, st, b.keys[src],           \
                       b.vals[src], b.keys[dst], b.vals[dst], n, shift, hist, b.d_desc, S, status, tc, offs, lt)
            if (large) {
                if (scene_pass) RT_SORT(true, kSortItemsLarge, false, kSortThreads, lt);
                else RT_SORT(false, kSortItemsLarge, false, kSortThreads, lt);
            } else {
                if (scene_pass) RT_SORT(true, kSortItemsSmall, true, kSortThreadsSmall, tiles);
                else RT_SORT(false, kSortItemsSmall, true, kSortThreadsSmall, tiles);
            }
#undef RT_SORT
        }
    }
    if (ev && ev[3]) hipEventRecord(ev[3], st);
    if (b.total_table > 0 && n > 0) {
        const uint32_t blocks = (uint32_t)min_u32((uint32_t)((b.total_table + 255) / 256), 4096u);
        hipLaunchKernelGGL(k_split_table, dim3(blocks), dim3(256), 0, st, b.d_desc, S, b.total_table,
                           b.keys[0], b.table);
    }
    // one big scene: the top levels over the whole chip, one launch per level
    const bool wide_top = S == 1 && n >= kTopWideMinN;
    const uint32_t depth = S == 1 ? b.desc[0].depth : 0u;
    auto level_grid = [&](uint32_t l) {  // at most 2^l nodes, and at most n / kSubtreeMax
        const uint64_t most = std::min<uint64_t>(l < 40 ? (1ull << l) : ~0ull, (uint64_t)n / kSubtreeMax + 1);
        return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((most + kTopLevelThreads - 1) / kTopLevelThreads,
                                                                   4ull * (uint32_t)cu_count));
    };
    if (wide_top) {
        for (uint32_t l = 0; l < depth && l < (uint32_t)kMaxLevels; ++l)
            hipLaunchKernelGGL(k_top_level, dim3(l == 0 ? 1u : level_grid(l)), dim3(kTopLevelThreads), 0, st,
                               b.d_desc, b.hdr, b.keys[0], b.table, b.big[0], b.big[1], b.small, b.counters,
                               b.top_nodes, b.top_counts, b.heap, b.trav, l);
    } else {
        hipLaunchKernelGGL(k_top, dim3(S), dim3(kTopThreads), 0, st, b.d_desc, b.hdr, b.keys[0], b.table,
                           b.big[0], b.big[1], b.small, b.counters, b.top_nodes, b.top_counts, b.heap,
                           b.trav);
    }
    if (ev && ev[4]) hipEventRecord(ev[4], st);
    if (n > 0) {
        const uint32_t blocks = min_u32(2u * (uint32_t)cu_count, (n + 63) / 64);
        hipLaunchKernelGGL(k_subtree, dim3(blocks > 0 ? blocks : 1), dim3(kSubtreeThreads), 0, st,
                           b.d_desc, S, b.hdr, b.keys[0], b.vals[0], b.trec, b.small,
                           b.counters, b.heap, b.trav, b.tris);
    }
    if (wide_top) {
        for (int32_t l = (int32_t)std::min<uint32_t>(depth, kMaxLevels) - 1; l >= 0; --l)
            hipLaunchKernelGGL(k_top_refit_level, dim3(level_grid((uint32_t)l)), dim3(kTopLevelThreads), 0, st,
                               b.hdr, b.d_desc, b.top_nodes, b.top_counts, b.heap, b.trav, (uint32_t)l);
    } else {
        hipLaunchKernelGGL(k_top_refit, dim3(S), dim3(kTopThreads), 0, st, b.d_desc, b.hdr, b.top_nodes,
                           b.top_counts, b.heap, b.trav);
    }
    if (ev && ev[5]) hipEventRecord(ev[5], st);
    return hipGetLastError();
}

}  // namespace rt
