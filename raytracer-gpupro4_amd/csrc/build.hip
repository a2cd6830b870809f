// LBVH build on gfx950: primitive bounds, 30-bit Morton codes, onesweep radix
// sort, level-by-level heap emission and a bottom-up refit with atomic flags.
//
// Reference stages (StructureConstruction::GPUBuildStructure,
// RayTracerCS/StructureConstruction.cpp:86-131):
//   LBVH_PreBuild   (LBVH.hlsl:15-132)  -> k_bounds + k_morton
//   BitonicSort     (BitonicSort.cpp:56-117) -> k_onesweep x 4 (stable LSD radix)
//   LBVH_Build      (LBVH.hlsl:134-226) -> k_emit_level x depth
//   LBVH_BuildAABB  (LBVH.hlsl:228-270) -> k_refit
// Node words follow the CPU builder (LBVH.cpp:190-314), including its axis
// rule; float arithmetic is IEEE single with no contraction (-ffp-contract=off).
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

// Reduces lo[3]/hi[3] over a 256-thread block; the result is valid in every
// thread.  red must hold 4 waves x 6 floats.
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
        for (int w = 1; w < 4; ++w) {
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

// Appends `mult` consecutive slots for every lane with pred set; all 64 lanes
// of the wave must call it.  Returns this lane's first slot.
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred, uint32_t mult)
{
    const uint64_t m = __ballot(pred);
    const int lane = __lane_id();
    if (m == 0) return 0;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m) * mult);
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)) * mult;
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

// ---------------------------------------------------------------------------
// Stage 1: per-triangle box and 3x centroid (LBVH.hlsl:26-36, LBVH.cpp:501-516);
// per-block partial scene boxes; clears the sort/emit scratch words.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBoundsThreads) void k_bounds(
    const float *__restrict__ xyz, const uint32_t *__restrict__ idx, uint32_t nt,
    PrimBox *__restrict__ pbox, float *__restrict__ cent, float *__restrict__ partials,
    uint32_t *__restrict__ zero_words, uint32_t nzero)
{
    __shared__ float red[24];
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const uint32_t base = blockIdx.x * (uint32_t)kBoundsTile;
#pragma unroll
    for (int r = 0; r < kBoundsItems; ++r) {
        const uint32_t i = base + r * kBoundsThreads + threadIdx.x;
        if (i < nt) {
            const uint32_t ia = idx[3 * (size_t)i + 0], ib = idx[3 * (size_t)i + 1],
                           ic = idx[3 * (size_t)i + 2];
            PrimBox b;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float a = xyz[3 * (size_t)ia + k], bb = xyz[3 * (size_t)ib + k],
                            c = xyz[3 * (size_t)ic + k];
                b.mn[k] = fminf(a, fminf(bb, c));
                b.mx[k] = fmaxf(a, fmaxf(bb, c));
                const float s = a + bb;
                cent[3 * (size_t)i + k] = s + c;
                lo[k] = fminf(lo[k], b.mn[k]);
                hi[k] = fmaxf(hi[k], b.mx[k]);
            }
            pbox[i] = b;
        }
    }
    block_box_reduce(lo, hi, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            partials[blockIdx.x * 6 + k] = lo[k];
            partials[blockIdx.x * 6 + 3 + k] = hi[k];
        }
    }
    for (uint32_t j = blockIdx.x * kBoundsThreads + threadIdx.x; j < nzero;
         j += gridDim.x * kBoundsThreads)
        zero_words[j] = 0u;
}

// ---------------------------------------------------------------------------
// Stage 2: scene box from the partials, Morton codes (LBVH.hlsl:92-114,
// LBVH.cpp:520-553), identity ids, digit histograms of all sort passes.
// Block 0 seeds the root of the emit frontier (LBVH.hlsl:123-131).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBoundsThreads) void k_morton(
    const float *__restrict__ cent, uint32_t nt, const float *__restrict__ partials,
    uint32_t nparts, uint32_t *__restrict__ keys, int32_t *__restrict__ vals,
    uint32_t *__restrict__ hist, SceneHdr *__restrict__ hdr, int4 *__restrict__ frontier0,
    uint32_t *__restrict__ counters)
{
    __shared__ float red[24];
    __shared__ uint32_t lhist[kSortPasses * kRadix];
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t j = threadIdx.x; j < nparts; j += kBoundsThreads) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], partials[j * 6 + k]);
            hi[k] = fmaxf(hi[k], partials[j * 6 + 3 + k]);
        }
    }
    for (int j = threadIdx.x; j < kSortPasses * kRadix; j += kBoundsThreads) lhist[j] = 0;
    block_box_reduce(lo, hi, red);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            hdr->scene_box[k] = lo[k];
            hdr->scene_box[3 + k] = hi[k];
        }
        hdr->root_ref = -1;
        frontier0[0] = make_int4(1, 0, (int)nt, -1);
        counters[0] = 1u;
    }
    float m[3], f[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float hk = 3.0f * hi[k];
        const float lk = 3.0f * lo[k];
        const float ext = hk - lk;
        m[k] = lk;
        f[k] = 1024.0f / ext;
    }
    const uint32_t base = blockIdx.x * (uint32_t)kBoundsTile;
#pragma unroll
    for (int r = 0; r < kBoundsItems; ++r) {
        const uint32_t i = base + r * kBoundsThreads + threadIdx.x;
        if (i < nt) {
            const uint32_t qx = quantise(cent[3 * (size_t)i + 0], m[0], f[0]);
            const uint32_t qy = quantise(cent[3 * (size_t)i + 1], m[1], f[1]);
            const uint32_t qz = quantise(cent[3 * (size_t)i + 2], m[2], f[2]);
            const uint32_t code = (part1by2(qx) << 2) | (part1by2(qy) << 1) | part1by2(qz);
            keys[i] = code;
            vals[i] = (int32_t)i;
#pragma unroll
            for (int p = 0; p < kSortPasses; ++p)
                atomicAdd(&lhist[p * kRadix + ((code >> (p * kRadixBits)) & (kRadix - 1))], 1u);
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kSortPasses * kRadix; j += kBoundsThreads)
        if (lhist[j]) atomicAdd(&hist[j], lhist[j]);
}

// ---------------------------------------------------------------------------
// Stage 3: one stable LSD pass of an onesweep radix sort (8-bit digit).
// Tiles are claimed in launch order through an atomic counter, so a tile only
// waits (decoupled look-back) on tiles that are already running.  Keys inside
// a tile are ranked in load order (wave, item, lane), which keeps the sort
// stable; ids start as 0..n-1, so the result is the (code, id) order.
// ---------------------------------------------------------------------------
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1u;

__global__ __launch_bounds__(kSortThreads) void k_onesweep(
    const uint32_t *__restrict__ kin, const int32_t *__restrict__ vin,
    uint32_t *__restrict__ kout, int32_t *__restrict__ vout, uint32_t n, uint32_t shift,
    const uint32_t *__restrict__ hist, uint32_t *__restrict__ status,
    uint32_t *__restrict__ tile_counter)
{
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave[kWaves][kRadix];
    __shared__ uint32_t s_excl[kRadix];
    __shared__ uint32_t s_gofs[kRadix];
    __shared__ uint32_t s_scan[kWaves];
    __shared__ uint32_t s_keys[kSortTile];
    __shared__ int32_t s_vals[kSortTile];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (int j = tid; j < kWaves * kRadix; j += kSortThreads) (&s_wave[0][0])[j] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * (uint32_t)kSortTile;

    uint32_t key[kSortItems];
    int32_t val[kSortItems];
    uint32_t rank[kSortItems];
    bool valid[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + wave * (64 * kSortItems) + r * 64 + lane;
        valid[r] = i < n;
        key[r] = valid[r] ? kin[i] : 0u;
        val[r] = valid[r] ? vin[i] : 0;
    }
    const uint64_t lt_mask = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t d = (key[r] >> shift) & (kRadix - 1);
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

    // Thread t owns digit t: per-wave exclusive offsets, tile count, tile scan.
    const uint32_t dg = (uint32_t)tid;
    uint32_t cnt = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const uint32_t c = s_wave[w][dg];
        s_wave[w][dg] = cnt;
        cnt += c;
    }
    uint32_t tile_total;
    const uint32_t excl = block_exclusive_scan(cnt, s_scan, &tile_total);
    s_excl[dg] = excl;

    // Decoupled look-back over earlier tiles for this digit.
    uint32_t *my = status + (size_t)tile * kRadix + dg;
    uint32_t prefix = 0;
    if (tile == 0) {
        __hip_atomic_store(my, kFlagInc | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(my, kFlagAgg | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int32_t j = (int32_t)tile - 1;
        while (j >= 0) {
            const uint32_t s = __hip_atomic_load(status + (size_t)j * kRadix + dg, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t flag = s & ~kValMask;
            if (flag == 0u) {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            prefix += s & kValMask;
            if (flag == kFlagInc) break;
            --j;
        }
        __hip_atomic_store(my, kFlagInc | (prefix + cnt), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // Global start of this digit: exclusive scan of the whole-array histogram.
    const uint32_t gexcl = block_exclusive_scan(hist[dg], s_scan, nullptr);
    s_gofs[dg] = gexcl + prefix;
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        if (valid[r]) {
            const uint32_t d = (key[r] >> shift) & (kRadix - 1);
            const uint32_t lp = s_excl[d] + s_wave[wave][d] + rank[r];
            s_keys[lp] = key[r];
            s_vals[lp] = val[r];
        }
    }
    __syncthreads();
    const uint32_t tile_n = (n - base) < (uint32_t)kSortTile ? (n - base) : (uint32_t)kSortTile;
    for (uint32_t i = tid; i < tile_n; i += kSortThreads) {
        const uint32_t k = s_keys[i];
        const uint32_t d = (k >> shift) & (kRadix - 1);
        const uint32_t dst = s_gofs[d] + (i - s_excl[d]);
        kout[dst] = k;
        vout[dst] = s_vals[i];
    }
}

// ---------------------------------------------------------------------------
// Stage 4: heap emission of one tree level (LBVH_Build, LBVH.hlsl:134-226),
// over the compacted frontier of that level instead of all 2^level slots.
// Frontier entry: {heap index, first, count, parent} with parent = -1 for the
// root, else (k << 6) | (side << 5) | split bit of the parent.
// Leaf rule and axis bits are those of build_recursive (LBVH.cpp:194-198,
// 231-266, 304-310): a node entered under a split at bit p becomes a leaf when
// count <= 1, p == 0 or node >= 2^(depth-1), with axis (31 - p) % 3; a node
// whose codes are all equal is a leaf with axis 1; an internal node split at
// bit p has axis (30 - p) % 3.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_emit_level(
    const uint32_t *__restrict__ codes, const int4 *__restrict__ fin, int4 *__restrict__ fout,
    uint32_t *__restrict__ counters, int level, uint32_t depth, rt_node32 *__restrict__ heap,
    TravNode *__restrict__ trav, int32_t *__restrict__ trav_parent, uint32_t *__restrict__ flags,
    int4 *__restrict__ leaves, SceneHdr *__restrict__ hdr)
{
    const uint32_t n_in = counters[level];
    const uint32_t leaf_level = 1u << (depth - 1);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t b = blockIdx.x * blockDim.x; b < n_in; b += stride) {
        const uint32_t i = b + threadIdx.x;
        const bool active = i < n_in;
        int4 e = active ? fin[i] : make_int4(0, 0, 0, 0);
        const uint32_t h = (uint32_t)e.x;
        const int32_t first = e.y, cnt = e.z, pr = e.w;
        const int32_t last = first + cnt - 1;
        const int pp = pr < 0 ? 31 : (pr & 31);
        bool leaf = false, internal = false;
        int axis = 0, split_bit = 0;
        uint32_t x = 0;
        if (active) {
            const bool at_entry = cnt <= 1 || h >= leaf_level || pp == 0;
            if (at_entry) {
                leaf = true;
                axis = (31 - pp) % 3;
            } else {
                x = codes[first] ^ codes[last];
                if (x == 0u) {
                    leaf = true;
                    axis = 1;
                } else {
                    internal = true;
                    split_bit = 31 - __clz(x);
                    axis = (30 - split_bit) % 3;
                }
            }
        }
        int32_t s = 0;
        if (internal) {
            int32_t lo = first, hi = last;
            while (hi - lo > 1) {
                const int32_t mid = (lo + hi) >> 1;
                if ((codes[mid] >> split_bit) & 1u) hi = mid;
                else lo = mid;
            }
            s = hi;
        }
        const uint32_t child_slot = wave_append(&counters[level + 1], internal, 2u);
        const uint32_t leaf_slot = wave_append(&counters[kCntLeaves], leaf, 1u);
        if (!active) continue;
        rt_node32 *hn = &heap[h];
        int32_t my_ref;
        if (leaf) {
            hn->prim_count = (cnt << 3) | 4 | axis;
            hn->prim_pos = first;
            my_ref = (int32_t)(kLeafBit | (uint32_t)first);
            leaves[leaf_slot] = make_int4((int)h, first, cnt, pr);
            if (h < leaf_level) {  // mark the two unused child slots, as LBVH.hlsl:166-176
                heap[2 * h].prim_count = kInvalid;
                heap[2 * h].prim_pos = kInvalid;
                heap[2 * h + 1].prim_count = kInvalid;
                heap[2 * h + 1].prim_pos = kInvalid;
            }
        } else {
            const int32_t k = s - 1;
            hn->prim_count = (cnt << 3) | axis;
            hn->prim_pos = first;
            my_ref = k;
            TravNode *tn = &trav[k];
            tn->axis = axis;
            tn->heap = (int32_t)h;
            trav_parent[k] = pr;
            flags[k] = 0u;
            fout[child_slot] = make_int4((int)(2 * h), first, s - first, (k << 6) | split_bit);
            fout[child_slot + 1] =
                make_int4((int)(2 * h + 1), s, last - s + 1, (k << 6) | (1 << 5) | split_bit);
        }
        if (pr < 0) hdr->root_ref = my_ref;
        else trav[pr >> 6].ref[(pr >> 5) & 1] = my_ref;
    }
}

// ---------------------------------------------------------------------------
// Stage 5: refit (LBVH_BuildAABB, LBVH.hlsl:228-270) bottom-up from the leaves.
// Each leaf box is the min/max of its primitives' boxes; at every internal
// node the second child to arrive (device-scope atomic counter) merges both
// children and climbs on, so each node is written exactly once.  Min/max is
// exact, so the boxes equal the CPU builder's (LBVH.cpp:200-206, 312-313).
// The leaf pass also lays the triangles out in sorted order for traversal.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_refit(
    const int4 *__restrict__ leaves, const uint32_t *__restrict__ counters,
    const int32_t *__restrict__ ids, const PrimBox *__restrict__ pbox,
    const float *__restrict__ xyz, const uint32_t *__restrict__ idx, rt_node32 *heap,
    TravNode *trav, const int32_t *__restrict__ trav_parent, uint32_t *flags,
    TravTri *__restrict__ tris, SceneHdr *hdr)
{
    const uint32_t nleaves = counters[kCntLeaves];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nleaves; i += stride) {
        const int4 e = leaves[i];
        const int32_t h = e.x, first = e.y, cnt = e.z;
        int32_t pr = e.w;
        float lo[3] = {INFINITY, INFINITY, INFINITY};
        float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int32_t j = 0; j < cnt; ++j) {
            const int32_t id = ids[first + j];
            const PrimBox b = pbox[id];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                lo[k] = fminf(lo[k], b.mn[k]);
                hi[k] = fmaxf(hi[k], b.mx[k]);
            }
            const uint32_t ia = idx[3 * (size_t)id + 0], ib = idx[3 * (size_t)id + 1],
                           ic = idx[3 * (size_t)id + 2];
            const float ax = xyz[3 * (size_t)ia + 0], ay = xyz[3 * (size_t)ia + 1],
                        az = xyz[3 * (size_t)ia + 2];
            TravTri t;
            t.a = make_float4(ax, ay, az, __int_as_float(id));
            t.e1 = make_float4(xyz[3 * (size_t)ib + 0] - ax, xyz[3 * (size_t)ib + 1] - ay,
                               xyz[3 * (size_t)ib + 2] - az, __int_as_float(j == 0 ? cnt : 0));
            t.e2 = make_float4(xyz[3 * (size_t)ic + 0] - ax, xyz[3 * (size_t)ic + 1] - ay,
                               xyz[3 * (size_t)ic + 2] - az, __int_as_float(j == 0 ? h : 0));
            tris[first + j] = t;
        }
        rt_node32 *hn = &heap[h];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            hn->mn[k] = lo[k];
            hn->mx[k] = hi[k];
        }
        while (true) {
            if (pr < 0) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    hdr->root_box[k] = lo[k];
                    hdr->root_box[3 + k] = hi[k];
                }
                break;
            }
            const int32_t k = pr >> 6, side = (pr >> 5) & 1;
            TravNode *tn = &trav[k];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                tn->box[side * 6 + c] = lo[c];
                tn->box[side * 6 + 3 + c] = hi[c];
            }
            __threadfence();
            const uint32_t old = atomicAdd(&flags[k], 1u);
            if (old == 0u) break;
            __threadfence();
            const int other = 1 - side;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                lo[c] = fminf(lo[c], tn->box[other * 6 + c]);
                hi[c] = fmaxf(hi[c], tn->box[other * 6 + 3 + c]);
            }
            rt_node32 *pn = &heap[tn->heap];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                pn->mn[c] = lo[c];
                pn->mx[c] = hi[c];
            }
            pr = trav_parent[k];
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
        const float s = a + b;
        const float t = s + c;
        xyz[3 * (size_t)i + r] = t + m[9 + r];
    }
}

__global__ void k_empty_scene(SceneHdr *hdr)
{
    if (threadIdx.x == 0) {
        hdr->root_ref = -1;
        hdr->nleaves = 0;
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
uint32_t sort_tiles(uint32_t n) { return (n + kSortTile - 1) / kSortTile; }
uint32_t bounds_blocks(uint32_t n) { return (n + kBoundsTile - 1) / kBoundsTile; }

static inline uint32_t min_u32(uint32_t a, uint32_t b) { return a < b ? a : b; }

hipError_t launch_transform(Scene &s, const Xform &x, hipStream_t st)
{
    if (s.nv == 0) return hipSuccess;
    hipLaunchKernelGGL(k_transform, dim3((s.nv + 255) / 256), dim3(256), 0, st, s.xyz_rest, s.xyz,
                       s.nv, x);
    return hipGetLastError();
}

hipError_t launch_build(Scene &s, hipStream_t st, hipEvent_t *ev)
{
    const uint32_t nt = s.nt;
    if (ev) hipEventRecord(ev[0], st);
    if (nt == 0) {
        hipLaunchKernelGGL(k_empty_scene, dim3(1), dim3(64), 0, st, s.hdr);
        if (ev)
            for (int i = 1; i < 6; ++i) hipEventRecord(ev[i], st);
        return hipGetLastError();
    }
    const uint32_t nb = bounds_blocks(nt);
    const uint32_t tiles = sort_tiles(nt);
    // hist | counters | status are one contiguous block of words
    const uint32_t nzero = kSortPasses * kRadix + kCounters + kSortPasses * tiles * kRadix;
    hipLaunchKernelGGL(k_bounds, dim3(nb), dim3(kBoundsThreads), 0, st, s.xyz, s.idx, nt, s.pbox,
                       s.cent, s.partials, s.hist, nzero);
    if (ev) hipEventRecord(ev[1], st);
    hipLaunchKernelGGL(k_morton, dim3(nb), dim3(kBoundsThreads), 0, st, s.cent, nt, s.partials, nb,
                       s.keys[0], s.vals[0], s.hist, s.hdr, s.frontier[0], s.counters);
    if (ev) hipEventRecord(ev[2], st);
    for (int p = 0; p < kSortPasses; ++p) {
        const int src = p & 1, dst = src ^ 1;
        hipLaunchKernelGGL(k_onesweep, dim3(tiles), dim3(kSortThreads), 0, st, s.keys[src],
                           s.vals[src], s.keys[dst], s.vals[dst], nt, (uint32_t)(p * kRadixBits),
                           s.hist + p * kRadix, s.status + (size_t)p * tiles * kRadix,
                           s.counters + kCntTiles + p);
    }
    if (ev) hipEventRecord(ev[3], st);
    for (uint32_t level = 0; level < s.depth; ++level) {
        const uint32_t width = level < 31 ? min_u32(1u << level, nt) : nt;
        const uint32_t blocks = min_u32((width + 255) / 256, 2048u);
        hipLaunchKernelGGL(k_emit_level, dim3(blocks), dim3(256), 0, st, s.keys[0],
                           s.frontier[level & 1], s.frontier[(level + 1) & 1], s.counters,
                           (int)level, s.depth, s.heap, s.trav, s.trav_parent, s.flags, s.leaves,
                           s.hdr);
    }
    if (ev) hipEventRecord(ev[4], st);
    {
        const uint32_t blocks = min_u32((nt + 255) / 256, 4096u);
        hipLaunchKernelGGL(k_refit, dim3(blocks), dim3(256), 0, st, s.leaves, s.counters, s.vals[0],
                           s.pbox, s.xyz, s.idx, s.heap, s.trav, s.trav_parent, s.flags, s.tris,
                           s.hdr);
    }
    if (ev) hipEventRecord(ev[5], st);
    return hipGetLastError();
}

}  // namespace rt
