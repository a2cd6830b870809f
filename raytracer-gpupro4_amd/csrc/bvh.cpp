// Second acceleration structure: the PBRT-derived SAH BVH of the reference
// (BVH::Build / RecursiveBuild / FlattenBVHTree,
// AccelerationStructures/BVH/bvh.cpp:64-349, bvh.h:33-101), built on the
// host as the reference builds it, flattened depth-first into its 32-byte
// LinearBVHNode (= rt_node32), then laid out for the device traversal.
//
// Restated rules (float arithmetic in the reference's operations and order):
//   - primitive box = Union(BBox(A, B), C); centroid = .5 pMin + .5 pMax;
//     empty boxes are (+inf, -inf); SurfaceArea = 2 (dx dy + dx dz + dy dz);
//     MaximumExtent: x if dx > dy and dx > dz, else y if dy > dz, else z;
//   - a node of exactly max_prims primitives is a leaf; with max_prims > 1 a
//     node whose centroids coincide on the split axis is a leaf;
//   - SAH: 4 or fewer primitives split at the median of the centroids on the
//     axis of largest centroid extent; otherwise 12 buckets over the centroid
//     bounds (b = int(12 ((c - min) / (max - min))), 12 -> 11), cost of a split
//     after bucket i = .125 + (n0 SA0 + n1 SA1) / SA, the first minimum wins
//     (IEEE comparisons, so a NaN cost is never taken over cost[0]), and the
//     node splits when it has more than max_prims primitives or the cost is
//     below its primitive count; MIDDLE splits at the centroid-bounds
//     midpoint, EQUAL at the median;
//   - partitions use the two-ended swapping scheme of std::partition (the
//     scheme of both the MSVC and GNU libraries); medians use a stable order
//     by centroid (the insertion sort MSVC's nth_element runs on small
//     ranges), so only exactly tied centroids can order differently;
//   - children are built left then right and flattened in pre-order: the
//     first child follows its parent, the second's index is stored in the
//     parent (secondChildOffset); inner nodes store -axis in nPrimitives.
// Deviations, each where the reference has undefined behaviour or cannot run:
//   - more than 4 primitives with coincident centroids (a 0 / 0 bucket index)
//     split at the median of their order;
//   - a MIDDLE split that leaves one side empty falls back to the median (PBRT
//     v2 does this; the reference's variant recurses on an empty range);
//   - the depth of a leaf is kept at most 26 below the root (the device stack
//     holds 26 entries; the reference's shader stack holds 35 and overflows
//     beyond): a split that would leave a child with more primitives than a
//     balanced subtree can hold in the remaining levels is replaced by the
//     median split, which keeps level + ceil(log2 n) <= 26 from the root
//     down.  rt_bvh_info.forced_medians counts such nodes.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "rt_internal.h"

namespace rt {

namespace {

struct Box {
    float mn[3], mx[3];
};

inline Box empty_box()
{
    const float inf = std::numeric_limits<float>::infinity();
    return Box{{inf, inf, inf}, {-inf, -inf, -inf}};
}

inline void box_union(Box &a, const Box &b)
{
    for (int k = 0; k < 3; ++k) {
        a.mn[k] = std::min(a.mn[k], b.mn[k]);
        a.mx[k] = std::max(a.mx[k], b.mx[k]);
    }
}

inline void box_add(Box &a, const float *p)
{
    for (int k = 0; k < 3; ++k) {
        a.mn[k] = std::min(a.mn[k], p[k]);
        a.mx[k] = std::max(a.mx[k], p[k]);
    }
}

inline float surface_area(const Box &b)
{
    const float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
    return 2.f * ((dx * dy + dx * dz) + dy * dz);
}

inline int max_extent(const Box &b)
{
    const float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
    if (dx > dy && dx > dz) return 0;
    if (dy > dz) return 1;
    return 2;
}

struct Prim {
    uint32_t id;
    float c[3];
    Box b;
};

// std::partition's two-ended scheme: returns the first element for which
// pred is false after swapping out-of-place pairs from both ends.
template <class Pred>
Prim *two_ended_partition(Prim *first, Prim *last, Pred pred)
{
    while (true) {
        while (true) {
            if (first == last) return first;
            if (pred(*first)) ++first;
            else break;
        }
        --last;
        while (true) {
            if (first == last) return first;
            if (!pred(*last)) --last;
            else break;
        }
        std::swap(*first, *last);
        ++first;
    }
}

float bits_float(int32_t x)
{
    float f;
    memcpy(&f, &x, 4);
    return f;
}

int ceil_log2(uint32_t n)
{
    int k = 0;
    while ((1ull << k) < n) ++k;
    return k;
}

constexpr int kMaxLeafLevel = 26;

struct Builder {
    std::vector<Prim> prims;
    std::vector<rt_node32> nodes;
    std::vector<int32_t> ordered;
    uint32_t max_prims = 1;
    uint32_t split = RT_BVH_SAH;
    uint32_t leaves = 0, max_level = 0, forced = 0;

    void leaf(uint32_t ni, uint32_t start, uint32_t end, const Box &bb, uint32_t level)
    {
        rt_node32 &n = nodes[ni];
        memcpy(n.mn, bb.mn, 12);
        memcpy(n.mx, bb.mx, 12);
        n.prim_count = (int32_t)(end - start);
        n.prim_pos = (int32_t)start;
        for (uint32_t i = start; i < end; ++i) ordered[i] = (int32_t)prims[i].id;
        ++leaves;
        max_level = std::max(max_level, level);
    }

    void median(uint32_t start, uint32_t end, int dim)
    {
        std::stable_sort(prims.begin() + start, prims.begin() + end,
                         [dim](const Prim &a, const Prim &b) { return a.c[dim] < b.c[dim]; });
    }

    // Returns the node's box (the union of its children's, or its primitives').
    Box build(uint32_t start, uint32_t end, uint32_t level)
    {
        const uint32_t ni = (uint32_t)nodes.size();
        nodes.push_back(rt_node32{});
        Box bb = empty_box();
        for (uint32_t i = start; i < end; ++i) box_union(bb, prims[i].b);
        const uint32_t n = end - start;
        if (n == max_prims) {
            leaf(ni, start, end, bb, level);
            return bb;
        }
        Box cb = empty_box();
        for (uint32_t i = start; i < end; ++i) box_add(cb, prims[i].c);
        const int dim = max_extent(cb);
        uint32_t mid = (start + end) / 2;
        if (max_prims > 1 && cb.mx[dim] == cb.mn[dim]) {
            leaf(ni, start, end, bb, level);
            return bb;
        }
        if (split == RT_BVH_MIDDLE) {
            const float pmid = .5f * (cb.mn[dim] + cb.mx[dim]);
            Prim *p = two_ended_partition(&prims[start], &prims[0] + end,
                                          [dim, pmid](const Prim &a) { return a.c[dim] < pmid; });
            mid = (uint32_t)(p - &prims[0]);
            if (mid == start || mid == end) {
                mid = (start + end) / 2;
                median(start, end, dim);
            }
        } else if (split == RT_BVH_EQUAL || n <= 4) {
            median(start, end, dim);
        } else if (!(cb.mx[dim] > cb.mn[dim])) {
            // coincident centroids: the bucket index would be 0 / 0
            mid = (start + end) / 2;
        } else {
            const float inv_sah = 1 / surface_area(bb);
            constexpr int kBuckets = 12;
            int count[kBuckets] = {};
            Box bounds[kBuckets];
            for (int k = 0; k < kBuckets; ++k) bounds[k] = empty_box();
            const float lo = cb.mn[dim], ext = cb.mx[dim] - cb.mn[dim];
            auto bucket = [&](const Prim &p) {
                int b = (int)(kBuckets * ((p.c[dim] - lo) / ext));
                return b == kBuckets ? kBuckets - 1 : b;
            };
            for (uint32_t i = start; i < end; ++i) {
                const int b = bucket(prims[i]);
                ++count[b];
                box_union(bounds[b], prims[i].b);
            }
            float cost[kBuckets - 1];
            for (int i = 0; i < kBuckets - 1; ++i) {
                Box b0 = empty_box(), b1 = empty_box();
                int c0 = 0, c1 = 0;
                for (int j = 0; j <= i; ++j) {
                    box_union(b0, bounds[j]);
                    c0 += count[j];
                }
                for (int j = i + 1; j < kBuckets; ++j) {
                    box_union(b1, bounds[j]);
                    c1 += count[j];
                }
                cost[i] = .125f + ((float)c0 * surface_area(b0) + (float)c1 * surface_area(b1)) * inv_sah;
            }
            float min_cost = cost[0];
            int best = 0;
            for (int i = 1; i < kBuckets - 1; ++i)
                if (cost[i] < min_cost) {
                    min_cost = cost[i];
                    best = i;
                }
            if (n > max_prims || min_cost < (float)n) {
                Prim *p = two_ended_partition(&prims[start], &prims[0] + end,
                                              [&](const Prim &q) { return bucket(q) <= best; });
                mid = (uint32_t)(p - &prims[0]);
            } else {
                leaf(ni, start, end, bb, level);
                return bb;
            }
        }
        // keep level + ceil(log2 n) <= 26 for both children (it holds at the
        // root for n <= 2^26, and a median split preserves it)
        if ((int)level + 1 + ceil_log2(std::max(mid - start, end - mid)) > kMaxLeafLevel || mid == start ||
            mid == end) {
            ++forced;
            median(start, end, dim);
            mid = (start + end) / 2;
        }
        const Box b0 = build(start, mid, level + 1);
        const uint32_t second = (uint32_t)nodes.size();
        const Box b1 = build(mid, end, level + 1);
        Box u = b0;
        box_union(u, b1);
        rt_node32 &nd = nodes[ni];
        memcpy(nd.mn, u.mn, 12);
        memcpy(nd.mx, u.mx, 12);
        nd.prim_count = -dim;
        nd.prim_pos = (int32_t)second;
        return u;
    }
};

}  // namespace

int bvh_build_host(const float *xyz, const uint32_t *idx, uint32_t nt, uint32_t max_prims, uint32_t split,
                   BvhHost &out, rt_bvh_info *info)
{
    const auto t0 = std::chrono::steady_clock::now();
    Builder B;
    B.max_prims = std::min(255u, max_prims);
    B.split = split;
    B.prims.resize(nt);
    B.ordered.assign(nt, -1);
    B.nodes.reserve(2 * (size_t)nt);
    for (uint32_t i = 0; i < nt; ++i) {
        const float *A = xyz + 3 * (size_t)idx[3 * (size_t)i];
        const float *Bv = xyz + 3 * (size_t)idx[3 * (size_t)i + 1];
        const float *C = xyz + 3 * (size_t)idx[3 * (size_t)i + 2];
        Box b;
        for (int k = 0; k < 3; ++k) {
            b.mn[k] = std::min(A[k], Bv[k]);
            b.mx[k] = std::max(A[k], Bv[k]);
        }
        box_add(b, C);
        Prim &p = B.prims[i];
        p.id = i;
        p.b = b;
        for (int k = 0; k < 3; ++k) p.c[k] = .5f * b.mn[k] + .5f * b.mx[k];
    }
    B.build(0, nt, 0);
    out.nodes.swap(B.nodes);
    out.ordered.swap(B.ordered);
    out.max_level = B.max_level;
    if (info) {
        info->nodes = (uint32_t)out.nodes.size();
        info->leaves = B.leaves;
        info->max_depth = B.max_level + 1;
        info->forced_medians = B.forced;
        // SAH cost of the tree: 1 per inner node and 1 per primitive, weighted
        // by the surface area relative to the root's
        const rt_node32 &r = out.nodes[0];
        const Box rb{{r.mn[0], r.mn[1], r.mn[2]}, {r.mx[0], r.mx[1], r.mx[2]}};
        const double ra = surface_area(rb);
        double cost = 0.0;
        for (const rt_node32 &nd : out.nodes) {
            const Box b{{nd.mn[0], nd.mn[1], nd.mn[2]}, {nd.mx[0], nd.mx[1], nd.mx[2]}};
            const double a = ra > 0 ? surface_area(b) / ra : 1.0;
            cost += nd.prim_count > 0 ? a * nd.prim_count : a * 0.125;
        }
        info->sah_cost = (float)cost;
        info->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return RT_OK;
}

// Device layout of a built BVH: one TravNode per inner node (both children's
// boxes and refs; inner nodes numbered in pre-order), a TravTri per
// primitive in leaf order, and the root in a SceneHdr.
void bvh_device_layout(const float *xyz, const uint32_t *idx, const BvhHost &h, std::vector<TravNode> &trav,
                       std::vector<TravTri> &tris, SceneHdr &hdr)
{
    const size_t nn = h.nodes.size();
    std::vector<int32_t> inner(nn, -1);
    int32_t k = 0;
    for (size_t i = 0; i < nn; ++i)
        if (h.nodes[i].prim_count <= 0) inner[i] = k++;
    auto ref_of = [&](uint32_t i) -> int32_t {
        const rt_node32 &n = h.nodes[i];
        return n.prim_count > 0 ? (int32_t)leaf_ref((uint32_t)n.prim_pos, (uint32_t)n.prim_count) : inner[i];
    };
    trav.assign(k > 0 ? (size_t)k : 1, TravNode{});
    for (size_t i = 0; i < nn; ++i) {
        const rt_node32 &n = h.nodes[i];
        if (n.prim_count > 0) continue;
        const uint32_t c0 = (uint32_t)i + 1, c1 = (uint32_t)n.prim_pos;
        TravNode &t = trav[inner[i]];
        memcpy(&t.box[0], h.nodes[c0].mn, 12);
        memcpy(&t.box[3], h.nodes[c0].mx, 12);
        memcpy(&t.box[6], h.nodes[c1].mn, 12);
        memcpy(&t.box[9], h.nodes[c1].mx, 12);
        t.ref[0] = ref_of(c0);
        t.ref[1] = ref_of(c1);
        t.axis = -n.prim_count;
        t.heap = (int32_t)i;
    }
    const size_t nt = h.ordered.size();
    tris.assign(nt, TravTri{});
    for (size_t i = 0; i < nn; ++i) {
        const rt_node32 &n = h.nodes[i];
        if (n.prim_count <= 0) continue;
        for (int32_t j = 0; j < n.prim_count; ++j) {
            const size_t pos = (size_t)n.prim_pos + j;
            const uint32_t id = (uint32_t)h.ordered[pos];
            const float *A = xyz + 3 * (size_t)idx[3 * (size_t)id];
            const float *Bv = xyz + 3 * (size_t)idx[3 * (size_t)id + 1];
            const float *C = xyz + 3 * (size_t)idx[3 * (size_t)id + 2];
            TravTri &t = tris[pos];
            t.a = make_float4(A[0], A[1], A[2], bits_float((int32_t)id));
            t.e1 = make_float4(Bv[0] - A[0], Bv[1] - A[1], Bv[2] - A[2], bits_float(j == 0 ? n.prim_count : 0));
            t.e2 = make_float4(C[0] - A[0], C[1] - A[1], C[2] - A[2], bits_float(j == 0 ? (int32_t)i : 0));
        }
    }
    memset(&hdr, 0, sizeof(hdr));
    hdr.root_ref = nn ? ref_of(0) : -1;
    hdr.levels = 0;
    if (nn) {
        memcpy(&hdr.root_box[0], h.nodes[0].mn, 12);
        memcpy(&hdr.root_box[3], h.nodes[0].mx, 12);
        memcpy(&hdr.scene_box[0], h.nodes[0].mn, 12);
        memcpy(&hdr.scene_box[3], h.nodes[0].mx, 12);
    }
}

}  // namespace rt
