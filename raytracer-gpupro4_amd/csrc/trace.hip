// Ray traversal on gfx950: closest-hit and any-hit over the LBVH, primary-ray
// generation, direct shading with shadow rays and diffuse path tracing.
//
// The visit order is the reference's stackless trail traversal
// (LBVH_IntersectP, LBVH.hlsl:363-440): depth first, at an internal node the
// child on side dirIsNeg[axis] first, a box culled when T0 > best t or T1 < 0,
// leaf triangles tested first to last, a hit taken only with t < best t.
// Here a node holds both child boxes, so both are tested when the node is
// entered; the second child is kept on a per-lane stack with its T0 and
// culled on pop against the then-current best t (an entry whose T0 already
// exceeds the best t is not pushed at all: best t only shrinks, so the pop
// would discard it without counting anything).  That is the same decision,
// at the same point of the sequence, as the trail algorithm makes when it
// reaches the sibling, so the hits, the tie-breaking and the visit counters
// (+1 per box, +65536 per triangle) equal the reference's.
// Float arithmetic is IEEE single with no contraction (-ffp-contract=off):
// IntersectBox and getIntersection (Intersection.hlsl:18-64) in source order.
//
// The stack keeps its first KLDS entries in LDS (one column per lane) and
// any pushed beyond them in private memory; KLDS and the block size are
// launch-time tunables (occupancy against spill traffic).  rt_trace can start
// a ray at any heap node, as LBVH_IntersectP takes its iRoot argument
// (trav_enter).  Frame kernels take their tiles heaviest first, from the
// last frame's per-tile costs (k_tile_order), and direct frames may trace the
// heaviest tiles with four lanes per pixel (quads, traverse_quad).
#include <algorithm>
#include <cstring>

#include "rt_internal.h"

namespace rt {

constexpr int kStackMax = 26;  // >= depth - 1 for depth <= 27

struct TraceOut {
    int32_t tri;
    float u, v, t;
    uint32_t boxes, tris;
    int32_t leaf_heap;
    int32_t pos;  // sorted slot of the hit triangle, -1 when none
};

// Stacks of at most 16 LDS entries spill deeper entries to private memory.
// Larger ones (20, 22, 26) hold the whole stack in LDS with no spill path: a
// traversal pushes at most one entry per level below the root, so a heap of
// depth d needs d - 1 entries, and the host only launches a shape whose stack
// covers the deepest heap of the launch (26 covers every allowed depth).
constexpr bool stack_spills(int klds) { return klds <= 16; }

template <int BLOCK, int KLDS>
struct RayStack {
    static constexpr bool kSpill = stack_spills(KLDS) && KLDS < kStackMax;
    uint2 *lds;  // this lane's LDS column: entry j at lds[j * BLOCK]
    uint2 spill[kSpill ? (kStackMax - KLDS) : 1];
    int sp;
    __device__ __forceinline__ void push(uint2 e)
    {
        if (!kSpill || sp < KLDS) lds[sp * BLOCK] = e;
        else spill[sp - KLDS] = e;
        ++sp;
    }
    __device__ __forceinline__ uint2 pop()
    {
        --sp;
        return (!kSpill || sp < KLDS) ? lds[sp * BLOCK] : spill[sp - KLDS];
    }
    // the entry pop() would return (entry 0 when empty: a read in bounds)
    __device__ __forceinline__ uint2 peek() const
    {
        const int i = sp > 0 ? sp - 1 : 0;
        return (!kSpill || i < KLDS) ? lds[i * BLOCK] : spill[i - KLDS];
    }
    // pop() with both words of the entry read by one LDS load: the node
    // comes with the T0 that culls it, instead of a second load after the test
    __device__ __forceinline__ uint2 pop_both()
    {
        uint2 e = pop();
        asm volatile("" : "+v"(e.x), "+v"(e.y));
        return e;
    }
};

__device__ __forceinline__ void box_test(const float *bx, float ox, float oy, float oz, float ix,
                                         float iy, float iz, float &t0, float &t1)
{
    const float dmaxx = (bx[3] - ox) * ix, dmaxy = (bx[4] - oy) * iy, dmaxz = (bx[5] - oz) * iz;
    const float dminx = (bx[0] - ox) * ix, dminy = (bx[1] - oy) * iy, dminz = (bx[2] - oz) * iz;
    float a = fminf(dminx, dmaxx);
    float b = fmaxf(dminx, dmaxx);
    a = fmaxf(a, fminf(dminy, dmaxy));
    b = fminf(b, fmaxf(dminy, dmaxy));
    a = fmaxf(a, fminf(dminz, dmaxz));
    b = fminf(b, fmaxf(dminz, dmaxz));
    if (a > b) {
        a = -1.0f;
        b = -1.0f;
    }
    t0 = a;
    t1 = b;
}

// A child box of a node against a ray, for the node step: its entry t0 and
// whether the child is live, i.e. box_test's (t0, t1) passes !(t0 > lim ||
// t1 < 0).  box_test turns a missed box (entry after exit) into (-1, -1),
// which fails that test; here the miss is one more term of it, so the
// caller selects the near child's flag instead of its two distances.  Same
// answer for every input, NaNs included (a comparison with NaN is false in
// both).
__device__ __forceinline__ void child_test(const float *bx, float ox, float oy, float oz, float ix, float iy,
                                           float iz, float lim, float &t0, bool &live)
{
    const float dmaxx = (bx[3] - ox) * ix, dmaxy = (bx[4] - oy) * iy, dmaxz = (bx[5] - oz) * iz;
    const float dminx = (bx[0] - ox) * ix, dminy = (bx[1] - oy) * iy, dminz = (bx[2] - oz) * iz;
    float a = fminf(dminx, dmaxx);
    float b = fmaxf(dminx, dmaxx);
    a = fmaxf(a, fminf(dminy, dmaxy));
    b = fminf(b, fmaxf(dminy, dmaxy));
    a = fmaxf(a, fminf(dminz, dmaxz));
    b = fminf(b, fmaxf(dminz, dmaxz));
    t0 = a;
    live = !(a > b) && !(a > lim) && !(b < 0.0f);
}

// One triangle test: getIntersection (Intersection.hlsl:47-64) and the
// acceptance rule of LBVH_IntersectP (LBVH.hlsl:400-420).  Returns true when
// the triangle becomes the current hit.
__device__ __forceinline__ bool tri_test(TraceOut &r, const float4 a, const float4 e1, const float4 e2, float ox,
                                         float oy, float oz, float dx, float dy, float dz, int32_t self,
                                         int32_t heap, int32_t pos)
{
    r.tris += 1;
    const int32_t id = __float_as_int(a.w);
    const float px = (dy * e2.z) - (dz * e2.y);
    const float py = (dz * e2.x) - (dx * e2.z);
    const float pz = (dx * e2.y) - (dy * e2.x);
    const float dd = (e1.x * px + e1.y * py) + e1.z * pz;
    const float det = 1.0f / dd;
    const float tx = ox - a.x, ty = oy - a.y, tz = oz - a.z;
    const float u = ((tx * px + ty * py) + tz * pz) * det;
    const float qx = (ty * e1.z) - (tz * e1.y);
    const float qy = (tz * e1.x) - (tx * e1.z);
    const float qz = (tx * e1.y) - (ty * e1.x);
    const float v = ((dx * qx + dy * qy) + dz * qz) * det;
    const float t = ((e2.x * qx + e2.y * qy) + e2.z * qz) * det;
    if (id != self && u >= 0.0f && v >= 0.0f && (u + v) <= 1.0f && t >= 0.0f && t < r.t) {
        r.tri = id;
        r.u = u;
        r.v = v;
        r.t = t;
        r.leaf_heap = heap;
        r.pos = pos;
        return true;
    }
    return false;
}

// The triangles of a leaf, in order.  The count hint in the reference lets
// the loads of up to four triangles go out together, before any test; longer
// leaves continue in chunks of four.  Returns true when an any-hit query
// accepted a triangle.
__device__ __forceinline__ bool leaf_test(TraceOut &r, const TravTri *__restrict__ tris, uint32_t ref, float ox,
                                          float oy, float oz, float dx, float dy, float dz, int32_t self,
                                          bool anyhit)
{
    constexpr int kChunk = 4;
    const uint32_t pos = ref & kLeafPosMask;
    const uint32_t hint = (ref >> kLeafCountShift) & 7u;
    const TravTri *tp = &tris[pos];
    int32_t cnt = (int32_t)hint;  // exact below the cap; the cap reads the count from the first record
    int32_t heap = 0;
    for (int32_t j = 0; j < cnt; j += kChunk) {
        const int32_t m = cnt - j < kChunk ? cnt - j : kChunk;
        float4 A[kChunk], B[kChunk], C[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) {
            if (k < m) {
                A[k] = tp[j + k].a;
                B[k] = tp[j + k].e1;
                C[k] = tp[j + k].e2;
            }
        }
        if (j == 0) {
            heap = __float_as_int(C[0].w);
            if (hint == kLeafCountCap) cnt = __float_as_int(B[0].w);
        }
#pragma unroll
        for (int k = 0; k < kChunk; ++k) {
            if (k < m && tri_test(r, A[k], B[k], C[k], ox, oy, oz, dx, dy, dz, self, heap, (int32_t)pos + j + k) &&
                anyhit)
                return true;
        }
    }
    return false;
}

// leaf_test with the leaf's first triangle (A0, B0, C0) already fetched by
// fetch_step; the rest follow in chunks of four, tested in the same order.
__device__ __forceinline__ bool leaf_test_first(TraceOut &r, const TravTri *__restrict__ tris, uint32_t ref,
                                                const float4 A0, const float4 B0, const float4 C0, float ox,
                                                float oy, float oz, float dx, float dy, float dz, int32_t self,
                                                bool anyhit)
{
    constexpr int kChunk = 4;
    const uint32_t pos = ref & kLeafPosMask;
    const uint32_t hint = (ref >> kLeafCountShift) & 7u;
    const int32_t cnt = hint == kLeafCountCap ? __float_as_int(B0.w) : (int32_t)hint;
    const int32_t heap = __float_as_int(C0.w);
    if (tri_test(r, A0, B0, C0, ox, oy, oz, dx, dy, dz, self, heap, (int32_t)pos) && anyhit) return true;
    const TravTri *tp = &tris[pos];
    for (int32_t j = 1; j < cnt; j += kChunk) {
        const int32_t m = cnt - j < kChunk ? cnt - j : kChunk;
        float4 A[kChunk], B[kChunk], C[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) {
            if (k < m) {
                A[k] = tp[j + k].a;
                B[k] = tp[j + k].e1;
                C[k] = tp[j + k].e2;
            }
        }
#pragma unroll
        for (int k = 0; k < kChunk; ++k) {
            if (k < m && tri_test(r, A[k], B[k], C[k], ox, oy, oz, dx, dy, dz, self, heap, (int32_t)pos + j + k) &&
                anyhit)
                return true;
        }
    }
    return false;
}

// One fetch per iteration for every lane: the 64-byte node of a lane at an
// internal node, or the first triangle record of a lane at a leaf (its fourth
// quad re-reads the record's first).  Lanes at nodes and lanes at leaves then
// wait on the same loads, so a wave holding both kinds pays one memory round
// trip per iteration instead of one per branch.  The empty asm reads the
// last quad, which keeps all four loads ahead of the branch.
__device__ __forceinline__ void fetch_step(const TravNode *__restrict__ nodes, const TravTri *__restrict__ tris,
                                           int32_t cur, float4 &q0, float4 &q1, float4 &q2, float4 &q3)
{
    const bool at_node = cur >= 0;
    const float4 *p = at_node ? reinterpret_cast<const float4 *>(&nodes[cur])
                              : reinterpret_cast<const float4 *>(&tris[(uint32_t)cur & kLeafPosMask]);
    const float4 *p3 = at_node ? p + 3 : p;
    q0 = p[0];
    q1 = p[1];
    q2 = p[2];
    q3 = *p3;
    asm volatile("" ::"v"(q3.x), "v"(q3.y), "v"(q3.z), "v"(q3.w));
}

__device__ __forceinline__ int4 as_int4(const float4 f)
{
    return make_int4(__float_as_int(f.x), __float_as_int(f.y), __float_as_int(f.z), __float_as_int(f.w));
}

template <bool ANYHIT, class Stack>
__device__ __forceinline__ TraceOut traverse(const TravNode *__restrict__ nodes,
                                             const TravTri *__restrict__ tris, int32_t root_ref,
                                             const float *root_box, float ox, float oy, float oz,
                                             float dx, float dy, float dz, float tmax, int32_t self,
                                             Stack &stk)
{
    TraceOut r;
    r.tri = -1;
    r.u = -1.0f;
    r.v = -1.0f;
    r.t = tmax;
    r.boxes = 0;
    r.tris = 0;
    r.leaf_heap = 1;
    r.pos = -1;
    if (root_ref == -1) return r;
    const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
    const uint32_t neg = (ix < 0.0f ? 1u : 0u) | (iy < 0.0f ? 2u : 0u) | (iz < 0.0f ? 4u : 0u);
    float t0, t1;
    box_test(root_box, ox, oy, oz, ix, iy, iz, t0, t1);
    r.boxes = 1;
    if (t0 > r.t || t1 < 0.0f) return r;
    int32_t cur = root_ref;
    stk.sp = 0;
    while (true) {
        if (cur >= 0) {
            const TravNode *n = &nodes[cur];
            const float4 q0 = *reinterpret_cast<const float4 *>(&n->box[0]);
            const float4 q1 = *reinterpret_cast<const float4 *>(&n->box[4]);
            const float4 q2 = *reinterpret_cast<const float4 *>(&n->box[8]);
            const int4 q3 = *reinterpret_cast<const int4 *>(&n->ref[0]);
            const float ba[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
            const float bb[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
            float a0, b0;
            bool a_ok, b_ok;
            child_test(ba, ox, oy, oz, ix, iy, iz, r.t, a0, a_ok);
            child_test(bb, ox, oy, oz, ix, iy, iz, r.t, b0, b_ok);
            r.boxes += 2;
            const int first_is_b = (neg >> q3.z) & 1u;  // the child on side dirIsNeg[axis]
            const float s0 = first_is_b ? a0 : b0;
            const int32_t fref = first_is_b ? q3.y : q3.x;
            const int32_t sref = first_is_b ? q3.x : q3.y;
            const bool f_ok = first_is_b ? b_ok : a_ok;
            const bool s_ok = first_is_b ? a_ok : b_ok;
            if (f_ok && s_ok) stk.push(make_uint2((uint32_t)sref, __float_as_uint(s0)));
            if (f_ok || s_ok) {
                cur = f_ok ? fref : sref;
                continue;
            }
        } else {
            if (leaf_test(r, tris, (uint32_t)cur, ox, oy, oz, dx, dy, dz, self, ANYHIT)) return r;
        }
        // pop the next live entry
        bool found = false;
        while (stk.sp > 0) {
            const uint2 e = stk.pop();
            if (!(__uint_as_float(e.y) > r.t)) {
                cur = (int32_t)e.x;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return r;
}

// traverse() with fetch_step: one round trip per iteration for waves whose
// lanes sit at nodes and at leaves, instead of a node fetch and a leaf fetch
// in their own branches.  Same visits, same results.
template <bool ANYHIT, class Stack>
__device__ __forceinline__ TraceOut traverse_shared(const TravNode *__restrict__ nodes,
                                             const TravTri *__restrict__ tris, int32_t root_ref,
                                             const float *root_box, float ox, float oy, float oz,
                                             float dx, float dy, float dz, float tmax, int32_t self,
                                             Stack &stk)
{
    TraceOut r;
    r.tri = -1;
    r.u = -1.0f;
    r.v = -1.0f;
    r.t = tmax;
    r.boxes = 0;
    r.tris = 0;
    r.leaf_heap = 1;
    r.pos = -1;
    if (root_ref == -1) return r;
    const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
    const uint32_t neg = (ix < 0.0f ? 1u : 0u) | (iy < 0.0f ? 2u : 0u) | (iz < 0.0f ? 4u : 0u);
    float t0, t1;
    box_test(root_box, ox, oy, oz, ix, iy, iz, t0, t1);
    r.boxes = 1;
    if (t0 > r.t || t1 < 0.0f) return r;
    int32_t cur = root_ref;
    stk.sp = 0;
    while (true) {
        // the stack top, read while the fetch is in flight: a pop after a
        // node with no live child takes it without another LDS round trip
        const uint2 ptop = stk.peek();
        float4 q0, q1, q2, q3f;
        fetch_step(nodes, tris, cur, q0, q1, q2, q3f);
        if (cur >= 0) {
            const int4 q3 = as_int4(q3f);
            const float ba[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
            const float bb[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
            float a0, b0;
            bool a_ok, b_ok;
            child_test(ba, ox, oy, oz, ix, iy, iz, r.t, a0, a_ok);
            child_test(bb, ox, oy, oz, ix, iy, iz, r.t, b0, b_ok);
            r.boxes += 2;
            const int first_is_b = (neg >> q3.z) & 1u;  // the child on side dirIsNeg[axis]
            const float s0 = first_is_b ? a0 : b0;
            const int32_t fref = first_is_b ? q3.y : q3.x;
            const int32_t sref = first_is_b ? q3.x : q3.y;
            const bool f_ok = first_is_b ? b_ok : a_ok;
            const bool s_ok = first_is_b ? a_ok : b_ok;
            // both children: the first now, the second pushed; one of them:
            // that one; neither: a pop
            if (f_ok && s_ok) stk.push(make_uint2((uint32_t)sref, __float_as_uint(s0)));
            if (f_ok || s_ok) {
                cur = f_ok ? fref : sref;
                continue;
            }
        } else {
            if (leaf_test_first(r, tris, (uint32_t)cur, q0, q1, q2, ox, oy, oz, dx, dy, dz, self, ANYHIT)) return r;
        }
        // pop the next live entry (a leaf's test may have pushed nothing, so
        // the top read before it is still the top)
        if (stk.sp == 0) break;
        --stk.sp;
        if (!(__uint_as_float(ptop.y) > r.t)) {
            cur = (int32_t)ptop.x;
            continue;
        }
        bool found = false;
        while (stk.sp > 0) {
            const uint2 e = stk.pop_both();
            if (!(__uint_as_float(e.y) > r.t)) {
                cur = (int32_t)e.x;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return r;
}

// The start of a traversal at heap node `start` (LBVH_IntersectP's start
// argument, LBVH.hlsl:363-370: 0 and 1 are the root).  The trail algorithm
// tests the start node's box, then, on its way back up, the box of the
// sibling at each level from the start's own up to the root's children.  Here
// the walk goes down from the root along the start's heap bits: each
// off-path sibling's box is tested and counted now and pushed when live (so
// the pops meet them deepest first, culled against the then-current best t,
// as the trail algorithm meets them), then the start node's box is tested.
// A start below a leaf stops at that leaf.  Returns the node to visit first,
// or -1 when nothing is left.
template <class Stack>
__device__ __forceinline__ int32_t trav_enter(const TravNode *__restrict__ nodes, int32_t root_ref,
                                              const float *root_box, uint32_t start, float ox, float oy,
                                              float oz, float ix, float iy, float iz, TraceOut &r, Stack &stk)
{
    stk.sp = 0;
    int32_t ref = root_ref;
    float box[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) box[k] = root_box[k];
    r.boxes = 0;
    if (start > 1u) {
        const int level = 31 - __clz(start);
        for (int k = level - 1; k >= 0 && ref >= 0; --k) {
            const TravNode *n = &nodes[ref];
            const float4 q0 = *reinterpret_cast<const float4 *>(&n->box[0]);
            const float4 q1 = *reinterpret_cast<const float4 *>(&n->box[4]);
            const float4 q2 = *reinterpret_cast<const float4 *>(&n->box[8]);
            const int4 q3 = *reinterpret_cast<const int4 *>(&n->ref[0]);
            const float ba[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};  // child 2n
            const float bb[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};  // child 2n + 1
            const bool right = (start >> k) & 1u;
            float sb[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) sb[j] = right ? ba[j] : bb[j];  // the sibling's box
            float t0, t1;
            box_test(sb, ox, oy, oz, ix, iy, iz, t0, t1);
            r.boxes += 1;
            if (!(t1 < 0.0f) && !(t0 > r.t)) stk.push(make_uint2((uint32_t)(right ? q3.x : q3.y), __float_as_uint(t0)));
            ref = right ? q3.y : q3.x;
#pragma unroll
            for (int j = 0; j < 6; ++j) box[j] = right ? bb[j] : ba[j];
        }
    }
    float t0, t1;
    box_test(box, ox, oy, oz, ix, iy, iz, t0, t1);
    r.boxes += 1;
    if (!(t0 > r.t || t1 < 0.0f)) return ref;
    while (stk.sp > 0) {
        const uint2 e = stk.pop();
        if (!(__uint_as_float(e.y) > r.t)) return (int32_t)e.x;
    }
    return -1;
}

// traverse_shared from heap node `start` (rt_trace with RT_FLAG_START_AT_IROOT):
// the same loop after trav_enter.  With start 1 it visits what traverse()
// visits; from a leaf it visits what the reference's trail loop visits from
// that leaf, in the same order, so hits and counters equal the oracle's.
template <bool ANYHIT, class Stack>
__device__ __forceinline__ TraceOut traverse_from(const TravNode *__restrict__ nodes,
                                                  const TravTri *__restrict__ tris, int32_t root_ref,
                                                  const float *root_box, uint32_t start, float ox, float oy,
                                                  float oz, float dx, float dy, float dz, float tmax, int32_t self,
                                                  Stack &stk)
{
    TraceOut r;
    r.tri = -1;
    r.u = -1.0f;
    r.v = -1.0f;
    r.t = tmax;
    r.boxes = 0;
    r.tris = 0;
    r.leaf_heap = 1;
    r.pos = -1;
    if (root_ref == -1) return r;
    const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
    const uint32_t neg = (ix < 0.0f ? 1u : 0u) | (iy < 0.0f ? 2u : 0u) | (iz < 0.0f ? 4u : 0u);
    int32_t cur = trav_enter(nodes, root_ref, root_box, start, ox, oy, oz, ix, iy, iz, r, stk);
    if (cur == -1) return r;
    while (true) {
        const uint2 ptop = stk.peek();  // read while the fetch is in flight
        float4 q0, q1, q2, q3f;
        fetch_step(nodes, tris, cur, q0, q1, q2, q3f);
        if (cur >= 0) {
            const int4 q3 = as_int4(q3f);
            const float ba[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
            const float bb[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
            float a0, b0;
            bool a_ok, b_ok;
            child_test(ba, ox, oy, oz, ix, iy, iz, r.t, a0, a_ok);
            child_test(bb, ox, oy, oz, ix, iy, iz, r.t, b0, b_ok);
            r.boxes += 2;
            const int first_is_b = (neg >> q3.z) & 1u;
            const float s0 = first_is_b ? a0 : b0;
            const int32_t fref = first_is_b ? q3.y : q3.x;
            const int32_t sref = first_is_b ? q3.x : q3.y;
            const bool f_ok = first_is_b ? b_ok : a_ok;
            const bool s_ok = first_is_b ? a_ok : b_ok;
            if (f_ok && s_ok) stk.push(make_uint2((uint32_t)sref, __float_as_uint(s0)));
            if (f_ok || s_ok) {
                cur = f_ok ? fref : sref;
                continue;
            }
        } else {
            if (leaf_test_first(r, tris, (uint32_t)cur, q0, q1, q2, ox, oy, oz, dx, dy, dz, self, ANYHIT)) return r;
        }
        if (stk.sp == 0) break;
        --stk.sp;
        if (!(__uint_as_float(ptop.y) > r.t)) {
            cur = (int32_t)ptop.x;
            continue;
        }
        bool found = false;
        while (stk.sp > 0) {
            const uint2 e = stk.pop();
            if (!(__uint_as_float(e.y) > r.t)) {
                cur = (int32_t)e.x;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return r;
}

// The same traversal written as a step function: trav_begin tests the root
// box; each trav_step fetches with fetch_step, handles one internal node (both
// child boxes) or one leaf and pops the stack; a ray is finished when either
// returns true.  It visits the same nodes in the same order as traverse(),
// but the compiler lays the loop out differently, and which layout is faster
// depends on the scene (coherent rays over deep leaves favour the loop with
// split fetches; incoherent rays through thin geometry, whose waves mix lanes
// at nodes and at leaves, favour the shared fetch), so the launch shape's
// traversal field picks one per launch: 0 loop, 1 step, 2 loop with the
// shared fetch.
struct Trav {
    float ox, oy, oz, dx, dy, dz, ix, iy, iz;
    int32_t self, cur;
    uint32_t neg;  // bit 0: 1/dx < 0, bit 1: 1/dy < 0, bit 2: 1/dz < 0
    bool anyhit;
    TraceOut r;
};

template <class Stack>
__device__ __forceinline__ bool trav_begin(Trav &T, int32_t root_ref, const float *root_box, float ox,
                                           float oy, float oz, float dx, float dy, float dz, float tmax,
                                           int32_t self, bool anyhit, Stack &stk)
{
    T.r.tri = -1;
    T.r.u = -1.0f;
    T.r.v = -1.0f;
    T.r.t = tmax;
    T.r.boxes = 0;
    T.r.tris = 0;
    T.r.leaf_heap = 1;
    T.r.pos = -1;
    T.ox = ox;
    T.oy = oy;
    T.oz = oz;
    T.dx = dx;
    T.dy = dy;
    T.dz = dz;
    T.self = self;
    T.anyhit = anyhit;
    if (root_ref == -1) return true;
    T.ix = 1.0f / dx;
    T.iy = 1.0f / dy;
    T.iz = 1.0f / dz;
    T.neg = (T.ix < 0.0f ? 1u : 0u) | (T.iy < 0.0f ? 2u : 0u) | (T.iz < 0.0f ? 4u : 0u);
    float t0, t1;
    box_test(root_box, ox, oy, oz, T.ix, T.iy, T.iz, t0, t1);
    T.r.boxes = 1;
    if (t0 > T.r.t || t1 < 0.0f) return true;
    T.cur = root_ref;
    stk.sp = 0;
    return false;
}

// MODE 0: closest hit, 1: any hit, 2: as T.anyhit says.
template <int MODE, class Stack>
__device__ __forceinline__ bool trav_step(Trav &T, const TravNode *__restrict__ nodes,
                                          const TravTri *__restrict__ tris, Stack &stk)
{
    TraceOut &r = T.r;
    const uint2 ptop = stk.peek();  // read while the fetch is in flight
    float4 q0, q1, q2, q3f;
    fetch_step(nodes, tris, T.cur, q0, q1, q2, q3f);
    if (T.cur >= 0) {
        const int4 q3 = as_int4(q3f);
        const float ba[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
        const float bb[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
        float a0, b0;
        bool a_ok, b_ok;
        child_test(ba, T.ox, T.oy, T.oz, T.ix, T.iy, T.iz, r.t, a0, a_ok);
        child_test(bb, T.ox, T.oy, T.oz, T.ix, T.iy, T.iz, r.t, b0, b_ok);
        r.boxes += 2;
        const int first_is_b = (T.neg >> q3.z) & 1u;
        const float s0 = first_is_b ? a0 : b0;
        const int32_t fref = first_is_b ? q3.y : q3.x;
        const int32_t sref = first_is_b ? q3.x : q3.y;
        const bool f_ok = first_is_b ? b_ok : a_ok;
        const bool s_ok = first_is_b ? a_ok : b_ok;
        if (f_ok && s_ok) stk.push(make_uint2((uint32_t)sref, __float_as_uint(s0)));
        if (f_ok || s_ok) {
            T.cur = f_ok ? fref : sref;
            return false;
        }
    } else {
        if (leaf_test_first(r, tris, (uint32_t)T.cur, q0, q1, q2, T.ox, T.oy, T.oz, T.dx, T.dy, T.dz, T.self,
                      MODE == 1 || (MODE == 2 && T.anyhit)))
            return true;
    }
    // pop the next live entry, the top first
    if (stk.sp == 0) return true;
    --stk.sp;
    if (!(__uint_as_float(ptop.y) > r.t)) {
        T.cur = (int32_t)ptop.x;
        return false;
    }
    while (stk.sp > 0) {
        const uint2 e = stk.pop();
        if (!(__uint_as_float(e.y) > r.t)) {
            T.cur = (int32_t)e.x;
            return false;
        }
    }
    return true;
}

template <bool ANYHIT, class Stack>
__device__ __forceinline__ TraceOut traverse_steps(const TravNode *__restrict__ nodes,
                                             const TravTri *__restrict__ tris, int32_t root_ref,
                                             const float *root_box, float ox, float oy, float oz,
                                             float dx, float dy, float dz, float tmax, int32_t self,
                                             Stack &stk)
{
    Trav T;
    if (!trav_begin(T, root_ref, root_box, ox, oy, oz, dx, dy, dz, tmax, self, ANYHIT, stk))
        while (!trav_step<ANYHIT ? 1 : 0>(T, nodes, tris, stk)) {
        }
    return T.r;
}


template <int FORM, bool ANYHIT, class Stack>
__device__ __forceinline__ TraceOut traverse_form(const TravNode *__restrict__ nodes,
                                                  const TravTri *__restrict__ tris, int32_t root_ref,
                                                  const float *root_box, float ox, float oy, float oz,
                                                  float dx, float dy, float dz, float tmax, int32_t self,
                                                  Stack &stk)
{
    if (FORM == 1)
        return traverse_steps<ANYHIT>(nodes, tris, root_ref, root_box, ox, oy, oz, dx, dy, dz, tmax, self, stk);
    if (FORM == 2)
        return traverse_shared<ANYHIT>(nodes, tris, root_ref, root_box, ox, oy, oz, dx, dy, dz, tmax, self, stk);
    return traverse<ANYHIT>(nodes, tris, root_ref, root_box, ox, oy, oz, dx, dy, dz, tmax, self, stk);
}

struct SceneView {
    const TravNode *nodes;
    const TravTri *tris;
    const SceneHdr *hdr;
};

__device__ __forceinline__ void load_root(const SceneHdr *hdr, int32_t &root_ref, float *root_box)
{
    root_ref = hdr->root_ref;
#pragma unroll
    for (int k = 0; k < 6; ++k) root_box[k] = hdr->root_box[k];
}

__device__ __forceinline__ void count_add(unsigned long long *stats, int slot, uint64_t v)
{
    // wave reduction, one atomic per wave
    uint64_t x = v;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if (__lane_id() == 0 && x) atomicAdd(&stats[slot], (unsigned long long)x);
}

__device__ __forceinline__ void store_hit(rt_hit24 *dst, const TraceOut &o)
{
    int2 *hp = reinterpret_cast<int2 *>(dst);
    hp[0] = make_int2(o.tri, __float_as_int(o.u));
    hp[1] = make_int2(__float_as_int(o.v), __float_as_int(o.t));
    hp[2] = make_int2((int32_t)(o.boxes + 65536u * o.tris), o.leaf_heap);
}

// rt_trace: one lane per ray.
template <bool ANYHIT, bool COUNT, int BLOCK, int KLDS, int FORM>
__global__ __launch_bounds__(BLOCK) void k_trace(SceneView sv, const rt_ray48 *__restrict__ rays,
                                                 uint32_t n, rt_hit24 *hits, int from_root_field,
                                                 unsigned long long *stats)
{
    __shared__ uint2 lds[KLDS * BLOCK];
    RayStack<BLOCK, KLDS> stk;
    stk.lds = &lds[threadIdx.x];
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    int32_t root_ref;
    float root_box[6];
    load_root(sv.hdr, root_ref, root_box);
    uint32_t boxes = 0, ntris = 0;
    if (i < n) {
        const float4 *rp = reinterpret_cast<const float4 *>(&rays[i]);
        const float4 r0 = rp[0], r1 = rp[1], r2 = rp[2];
        // o = r0.xyz, d = (r0.w, r1.x, r1.y), refl = (r1.z, r1.w, r2.x), tmin = r2.y,
        // tmax = r2.z, tri = r2.w
        const int32_t self = __float_as_int(r2.w);
        if (self <= -2) {  // dead ray (CSIntersections.hlsl:23,37)
            TraceOut d;
            d.tri = -2;
            d.u = -1.0f;
            d.v = -1.0f;
            d.t = r2.z;
            d.boxes = 0;
            d.tris = 0;
            d.leaf_heap = 1;
            store_hit(&hits[i], d);
        } else {
            TraceOut o;
            if (from_root_field) {  // the previous intersection's iRoot, read before it is overwritten
                const uint32_t start = (uint32_t)hits[i].root;
                o = traverse_from<ANYHIT>(sv.nodes, sv.tris, root_ref, root_box, start, r0.x, r0.y, r0.z, r0.w,
                                          r1.x, r1.y, r2.z, self, stk);
            } else {
                o = traverse_form<FORM, ANYHIT>(sv.nodes, sv.tris, root_ref, root_box, r0.x, r0.y, r0.z, r0.w,
                                                r1.x, r1.y, r2.z, self, stk);
            }
            store_hit(&hits[i], o);
            boxes = o.boxes;
            ntris = o.tris;
        }
    }
    if (COUNT) {
        count_add(stats, 0, boxes);
        count_add(stats, 1, ntris);
    }
}

// Row r of the reference's camera Matrix4 applied to (v, 1), summed left to
// right: HLSL mul(float4(v, 1), g_mfWorld) of the column-major cbuffer.
__device__ __forceinline__ float world_row(const float *m, int r, float v0, float v1, float v2)
{
    return ((m[4 * r + 0] * v0 + m[4 * r + 1] * v1) + m[4 * r + 2] * v2) + m[4 * r + 3];
}

// The primary ray's direction (its origin is c.pos).  jx = jy = 0 gives the
// pixel centre; render jobs pass a jitter in [-1, 1) (CSPrimaryRays.hlsl:25-36).
//  - RT_CAMERA_BASIS: through (px + 0.5 + jx/2, py + 0.5 + jy/2) of the image
//    plane at distance 1 along fwd, half extents tan_x, tan_y;
//  - RT_CAMERA_WORLD: the reference's matrix rule (CSPrimaryRays.hlsl:39-52):
//    pixel pitch 2/H, z = 2, scale = 0.008 / |0.035 z|, d = normalize(
//    (M (scale x, scale y, scale z, 1)).xyz / w - origin); the host has put
//    the origin (M (0, 0, 0, 1)).xyz / w into c.pos.
__device__ __forceinline__ void primary_ray(const rt_camera &c, uint32_t px, uint32_t py, float fw,
                                            float fh, float jx, float jy, float &dx, float &dy,
                                            float &dz)
{
    float ddx, ddy, ddz;
    if (c.kind == RT_CAMERA_WORLD) {
        const float inverse = 1.0f / fh;
        const float y = -((((2.0f * (float)py + jy) + 1.0f) - fh) * inverse);
        const float x = (((2.0f * (float)px + jx) + 1.0f) - fw) * inverse;
        const float z = 2.0f;
        const float scale = 0.008f / fabsf(0.035f * z);
        const float p0 = scale * x, p1 = scale * y, p2 = scale * z;
        const float w = world_row(c.world, 3, p0, p1, p2);
        ddx = world_row(c.world, 0, p0, p1, p2) / w - c.pos[0];
        ddy = world_row(c.world, 1, p0, p1, p2) / w - c.pos[1];
        ddz = world_row(c.world, 2, p0, p1, p2) / w - c.pos[2];
    } else {
        const float x = (((float)(2u * px + 1u) + jx) - fw) / fw;
        const float y = -((((float)(2u * py + 1u) + jy) - fh) / fh);
        const float sx = x * c.tan_x, sy = y * c.tan_y;
        const float ax = c.fwd[0] + sx * c.right[0];
        const float ay = c.fwd[1] + sx * c.right[1];
        const float az = c.fwd[2] + sx * c.right[2];
        ddx = ax + sy * c.up[0];
        ddy = ay + sy * c.up[1];
        ddz = az + sy * c.up[2];
    }
    const float len = sqrtf((ddx * ddx + ddy * ddy) + ddz * ddz);
    dx = ddx / len;
    dy = ddy / len;
    dz = ddz / len;
}

// Flat unit normal of a sorted triangle record: normalize(cross(E1, E2))
// (CSComputeColorRT.hlsl:61-77).
__device__ __forceinline__ void flat_normal(const TravTri &T, float &nx, float &ny, float &nz)
{
    const float4 e1 = T.e1, e2 = T.e2;
    nx = (e1.y * e2.z) - (e1.z * e2.y);
    ny = (e1.z * e2.x) - (e1.x * e2.z);
    nz = (e1.x * e2.y) - (e1.y * e2.x);
    const float nl = sqrtf((nx * nx + ny * ny) + nz * nz);
    nx = nx / nl;
    ny = ny / nl;
    nz = nz / nl;
}

// ---------------------------------------------------------------------------
// Quads: four lanes per ray for the heaviest tiles of a frame.  The subtrees
// rooted at heap depth D of the LBVH ("items"; a leaf above depth D is an
// item too) are numbered in the ray's front-to-back visit order (the side
// dirIsNeg[axis] first at each node), and lane s of the quad takes the items
// numbered s, s + 4, s + 8, ...  Above depth D the four lanes walk the same
// nodes: there a box is culled only against the ray's tmax and an entry
// stacked there is never culled on pop, so the numbering is the same in all
// four lanes whatever each lane's best t.  Inside its own items a lane culls
// as the whole traversal does (a box it culls cannot hold a hit closer than
// its own best, hence none closer than the quad's).  The quad's closest hit
// is the smallest t over its lanes, on equal t the one of the lowest item
// number: the hit the whole traversal keeps, since it takes a hit only with
// t < best t.  A shadow ray is blocked when any lane finds a blocker.  A
// pixel's chain of dependent iterations becomes the longest of its four
// shares (the items along a ray's path alternate between the lanes); box
// and triangle counts differ from a single traversal's, so quads render
// frames and never count.
// ---------------------------------------------------------------------------
constexpr uint32_t kQuadDepth = 8;  // heap depth of the subtrees a quad shares out

template <bool ANYHIT, class Stack>
__device__ __forceinline__ TraceOut traverse_quad(const TravNode *__restrict__ nodes,
                                                  const TravTri *__restrict__ tris, int32_t root_ref,
                                                  const float *root_box, uint32_t s, uint32_t D, float ox,
                                                  float oy, float oz, float dx, float dy, float dz, float tmax,
                                                  int32_t self, Stack &stk, uint32_t &rank, uint32_t kmask = 3u)
{
    TraceOut r;
    r.tri = -1;
    r.u = -1.0f;
    r.v = -1.0f;
    r.t = tmax;
    r.boxes = 0;
    r.tris = 0;
    r.leaf_heap = 1;
    r.pos = -1;
    rank = 0;
    if (root_ref == -1) return r;
    const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
    const uint32_t neg = (ix < 0.0f ? 1u : 0u) | (iy < 0.0f ? 2u : 0u) | (iz < 0.0f ? 4u : 0u);
    float t0, t1;
    box_test(root_box, ox, oy, oz, ix, iy, iz, t0, t1);
    r.boxes = s == 0 ? 1u : 0u;  // the walk above depth D is counted once per quad
    if (t0 > r.t || t1 < 0.0f) return r;
    const uint32_t kNeverCulled = 0xff800000u;  // -inf: no best t is below it
    int32_t cur = root_ref;
    uint32_t items = 0, item = 0;
    stk.sp = 0;
    while (true) {
        const uint2 ptop = stk.peek();  // read while the fetch is in flight
        float4 q0, q1, q2, q3f;
        fetch_step(nodes, tris, cur, q0, q1, q2, q3f);
        const bool at_node = cur >= 0;
        const uint32_t depth = 31u - __clz((uint32_t)__float_as_int(at_node ? q3f.w : q2.w));
        const bool top = depth < D;
        bool mine = true;
        if (depth == D || (top && !at_node)) {
            item = items++;
            mine = (item & kmask) == s;
        }
        if (mine) {
            if (at_node) {
                const int4 q3 = as_int4(q3f);
                const float ba[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
                const float bb[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
                const float lim = top ? tmax : r.t;
                float a0, b0;
                bool a_ok, b_ok;
                child_test(ba, ox, oy, oz, ix, iy, iz, lim, a0, a_ok);
                child_test(bb, ox, oy, oz, ix, iy, iz, lim, b0, b_ok);
                r.boxes += (!top || s == 0) ? 2u : 0u;
                const int first_is_b = (neg >> q3.z) & 1u;
                const float s0 = first_is_b ? a0 : b0;
                const int32_t fref = first_is_b ? q3.y : q3.x;
                const int32_t sref = first_is_b ? q3.x : q3.y;
                const bool f_ok = first_is_b ? b_ok : a_ok;
                const bool s_ok = first_is_b ? a_ok : b_ok;
                if (f_ok && s_ok) stk.push(make_uint2((uint32_t)sref, top ? kNeverCulled : __float_as_uint(s0)));
                if (f_ok || s_ok) {
                    cur = f_ok ? fref : sref;
                    continue;
                }
            } else {
                const float before = r.t;
                const bool stop = leaf_test_first(r, tris, (uint32_t)cur, q0, q1, q2, ox, oy, oz, dx, dy, dz,
                                                  self, ANYHIT);
                if (r.t < before) rank = item;
                if (stop) return r;
            }
        }
        if (stk.sp == 0) break;
        --stk.sp;
        if (!(__uint_as_float(ptop.y) > r.t)) {
            cur = (int32_t)ptop.x;
            continue;
        }
        bool found = false;
        while (stk.sp > 0) {
            const uint2 e = stk.pop();
            if (!(__uint_as_float(e.y) > r.t)) {
                cur = (int32_t)e.x;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return r;
}

// The quad's closest hit in every lane of the quad: smallest t, on equal t
// the lowest rank.  Lanes without a hit hold t = tmax, above every hit.
__device__ __forceinline__ void quad_closest(TraceOut &h, uint32_t rank)
{
#pragma unroll
    for (int o = 1; o <= 2; o <<= 1) {
        const float t = __shfl_xor(h.t, o);
        const int32_t tri = __shfl_xor(h.tri, o);
        const int32_t pos = __shfl_xor(h.pos, o);
        const float u = __shfl_xor(h.u, o);
        const float v = __shfl_xor(h.v, o);
        const int32_t heap = __shfl_xor(h.leaf_heap, o);
        const uint32_t rk = __shfl_xor(rank, o);
        if (t < h.t || (t == h.t && rk < rank)) {
            h.t = t;
            h.tri = tri;
            h.pos = pos;
            h.u = u;
            h.v = v;
            h.leaf_heap = heap;
            rank = rk;
        }
    }
}

__device__ __forceinline__ bool quad_any(bool x)
{
    int b = x ? 1 : 0;
    b |= __shfl_xor(b, 1);
    b |= __shfl_xor(b, 2);
    return b != 0;
}

__device__ __forceinline__ uint32_t quad_sum(uint32_t x)
{
    x += __shfl_xor(x, 1);
    x += __shfl_xor(x, 2);
    return x;
}

// quad_closest and quad_any over an aligned group of k lanes (k a power of
// two, at most 64).
__device__ __forceinline__ void group_closest(TraceOut &h, uint32_t rank, uint32_t k)
{
    for (uint32_t o = 1; o < k; o <<= 1) {
        const float t = __shfl_xor(h.t, o);
        const int32_t tri = __shfl_xor(h.tri, o);
        const int32_t pos = __shfl_xor(h.pos, o);
        const float u = __shfl_xor(h.u, o);
        const float v = __shfl_xor(h.v, o);
        const int32_t heap = __shfl_xor(h.leaf_heap, o);
        const uint32_t rk = __shfl_xor(rank, o);
        if (t < h.t || (t == h.t && rk < rank)) {
            h.t = t;
            h.tri = tri;
            h.pos = pos;
            h.u = u;
            h.v = v;
            h.leaf_heap = heap;
            rank = rk;
        }
    }
}

__device__ __forceinline__ bool group_any(bool x, uint32_t k)
{
    int b = x ? 1 : 0;
    for (uint32_t o = 1; o < k; o <<= 1) b |= __shfl_xor(b, o);
    return b != 0;
}

// Direct term of one primary hit (CSComputeColorRT.hlsl:125-198, white
// material of alpha 1): shadow ray {o = hit, d = -Ldir, maxT = |hit - L|,
// self = hit triangle} per light, cast when dot(n, Ldir) * dot(n, d) > 0,
// traced any-hit from the root (the occlusion predicate of the reference's
// closest-hit shadow ray).  Returns (ambient + sum max(0, -n'.Ldir)) * albedo.
// With QUAD the four lanes of a quad hold the same hit and each traces its
// share of every shadow ray (quad_lane: the lane's place in the quad;
// quad_depth: the heap depth of the items shared out).
template <int FORM, bool QUAD = false, class Stack>
__device__ __forceinline__ float direct_term(const SceneView &sv, int32_t root_ref,
                                             const float *root_box, const TraceOut &h, float ox,
                                             float oy, float oz, float dx, float dy, float dz,
                                             const float *lights, uint32_t nlights, float ambient,
                                             float albedo, Stack &stk, uint32_t &boxes,
                                             uint32_t &ntris, uint32_t &nshadow, uint8_t *occ_out,
                                             uint32_t quad_lane = 0, uint32_t quad_depth = 0)
{
    float nx, ny, nz;
    flat_normal(sv.tris[h.pos], nx, ny, nz);
    float mx = nx, my = ny, mz = nz;  // normal facing the ray (:125-131)
    if (((mx * dx + my * dy) + mz * dz) > 0.0f) {
        mx = -mx;
        my = -my;
        mz = -mz;
    }
    const float hx = ox + h.t * dx, hy = oy + h.t * dy, hz = oz + h.t * dz;
    float diffuse = 0.0f;
    for (uint32_t l = 0; l < nlights; ++l) {
        float lx = hx - lights[3 * l + 0];
        float ly = hy - lights[3 * l + 1];
        float lz = hz - lights[3 * l + 2];
        const float dist = sqrtf((lx * lx + ly * ly) + lz * lz);
        lx = lx / dist;
        ly = ly / dist;
        lz = lz / dist;
        const float ls = (nx * lx + ny * ly) + nz * lz;
        const float rs = (nx * dx + ny * dy) + nz * dz;
        uint8_t o = 2;
        if ((ls * rs) > 0.0f) {
            TraceOut s;
            if constexpr (QUAD) {
                uint32_t rank;
                s = traverse_quad<true>(sv.nodes, sv.tris, root_ref, root_box, quad_lane, quad_depth, hx, hy,
                                        hz, -lx, -ly, -lz, dist, h.tri, stk, rank);
                s.tri = quad_any(s.tri >= 0) ? 0 : -1;
            } else {
                s = traverse_form<FORM, true>(sv.nodes, sv.tris, root_ref, root_box, hx, hy, hz, -lx, -ly, -lz,
                                              dist, h.tri, stk);
            }
            boxes += s.boxes;
            ntris += s.tris;
            nshadow += 1;
            o = s.tri < 0 ? 0 : 1;
            if (s.tri < 0) {
                const float dd = -((mx * lx + my * ly) + mz * lz);
                diffuse = diffuse + (dd > 0.0f ? dd : 0.0f);
            }
        }
        if (occ_out) occ_out[l] = o;
    }
    return (ambient + diffuse) * albedo;
}

// An any-hit ray traced by a group of k lanes (k a power of two, kmask = k -
// 1, lane s of the group): every lane walks the levels above heap depth D
// (culled against tmax only, as an any-hit traversal culls until its first
// hit, so the lanes walk them alike and number the subtrees at depth D and
// the leaves above it alike); lane s searches the items numbered s modulo k.
// A lane that accepts a triangle stops, and the others of its group (gmask)
// stop at their next step.  The lanes of the wave run the loop together
// (its ballots need them); lanes of no group pass member = false.  Returns
// whether this lane found the occluder.  The group's answer, "some triangle
// of the ray's reachable leaves passes the test", is the one-lane traversal's,
// whichever triangle each finds first.
template <class Stack>
__device__ __forceinline__ bool traverse_group_anyhit(const TravNode *__restrict__ nodes,
                                                      const TravTri *__restrict__ tris, int32_t root_ref,
                                                      const float *root_box, uint32_t s, uint32_t kmask, uint32_t D,
                                                      bool member, uint64_t gmask, float ox, float oy, float oz,
                                                      float dx, float dy, float dz, float tmax, int32_t self,
                                                      Stack &stk, uint32_t &boxes, uint32_t &ntris)
{
    bool found = false, active = member && root_ref != -1;
    float ix = 0.0f, iy = 0.0f, iz = 0.0f;
    uint32_t neg = 0;
    if (active) {
        ix = 1.0f / dx;
        iy = 1.0f / dy;
        iz = 1.0f / dz;
        neg = (ix < 0.0f ? 1u : 0u) | (iy < 0.0f ? 2u : 0u) | (iz < 0.0f ? 4u : 0u);
        float t0, t1;
        box_test(root_box, ox, oy, oz, ix, iy, iz, t0, t1);
        boxes += s == 0 ? 1u : 0u;
        if (t0 > tmax || t1 < 0.0f) active = false;
    }
    TraceOut r;
    r.tri = -1;
    r.u = -1.0f;
    r.v = -1.0f;
    r.t = tmax;
    r.boxes = 0;
    r.tris = 0;
    r.leaf_heap = 1;
    r.pos = -1;
    int32_t cur = root_ref;
    uint32_t items = 0;
    stk.sp = 0;
    while (true) {
        if (__ballot(found) & gmask) active = false;
        if (__ballot(active) == 0) break;
        if (active) {
            const uint2 ptop = stk.peek();  // read while the fetch is in flight
            float4 q0, q1, q2, q3f;
            fetch_step(nodes, tris, cur, q0, q1, q2, q3f);
            const bool at_node = cur >= 0;
            const uint32_t depth = 31u - __clz((uint32_t)__float_as_int(at_node ? q3f.w : q2.w));
            const bool top = depth < D;
            bool mine = true;
            if (depth == D || (top && !at_node)) {
                mine = (items & kmask) == s;
                ++items;
            }
            bool go = false;
            if (mine) {
                if (at_node) {
                    const int4 q3 = as_int4(q3f);
                    const float ba[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
                    const float bb[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
                    float a0, b0;
                    bool a_ok, b_ok;
                    child_test(ba, ox, oy, oz, ix, iy, iz, tmax, a0, a_ok);
                    child_test(bb, ox, oy, oz, ix, iy, iz, tmax, b0, b_ok);
                    boxes += (!top || s == 0) ? 2u : 0u;
                    const int first_is_b = (neg >> q3.z) & 1u;
                    const float s0 = first_is_b ? a0 : b0;
                    const int32_t fref = first_is_b ? q3.y : q3.x;
                    const int32_t sref = first_is_b ? q3.x : q3.y;
                    const bool f_ok = first_is_b ? b_ok : a_ok;
                    const bool s_ok = first_is_b ? a_ok : b_ok;
                    if (f_ok && s_ok) stk.push(make_uint2((uint32_t)sref, __float_as_uint(s0)));
                    if (f_ok || s_ok) {
                        cur = f_ok ? fref : sref;
                        go = true;
                    }
                } else if (leaf_test_first(r, tris, (uint32_t)cur, q0, q1, q2, ox, oy, oz, dx, dy, dz, self, true)) {
                    found = true;
                    active = false;
                }
            }
            // a pop (an any-hit traversal culls nothing it pushed: every
            // entry's T0 is at most tmax, the best t until the first hit)
            if (active && !go) {
                if (stk.sp == 0) {
                    active = false;
                } else {
                    --stk.sp;
                    cur = (int32_t)ptop.x;
                }
            }
        }
    }
    ntris += r.tris;
    return found;
}

// direct_term for the lanes of a wave together (all 64 lanes call it; lit:
// the lane holds a primary hit to shade).  Per light, the lanes whose hit
// casts a shadow ray (m of them) each lend their ray to a group of k =
// 64 / m lanes (a power of two, at most 16; k = 1 above 32 rays), which
// traces it with traverse_group_anyhit: the wave's idle lanes share the
// shadow rays' subtrees instead of waiting for the longest ray.  Only
// occlusion reaches the image, so the result equals direct_term's; the box
// and triangle counts do not (counted frames do not come here).
template <int FORM, bool KEEP, class Stack>
__device__ __forceinline__ float shade_wave(const SceneView &sv, int32_t root_ref, const float *root_box,
                                            const TraceOut &h, bool lit, float ox, float oy, float oz, float dx,
                                            float dy, float dz, const float *lights, uint32_t nlights, float ambient,
                                            float albedo, Stack &stk, uint32_t &boxes, uint32_t &ntris,
                                            uint32_t &nshadow, uint8_t *occ_out)
{
    const uint32_t lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    float nx = 0.0f, ny = 0.0f, nz = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
    float hx = 0.0f, hy = 0.0f, hz = 0.0f, rs = 0.0f;
    if (lit) {
        flat_normal(sv.tris[h.pos], nx, ny, nz);
        mx = nx;
        my = ny;
        mz = nz;
        if (((mx * dx + my * dy) + mz * dz) > 0.0f) {
            mx = -mx;
            my = -my;
            mz = -mz;
        }
        hx = ox + h.t * dx;
        hy = oy + h.t * dy;
        hz = oz + h.t * dz;
        rs = (nx * dx + ny * dy) + nz * dz;
    }
    float diffuse = 0.0f;
    for (uint32_t l = 0; l < nlights; ++l) {
        float lx = 0.0f, ly = 0.0f, lz = 0.0f, dist = 0.0f;
        bool need = false;
        if (lit) {
            lx = hx - lights[3 * l + 0];
            ly = hy - lights[3 * l + 1];
            lz = hz - lights[3 * l + 2];
            dist = sqrtf((lx * lx + ly * ly) + lz * lz);
            lx = lx / dist;
            ly = ly / dist;
            lz = lz / dist;
            const float ls = (nx * lx + ny * ly) + nz * lz;
            need = (ls * rs) > 0.0f;
            if (KEEP && !need) occ_out[l] = 2;
        }
        const uint64_t needm = __ballot(need);
        if (needm == 0) continue;
        const uint32_t m = (uint32_t)__popcll(needm);
        bool occluded = false;
        if (m > 32) {
            if (need) {
                const TraceOut sh = traverse_form<FORM, true>(sv.nodes, sv.tris, root_ref, root_box, hx, hy, hz, -lx,
                                                              -ly, -lz, dist, h.tri, stk);
                boxes += sh.boxes;
                ntris += sh.tris;
                occluded = sh.tri >= 0;
            }
        } else {
            const uint32_t lk = m > 16 ? 1u : m > 8 ? 2u : m > 4 ? 3u : 4u;  // log2 k
            const uint32_t k = 1u << lk, g = lane >> lk;
            // the g-th lane casting a ray lends it to group g
            uint64_t rest = needm;
            uint32_t src = 0;
            for (uint32_t j = 0; j < m; ++j) {
                const uint32_t b = (uint32_t)__builtin_ctzll(rest);
                rest &= rest - 1ull;
                if (j == g) src = b;
            }
            const float gx = __shfl(hx, src), gy = __shfl(hy, src), gz = __shfl(hz, src);
            const float gdx = __shfl(-lx, src), gdy = __shfl(-ly, src), gdz = __shfl(-lz, src);
            const float gdist = __shfl(dist, src);
            const int32_t gself = __shfl(h.tri, src);
            const bool member = g < m;
            const uint64_t kbits = (1ull << k) - 1ull;
            const uint64_t gmask = member ? kbits << (g * k) : 0ull;
            const bool found = traverse_group_anyhit(sv.nodes, sv.tris, root_ref, root_box, lane & (k - 1u), k - 1u,
                                                     kQuadDepth, member, gmask, gx, gy, gz, gdx, gdy, gdz, gdist,
                                                     gself, stk, boxes, ntris);
            const uint64_t F = __ballot(found);
            const uint32_t rank = (uint32_t)__popcll(needm & below);
            occluded = need && ((F >> (rank * k)) & kbits) != 0ull;
        }
        if (need) {
            nshadow += 1;
            if (!occluded) {
                const float dd = -((mx * lx + my * ly) + mz * lz);
                diffuse = diffuse + (dd > 0.0f ? dd : 0.0f);
            }
            if (KEEP) occ_out[l] = occluded ? 1 : 0;
        }
    }
    return (ambient + diffuse) * albedo;
}

struct RenderParams {
    rt_camera cam;
    uint32_t width, height, nlights;
    float ambient, albedo, env;
    float lights[24];           // up to 8 point lights
    const uint32_t *order;      // tile of each block, heaviest first; null: block index order
    uint32_t *cost;             // per tile: max over its lanes of box + triangle tests (atomicMax)
    uint32_t nsplit;            // k_render_direct_quad: the first nsplit tiles of the order get quads
    uint32_t quad_depth;        // the heap depth of the subtrees a quad shares out
    uint32_t group;             // k_render_direct, form 2, no counts: shade_wave (shadow rays on lane groups)
};

// A block renders a tile of 16 x BLOCK/16 pixels.  With two or more waves
// per block each wave takes an 8 x 8 square of it (a smaller footprint than a
// 16 x 4 strip, so its rays stay together deeper into the tree).
template <int BLOCK>
struct Tile {
    static constexpr int W = 16, H = BLOCK / 16;
    static __device__ __forceinline__ void offset(uint32_t t, uint32_t &x, uint32_t &y)
    {
        if (BLOCK >= 128) {
            const uint32_t w = t >> 6, l = t & 63u;
            x = (w & 1u) * 8u + (l & 7u);
            y = (w >> 1) * 8u + (l >> 3);
        } else {
            x = t % W;
            y = t / W;
        }
    }
};

// A tile's cost for the next frame's order: the box + triangle tests of the
// longest lane of its waves (a wave runs as long as its longest lane).
__device__ __forceinline__ void tile_cost_max(uint32_t *cost, uint32_t tile, uint32_t c)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t y = __shfl_xor(c, o);
        c = y > c ? y : c;
    }
    if (__lane_id() == 0 && c) atomicMax(&cost[tile], c);
}

// k_render_direct's work for one tile, without kept hits or counts: the
// tiles of k_render_direct_quad that have one lane per pixel.  (The kernel
// keeps its own copy, so that its code stays as it was measured.)
template <int BLOCK, int KLDS, int FORM>
__device__ __forceinline__ void render_tile(const SceneView &sv, const RenderParams &rp, uint32_t tile,
                                            float4 *__restrict__ rgba, unsigned long long *stats,
                                            RayStack<BLOCK, KLDS> &stk)
{
    constexpr int TW = Tile<BLOCK>::W, TH = Tile<BLOCK>::H;
    const uint32_t tiles_x = (rp.width + TW - 1) / TW;
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    uint32_t ox_, oy_;
    Tile<BLOCK>::offset(threadIdx.x, ox_, oy_);
    const uint32_t px = tx * TW + ox_;
    const uint32_t py = ty * TH + oy_;
    const bool valid = px < rp.width && py < rp.height;
    int32_t root_ref;
    float root_box[6];
    load_root(sv.hdr, root_ref, root_box);
    uint32_t boxes = 0, ntris = 0, nshadow = 0;
    if (valid) {
        const uint32_t pix = py * rp.width + px;
        float dx, dy, dz;
        primary_ray(rp.cam, px, py, (float)rp.width, (float)rp.height, 0.0f, 0.0f, dx, dy, dz);
        const float ox = rp.cam.pos[0], oy = rp.cam.pos[1], oz = rp.cam.pos[2];
        const TraceOut h = traverse_form<FORM, false>(sv.nodes, sv.tris, root_ref, root_box, ox, oy, oz,
                                                      dx, dy, dz, 10000000000000000.0f, -1, stk);
        boxes += h.boxes;
        ntris += h.tris;
        float color = rp.env;
        if (h.tri >= 0)
            color = direct_term<FORM>(sv, root_ref, root_box, h, ox, oy, oz, dx, dy, dz, rp.lights, rp.nlights,
                                      rp.ambient, rp.albedo, stk, boxes, ntris, nshadow, nullptr);
        rgba[pix] = make_float4(color, color, color, 1.0f);
    }
    count_add(stats, 2, nshadow);
    if (rp.cost) tile_cost_max(rp.cost, tile, boxes + ntris);
}

// Primary + shadow rays + direct term for one image, one tile per block.
template <bool KEEP, bool COUNT, int BLOCK, int KLDS, int FORM>
__global__ __launch_bounds__(BLOCK) void k_render_direct(SceneView sv, RenderParams rp,
                                                         float4 *__restrict__ rgba,
                                                         rt_hit24 *__restrict__ hits,
                                                         uint8_t *__restrict__ occ,
                                                         unsigned long long *stats)
{
    __shared__ uint2 lds[KLDS * BLOCK];
    RayStack<BLOCK, KLDS> stk;
    stk.lds = &lds[threadIdx.x];
    constexpr int TW = Tile<BLOCK>::W, TH = Tile<BLOCK>::H;
    const uint32_t tiles_x = (rp.width + TW - 1) / TW;
    const uint32_t tile = rp.order ? rp.order[blockIdx.x] : blockIdx.x;
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    uint32_t ox_, oy_;
    Tile<BLOCK>::offset(threadIdx.x, ox_, oy_);
    const uint32_t px = tx * TW + ox_;
    const uint32_t py = ty * TH + oy_;
    const bool valid = px < rp.width && py < rp.height;
    int32_t root_ref;
    float root_box[6];
    load_root(sv.hdr, root_ref, root_box);
    uint32_t boxes = 0, ntris = 0, nshadow = 0;
    if (!COUNT && FORM == 2 && rp.group) {
        // the shadow rays on lane groups: shade_wave needs every lane
        const uint32_t pix = valid ? py * rp.width + px : 0u;
        float dx = 0.0f, dy = 0.0f, dz = 0.0f;
        const float ox = rp.cam.pos[0], oy = rp.cam.pos[1], oz = rp.cam.pos[2];
        TraceOut h;
        h.tri = -1;
        if (valid) {
            primary_ray(rp.cam, px, py, (float)rp.width, (float)rp.height, 0.0f, 0.0f, dx, dy, dz);
            h = traverse_form<FORM, false>(sv.nodes, sv.tris, root_ref, root_box, ox, oy, oz, dx, dy, dz,
                                           10000000000000000.0f, -1, stk);
            boxes += h.boxes;
            ntris += h.tris;
            if (KEEP) store_hit(&hits[pix], h);
            if (KEEP && h.tri < 0)
                for (uint32_t l = 0; l < rp.nlights; ++l) occ[(size_t)pix * rp.nlights + l] = 2;
        }
        const bool lit = valid && h.tri >= 0;
        const float shaded = shade_wave<FORM, KEEP>(sv, root_ref, root_box, h, lit, ox, oy, oz, dx, dy, dz, rp.lights,
                                                    rp.nlights, rp.ambient, rp.albedo, stk, boxes, ntris, nshadow,
                                                    KEEP && lit ? &occ[(size_t)pix * rp.nlights] : nullptr);
        if (valid) {
            const float color = lit ? shaded : rp.env;
            rgba[pix] = make_float4(color, color, color, 1.0f);
        }
    } else if (valid) {
        const uint32_t pix = py * rp.width + px;
        float dx, dy, dz;
        primary_ray(rp.cam, px, py, (float)rp.width, (float)rp.height, 0.0f, 0.0f, dx, dy, dz);
        const float ox = rp.cam.pos[0], oy = rp.cam.pos[1], oz = rp.cam.pos[2];
        const TraceOut h = traverse_form<FORM, false>(sv.nodes, sv.tris, root_ref, root_box, ox, oy, oz,
                                                      dx, dy, dz, 10000000000000000.0f, -1, stk);
        boxes += h.boxes;
        ntris += h.tris;
        if (KEEP) store_hit(&hits[pix], h);
        float color;
        if (h.tri < 0) {
            color = rp.env;
            if (KEEP)
                for (uint32_t l = 0; l < rp.nlights; ++l) occ[(size_t)pix * rp.nlights + l] = 2;
        } else {
            color = direct_term<FORM>(sv, root_ref, root_box, h, ox, oy, oz, dx, dy, dz, rp.lights,
                                rp.nlights, rp.ambient, rp.albedo, stk, boxes, ntris, nshadow,
                                KEEP ? &occ[(size_t)pix * rp.nlights] : nullptr);
        }
        rgba[pix] = make_float4(color, color, color, 1.0f);
    }
    if (COUNT) {
        count_add(stats, 0, boxes);
        count_add(stats, 1, ntris);
    }
    count_add(stats, 2, nshadow);
    if (rp.cost) tile_cost_max(rp.cost, tile, boxes + ntris);
}

// Part `part` (0..3) of a tile with a quad per pixel: the tile's pixels
// part * BLOCK/4 .. (part + 1) * BLOCK/4 - 1 in its lane order, pixel
// threadIdx.x / 4 of the part on the lanes 4k .. 4k + 3.  The image equals
// render_tile's.  The tile's cost is the largest sum over a quad's lanes,
// with the walk above the items counted once: about the single lane's
// count (at least it: each lane culls less), so the tile keeps its place in
// the order as the scene or view moves.
template <int BLOCK, int KLDS, int FORM>
__device__ __forceinline__ void render_quad_part(const SceneView &sv, const RenderParams &rp, uint32_t tile,
                                                 uint32_t part, float4 *__restrict__ rgba,
                                                 unsigned long long *stats, RayStack<BLOCK, KLDS> &stk)
{
    constexpr int TW = Tile<BLOCK>::W, TH = Tile<BLOCK>::H;
    const uint32_t tiles_x = (rp.width + TW - 1) / TW;
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    const uint32_t s = threadIdx.x & 3u;
    uint32_t ox_, oy_;
    Tile<BLOCK>::offset(part * (BLOCK / 4) + (threadIdx.x >> 2), ox_, oy_);
    const uint32_t px = tx * TW + ox_;
    const uint32_t py = ty * TH + oy_;
    const bool valid = px < rp.width && py < rp.height;  // the same for the four lanes of a quad
    int32_t root_ref;
    float root_box[6];
    load_root(sv.hdr, root_ref, root_box);
    uint32_t boxes = 0, ntris = 0, nshadow = 0;
    if (valid) {
        const uint32_t pix = py * rp.width + px;
        float dx, dy, dz;
        primary_ray(rp.cam, px, py, (float)rp.width, (float)rp.height, 0.0f, 0.0f, dx, dy, dz);
        const float ox = rp.cam.pos[0], oy = rp.cam.pos[1], oz = rp.cam.pos[2];
        uint32_t rank;
        TraceOut h = traverse_quad<false>(sv.nodes, sv.tris, root_ref, root_box, s, rp.quad_depth, ox, oy, oz,
                                          dx, dy, dz, 10000000000000000.0f, -1, stk, rank);
        boxes += h.boxes;
        ntris += h.tris;
        quad_closest(h, rank);
        float color;
        if (h.tri < 0) {
            color = rp.env;
        } else {
            color = direct_term<FORM, true>(sv, root_ref, root_box, h, ox, oy, oz, dx, dy, dz, rp.lights,
                                            rp.nlights, rp.ambient, rp.albedo, stk, boxes, ntris, nshadow,
                                            nullptr, s, rp.quad_depth);
        }
        if (s == 0) rgba[pix] = make_float4(color, color, color, 1.0f);
    }
    count_add(stats, 2, s == 0 ? nshadow : 0u);
    if (rp.cost) tile_cost_max(rp.cost, tile, quad_sum(boxes + ntris));
}

// A frame whose rp.nsplit heaviest tiles (the head of rp.order) take four
// blocks each, a quad per pixel, ahead of the other tiles, one block each.
template <int BLOCK, int KLDS, int FORM>
__global__ __launch_bounds__(BLOCK) void k_render_direct_quad(SceneView sv, RenderParams rp,
                                                              float4 *__restrict__ rgba,
                                                              unsigned long long *stats)
{
    __shared__ uint2 lds[KLDS * BLOCK];
    RayStack<BLOCK, KLDS> stk;
    stk.lds = &lds[threadIdx.x];
    const uint32_t b = blockIdx.x, nq = 4u * rp.nsplit;
    if (b < nq)
        render_quad_part<BLOCK, KLDS, FORM>(sv, rp, rp.order[b >> 2], b & 3u, rgba, stats, stk);
    else
        render_tile<BLOCK, KLDS, FORM>(sv, rp, rp.order[b - nq + rp.nsplit], rgba, stats, stk);
}

// The next frame's tile order from this frame's tile costs, heaviest first:
// a wave's time is the length of its longest lane's chain of dependent
// iterations, and the frame ends with its last wave, so the longest tiles
// should start first.  Counting sorts over 16 buckets per octave of cost
// (descending; the order inside a bucket is whatever the atomics give), then
// the costs are cleared for the next frame.
constexpr int kOrderBuckets = 33 * 16;

__device__ __forceinline__ int order_bucket(uint32_t c)
{
    if (c == 0) return kOrderBuckets - 1;
    const int e = 31 - __clz(c);                            // 0..31
    const uint32_t m = e >= 4 ? (c >> (e - 4)) & 15u : (c << (4 - e)) & 15u;
    return kOrderBuckets - 1 - (e + 1) * 16 - (int)m;       // larger cost, smaller bucket
}

// Lanes of the wave whose bucket equals this lane's (ten ballots, one per
// bit of the bucket index): one atomic per distinct bucket per wave instead
// of one per lane, which matters because most tiles of a frame (rays that
// miss) fall into the same few buckets.
__device__ __forceinline__ uint64_t same_bucket_lanes(int key)
{
    uint64_t m = __ballot(1);
#pragma unroll
    for (int b = 0; b < 10; ++b) {
        const bool bit = (key >> b) & 1;
        const uint64_t v = __ballot(bit);
        m &= bit ? v : ~v;
    }
    return m;
}

// G blocks of 1024 threads (about 1024 tiles each).  Block g orders the
// tiles i = g + r * G, an interleaved sample of the whole image, and writes
// its r-th heaviest to position r * G + g: every block's heaviest come
// first, so the order is close to a global one without any coordination
// between blocks, and the positions of block g are exactly its tiles'
// indices, so the result is a permutation.  Each pass takes the costs in
// chunks of kPer per thread, loads issued before any is used.
__global__ __launch_bounds__(1024) void k_tile_order(uint32_t *__restrict__ cost, uint32_t n,
                                                     uint32_t *__restrict__ order)
{
    static_assert(kOrderBuckets <= 1024, "ten-bit bucket index");
    constexpr uint32_t kPer = 4, kChunk = 1024 * kPer;
    __shared__ uint32_t hist[kOrderBuckets];
    const uint32_t t = threadIdx.x, lane = t & 63u;
    const uint32_t g = blockIdx.x, G = gridDim.x;
    const uint32_t ng = n > g ? (n - g + G - 1) / G : 0u;  // tiles of this block
    for (int b = t; b < kOrderBuckets; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    for (uint32_t base = 0; base < ng; base += kChunk) {
        uint32_t c[kPer];
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t r = base + k * 1024u + t;
            c[k] = r < ng ? cost[g + r * G] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            if (base + k * 1024u + t < ng) {
                const int key = order_bucket(c[k]);
                const uint64_t m = same_bucket_lanes(key);
                if (lane == (uint32_t)(__ffsll((unsigned long long)m) - 1))
                    atomicAdd(&hist[key], (uint32_t)__popcll(m));
            }
        }
    }
    __syncthreads();
    {  // exclusive prefix sum over the buckets, in place: wave scans, then the wave totals
        __shared__ uint32_t wsum[16];
        const uint32_t w = t >> 6;
        const uint32_t v = t < (uint32_t)kOrderBuckets ? hist[t] : 0u;
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= (uint32_t)o) x += y;
        }
        if (lane == 63u) wsum[w] = x;
        __syncthreads();
        uint32_t before = 0;
        for (uint32_t k = 0; k < w; ++k) before += wsum[k];
        if (t < (uint32_t)kOrderBuckets) hist[t] = before + x - v;
    }
    __syncthreads();
    for (uint32_t base = 0; base < ng; base += kChunk) {
        uint32_t c[kPer];
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t r = base + k * 1024u + t;
            c[k] = r < ng ? cost[g + r * G] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t r = base + k * 1024u + t;
            if (r < ng) {
                const int key = order_bucket(c[k]);
                const uint64_t m = same_bucket_lanes(key);
                const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)m) - 1);
                uint32_t at = 0;
                if (lane == leader) at = atomicAdd(&hist[key], (uint32_t)__popcll(m));
                at = __shfl(at, (int)leader);
                const uint32_t rank = at + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                order[rank * G + g] = g + r * G;
                cost[g + r * G] = 0u;
            }
        }
    }
}

// Blocks for k_tile_order over n tiles: about 1024 tiles each, at most 256.
static inline uint32_t tile_order_blocks(uint32_t n)
{
    uint32_t b = (n + 1023u) / 1024u;
    return b < 1u ? 1u : (b > 256u ? 256u : b);
}

// ---------------------------------------------------------------------------
// Render jobs: many images of many scenes in one launch.  Each job renders
// spp samples per pixel into its accumulation buffer (+= (r, g, b, samples)),
// either direct lighting or diffuse path tracing.  Tiles of all jobs form one
// queue; persistent blocks take tiles from an atomic counter until the queue
// is empty (every block reaches the exit), or one block per tile.
//
// Path tracing (CSComputeColorGI.hlsl:153-230 loop structure): per sample a
// jittered primary ray, then up to max_bounces diffuse bounces with
// cosine-weighted directions around the face-forward flat normal, throughput
// *= albedo per bounce, a miss adds throughput * sky.  Random numbers come from
// PCG32 (O'Neill 2014) with stream = pixel index and a seed derived from the
// job seed, frame and sample, so a pixel's samples do not depend on which
// other pixels or jobs share the launch.
// ---------------------------------------------------------------------------
struct Pcg32 {
    uint64_t state, inc;
};

__device__ __forceinline__ uint32_t pcg_next(Pcg32 &r)
{
    const uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + r.inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

__device__ __forceinline__ Pcg32 pcg_seed(uint64_t seed, uint64_t stream)
{
    Pcg32 r;
    r.state = 0u;
    r.inc = (stream << 1u) | 1u;
    pcg_next(r);
    r.state += seed;
    pcg_next(r);
    return r;
}

__device__ __forceinline__ float pcg_unit(Pcg32 &r)
{
    return (float)(pcg_next(r) >> 8) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ uint64_t sample_seed(uint64_t seed, uint32_t frame, uint32_t sample)
{
    return seed + 0x9E3779B97F4A7C15ull * ((uint64_t)frame * 65536ull + sample + 1ull);
}

// Cosine-weighted direction around unit n (orthonormal basis of Duff et al.,
// "Building an Orthonormal Basis, Revisited", JCGT 2017).
__device__ __forceinline__ void cosine_dir(float nx, float ny, float nz, float r1, float r2,
                                           float &dx, float &dy, float &dz)
{
    const float phi = 6.28318530717958647692f * r1;
    const float rr = sqrtf(r2);
    const float lx = rr * cosf(phi), ly = rr * sinf(phi);
    const float lz = sqrtf(fmaxf(0.0f, 1.0f - r2));
    const float sign = nz >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + nz);
    const float b = nx * ny * a;
    const float tx = 1.0f + sign * nx * nx * a, ty = sign * b, tz = -sign * nx;
    const float bx = b, by = sign + ny * ny * a, bz = -ny;
    dx = (lx * tx + ly * bx) + lz * nx;
    dy = (lx * ty + ly * by) + lz * ny;
    dz = (lx * tz + ly * bz) + lz * nz;
}

// ---------------------------------------------------------------------------
// Whitted frames: the reference's ray-tracing mode (RayTracerCS.cpp:286-295)
// runs CSIntersections and CSComputeColor (CSComputeColorRT.hlsl:18-262) once
// per pass over a ray buffer of every pixel.  Two forms here, same results:
//   - one pass (k_render_whitted): one lane carries its pixel through every
//     pass, rays and the accumulator in registers;
//   - wavefront (k_whitted_first + k_whitted_next per bounce): pass 0 per
//     pixel, then each bounce pass over a compacted queue of the rays still
//     alive, so waves stay full when most paths have ended.  The accumulator
//     lives in the frame buffer; each queued ray is one 64-byte record.
// ---------------------------------------------------------------------------
struct ShadeView {
    const float *nrm;         // the scene's vertex normals (null: flat shading only)
    const uint32_t *idx;      // the scene's vertex indices, 3 per triangle
    const uint32_t *mat;      // the scene's material id per triangle (null: material 0)
    const rt_material *mats;  // the scene's materials (null: defmat)
};

struct WhittedParams {
    rt_camera cam;
    uint32_t width, height, bounces, switches;
    float light[3];
    int32_t env_flag;
    float env[3];
    uint32_t frame;
    uint64_t seed;
    rt_material defmat;
    const uint32_t *order;  // tile of each block, heaviest first; null: block index order
    uint32_t *cost;         // per tile: max over its lanes of box + triangle tests (atomicMax)
    uint32_t groups;        // one-pass kernel: live paths on lane groups (uncounted LBVH frames)
};

// A live path between passes: origin, direction, reflective factor, the id
// of the triangle it leaves (self), its pixel and its random stream.
struct WRay {
    float ox, oy, oz, dx, dy, dz, fx, fy, fz;
    int32_t self;
    uint32_t pix;
    Pcg32 rng;
};

// Queue record of a WRay (64 bytes; the stream increment is 2 pix + 1).
struct alignas(16) RayRec {
    float4 a;  // ox oy oz dx
    float4 b;  // dy dz fx fy
    float4 c;  // fz, self, pix, unused
    uint4 d;   // rng state lo, hi, unused
};

__device__ __forceinline__ void put_ray(RayRec *q, uint32_t i, const WRay &r)
{
    RayRec e;
    e.a = make_float4(r.ox, r.oy, r.oz, r.dx);
    e.b = make_float4(r.dy, r.dz, r.fx, r.fy);
    e.c = make_float4(r.fz, __int_as_float(r.self), __uint_as_float(r.pix), 0.0f);
    e.d = make_uint4((uint32_t)r.rng.state, (uint32_t)(r.rng.state >> 32), 0u, 0u);
    q[i] = e;
}

__device__ __forceinline__ WRay get_ray(const RayRec *q, uint32_t i)
{
    const RayRec e = q[i];
    WRay r;
    r.ox = e.a.x;
    r.oy = e.a.y;
    r.oz = e.a.z;
    r.dx = e.a.w;
    r.dy = e.b.x;
    r.dz = e.b.y;
    r.fx = e.b.z;
    r.fy = e.b.w;
    r.fz = e.c.x;
    r.self = __float_as_int(e.c.y);
    r.pix = __float_as_uint(e.c.z);
    r.rng.state = (uint64_t)e.d.x | ((uint64_t)e.d.y << 32);
    r.rng.inc = ((uint64_t)r.pix << 1u) | 1u;
    return r;
}

// Slot for this lane in an append-only queue: one atomic per wave.  Every
// lane of the wave must call it; lanes with want == false get no slot.
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool want)
{
    const uint64_t mask = __ballot(want);
    if (mask == 0) return 0u;
    const uint32_t lane = __lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    return (ax * bx + ay * by) + az * bz;
}

// x^49 as x^32 * x^16 * x by repeated squaring (the shader's pow(x, 49)).
__device__ __forceinline__ float pow49(float x)
{
    const float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8, x32 = x16 * x16;
    return (x32 * x16) * x;
}

struct WCount {
    uint32_t boxes = 0, ntris = 0, nshadow = 0;
};

// One intersection + colour pass of ray r.  Adds the pass's colour into acc
// (or sets white), and returns true when the path continues, with r turned
// into the continuation ray.  first: store the hit and shadow result.
// With QUAD the qk lanes of an aligned group (four by default; quad_lane:
// this lane's place) hold the same path; each traces its share of the pass's
// rays (traverse_quad), the group takes the closest hit, and the rest of the
// pass runs on all of them.
template <int FORM, bool QUAD = false, class Stack>
__device__ __forceinline__ bool whitted_pass(const SceneView &sv, const ShadeView &sh, const WhittedParams &wp,
                                             int32_t root_ref, const float *root_box, WRay &r, float *acc,
                                             bool &white, rt_hit24 *hit_out, uint8_t *occ_out, Stack &stk,
                                             WCount &cnt, uint32_t quad_lane = 0, uint32_t qk = 4)
{
    TraceOut h;
    if constexpr (QUAD) {
        uint32_t rank;
        h = traverse_quad<false>(sv.nodes, sv.tris, root_ref, root_box, quad_lane, kQuadDepth, r.ox, r.oy, r.oz,
                                 r.dx, r.dy, r.dz, 10000000000000000.0f, r.self, stk, rank, qk - 1u);
        cnt.boxes += h.boxes;
        cnt.ntris += h.tris;
        group_closest(h, rank, qk);
    } else {
        h = traverse_form<FORM, false>(sv.nodes, sv.tris, root_ref, root_box, r.ox, r.oy, r.oz, r.dx, r.dy, r.dz,
                                       10000000000000000.0f, r.self, stk);
        cnt.boxes += h.boxes;
        cnt.ntris += h.tris;
    }
    if (hit_out) {
        store_hit(hit_out, h);
        *occ_out = 2;
    }
    // A miss under environment-mapping flag 0 paints the pixel white and
    // leaves the ray as it is, so every later pass repeats it (:30-35).
    if (h.tri < wp.env_flag) {
        white = true;
        return false;
    }
    if (h.tri < 0) {  // environment colour, ray ended (:238-248)
        acc[0] = acc[0] + r.fx * wp.env[0];
        acc[1] = acc[1] + r.fy * wp.env[1];
        acc[2] = acc[2] + r.fz * wp.env[2];
        acc[3] = acc[3] + 1.0f;
        return false;
    }
    const uint32_t mid = sh.mat ? sh.mat[h.tri] : 0u;
    const rt_material M = sh.mats ? sh.mats[mid] : wp.defmat;
    // shading normal: interpolated vertex normals or the face normal (:56-77)
    float nx, ny, nz;
    if (wp.switches & RT_WHITTED_PHONG) {
        const float *na = sh.nrm + 3 * (size_t)sh.idx[3 * (size_t)h.tri];
        const float *nb = sh.nrm + 3 * (size_t)sh.idx[3 * (size_t)h.tri + 1];
        const float *nc = sh.nrm + 3 * (size_t)sh.idx[3 * (size_t)h.tri + 2];
        const float w = 1.0f - (h.u + h.v);  // Vec3BaryCentric (Intersection.hlsl:98-108)
        nx = (na[0] * w + nb[0] * h.u) + nc[0] * h.v;
        ny = (na[1] * w + nb[1] * h.u) + nc[1] * h.v;
        nz = (na[2] * w + nb[2] * h.u) + nc[2] * h.v;
        const float nl = sqrtf(dot3(nx, ny, nz, nx, ny, nz));
        nx = nx / nl;
        ny = ny / nl;
        nz = nz / nl;
    } else {
        flat_normal(sv.tris[h.pos], nx, ny, nz);
    }
    // facing the ray (:124-131)
    float mx = nx, my = ny, mz = nz;
    const bool inverse = dot3(mx, my, mz, r.dx, r.dy, r.dz) > 0.0f;
    if (inverse) {
        mx = -mx;
        my = -my;
        mz = -mz;
    }
    const float hx = r.ox + h.t * r.dx, hy = r.oy + h.t * r.dy, hz = r.oz + h.t * r.dz;
    float lx = hx - wp.light[0], ly = hy - wp.light[1], lz = hz - wp.light[2];
    const float dist = sqrtf(dot3(lx, ly, lz, lx, ly, lz));
    lx = lx / dist;
    ly = ly / dist;
    lz = lz / dist;
    // normalize(reflect(d, n')), reflect(i, n) = i - 2 dot(n, i) n (:141-142)
    const float k2 = 2.0f * dot3(mx, my, mz, r.dx, r.dy, r.dz);
    float rx = r.dx - k2 * mx, ry = r.dy - k2 * my, rz = r.dz - k2 * mz;
    const float rl = sqrtf(dot3(rx, ry, rz, rx, ry, rz));
    rx = rx / rl;
    ry = ry / rl;
    rz = rz / rl;
    // light and viewer on the same side of the surface: shadow ray to the
    // light, then diffuse and specular terms (:145-176)
    bool lit = false;
    float D = 0.0f, S = 0.0f;
    if ((dot3(nx, ny, nz, lx, ly, lz) * dot3(nx, ny, nz, r.dx, r.dy, r.dz)) > 0.0f) {
        bool unoccluded = true;
        if (wp.switches & RT_WHITTED_SHADOWS) {
            TraceOut s;
            if constexpr (QUAD) {
                uint32_t rank;
                s = traverse_quad<true>(sv.nodes, sv.tris, root_ref, root_box, quad_lane, kQuadDepth, hx, hy, hz,
                                        -lx, -ly, -lz, dist, h.tri, stk, rank, qk - 1u);
                cnt.boxes += s.boxes;
                cnt.ntris += s.tris;
                s.tri = group_any(s.tri >= 0, qk) ? 0 : -1;
            } else {
                s = traverse_form<FORM, true>(sv.nodes, sv.tris, root_ref, root_box, hx, hy, hz, -lx, -ly, -lz, dist,
                                              h.tri, stk);
                cnt.boxes += s.boxes;
                cnt.ntris += s.tris;
            }
            cnt.nshadow += quad_lane == 0 ? 1u : 0u;
            unoccluded = s.tri < 0;
            if (occ_out) *occ_out = unoccluded ? 0 : 1;
        }
        if (unoccluded) {
            lit = true;
            const float sp = -dot3(lx, ly, lz, rx, ry, rz);
            S = pow49(sp > 0.0f ? sp : 0.0f);
            const float dd = -dot3(mx, my, mz, lx, ly, lz);
            D = dd > 0.0f ? dd : 0.0f;
        }
    }
    // colour (:181-198): (ambient + D * diffuse.w) * diffuse
    // + S * (1 - diffuse.w) * specular, times (factor, 1)
    const float tw = M.diffuse[3];
    const float gloss = (wp.switches & RT_WHITTED_GLOSS) ? 1.0f : 0.0f;
    const float f4[4] = {r.fx, r.fy, r.fz, 1.0f};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float Dc = lit ? D : (c == 3 ? 1.0f : 0.0f);
        const float Sc = lit ? S : (c == 3 ? 1.0f : 0.0f);
        const float spc = c == 3 ? M.specular[3] * gloss : M.specular[c];
        const float dcol = (M.ambient[c] + Dc * tw) * M.diffuse[c];
        const float scol = (Sc * (1.0f - tw)) * spc;
        acc[c] = acc[c] + (dcol + scol) * f4[c];
    }
    // continuation ray (:200-235): reflect with probability alpha, else
    // refract(d, n', eta) with eta = ambient.w entering, its inverse leaving;
    // total internal reflection reflects
    const float u = pcg_unit(r.rng);
    if (M.alpha >= u) {
        r.fx = r.fx * M.specular[0];
        r.fy = r.fy * M.specular[1];
        r.fz = r.fz * M.specular[2];
        r.dx = rx;
        r.dy = ry;
        r.dz = rz;
    } else {
        float eta = M.ambient[3];
        if (!inverse) eta = 1.0f / eta;
        // refract(i, n, eta): cosi = dot(-i, n), cost2 = 1 - eta^2 (1 - cosi^2),
        // t = (eta i + (eta cosi - sqrt|cost2|) n) * (cost2 > 0)
        const float cosi = -dot3(r.dx, r.dy, r.dz, mx, my, mz);
        const float cost2 = 1.0f - (eta * eta) * (1.0f - cosi * cosi);
        const float g = eta * cosi - sqrtf(fabsf(cost2));
        const float keepr = cost2 > 0.0f ? 1.0f : 0.0f;
        const float tx_ = (eta * r.dx + g * mx) * keepr;
        const float ty_ = (eta * r.dy + g * my) * keepr;
        const float tz_ = (eta * r.dz + g * mz) * keepr;
        const float tl = sqrtf(dot3(tx_, ty_, tz_, tx_, ty_, tz_));
        if (tl < 0.000001f) {
            r.dx = rx;
            r.dy = ry;
            r.dz = rz;
        } else {
            r.dx = tx_ / tl;
            r.dy = ty_ / tl;
            r.dz = tz_ / tl;
            r.fx = r.fx * M.diffuse[0];
            r.fy = r.fy * M.diffuse[1];
            r.fz = r.fz * M.diffuse[2];
        }
    }
    r.ox = hx;
    r.oy = hy;
    r.oz = hz;
    r.self = h.tri;
    return true;
}

// The pixel's primary ray (CSPrimaryRays.hlsl:35-53, no jitter) and stream.
__device__ __forceinline__ WRay whitted_primary(const WhittedParams &wp, uint32_t px, uint32_t py)
{
    WRay r;
    r.pix = py * wp.width + px;
    primary_ray(wp.cam, px, py, (float)wp.width, (float)wp.height, 0.0f, 0.0f, r.dx, r.dy, r.dz);
    r.ox = wp.cam.pos[0];
    r.oy = wp.cam.pos[1];
    r.oz = wp.cam.pos[2];
    r.fx = r.fy = r.fz = 1.0f;  // reflective factor (:52)
    r.self = -1;
    r.rng = pcg_seed(sample_seed(wp.seed, wp.frame, 0u), r.pix);
    return r;
}

// One-pass form: every pass of a pixel in one lane; for uncounted frames
// over the LBVH, a wave's live paths are shared out to lane groups before
// each pass (whitted_group_passes), as the split form's remaining passes do.
// Same frame either way.
template <int BLOCK, int KLDS, int FORM>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_render_whitted(SceneView sv, ShadeView sh, WhittedParams wp,
                                                          float4 *__restrict__ rgba,
                                                          rt_hit24 *__restrict__ hits,
                                                          uint8_t *__restrict__ occ, int keep, int count,
                                                          unsigned long long *stats)
{
    __shared__ uint2 lds[KLDS * BLOCK];
    RayStack<BLOCK, KLDS> stk;
    stk.lds = &lds[threadIdx.x];
    constexpr int TW = Tile<BLOCK>::W, TH = Tile<BLOCK>::H;
    const uint32_t tiles_x = (wp.width + TW - 1) / TW;
    const uint32_t tile = wp.order ? wp.order[blockIdx.x] : blockIdx.x;
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    uint32_t ox_, oy_;
    Tile<BLOCK>::offset(threadIdx.x, ox_, oy_);
    const uint32_t px = tx * TW + ox_;
    const uint32_t py = ty * TH + oy_;
    int32_t root_ref;
    float root_box[6];
    load_root(sv.hdr, root_ref, root_box);
    WCount cnt;
    uint32_t nbounce = 0;
    const bool valid = px < wp.width && py < wp.height;
    if (wp.groups) {  // the wave's live paths on lane groups
        WRay r;
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        bool white = false, alive = valid;
        if (valid) r = whitted_primary(wp, px, py);
        if (keep && alive) {  // pass 0 keeps its hit: one lane per pixel
            alive = whitted_pass<FORM>(sv, sh, wp, root_ref, root_box, r, acc, white, &hits[r.pix], &occ[r.pix], stk,
                                       cnt);
            if (!alive)
                rgba[r.pix] = white ? make_float4(1.0f, 1.0f, 1.0f, 1.0f)
                                    : make_float4(acc[0], acc[1], acc[2], acc[3]);
        }
        whitted_group_passes<FORM>(sv, sh, wp, root_ref, root_box, keep ? 1u : 0u, r, acc, white, alive, rgba, stk,
                                   cnt, nbounce);
        if (alive)
            rgba[r.pix] = white ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : make_float4(acc[0], acc[1], acc[2], acc[3]);
    } else if (valid) {
        WRay r = whitted_primary(wp, px, py);
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // g_uAccumulation
        bool white = false;
        for (uint32_t pass = 0; pass <= wp.bounces; ++pass) {
            nbounce += pass > 0 ? 1u : 0u;
            const bool first = keep && pass == 0;
            if (!whitted_pass<FORM>(sv, sh, wp, root_ref, root_box, r, acc, white, first ? &hits[r.pix] : nullptr,
                                    first ? &occ[r.pix] : nullptr, stk, cnt))
                break;
        }
        rgba[r.pix] = white ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : make_float4(acc[0], acc[1], acc[2], acc[3]);
    }
    if (count) {
        count_add(stats, 0, cnt.boxes);
        count_add(stats, 1, cnt.ntris);
    }
    count_add(stats, 2, cnt.nshadow);
    count_add(stats, 3, nbounce);
    if (wp.cost) tile_cost_max(wp.cost, tile, cnt.boxes + cnt.ntris);
}

// Wavefront form, pass 0: one lane per pixel; live paths go to queue q.
template <int BLOCK, int KLDS, int FORM>
__global__ __launch_bounds__(BLOCK) void k_whitted_first(SceneView sv, ShadeView sh, WhittedParams wp,
                                                         float4 *__restrict__ rgba, rt_hit24 *__restrict__ hits,
                                                         uint8_t *__restrict__ occ, int keep, int count,
                                                         RayRec *__restrict__ q, uint32_t *qcount,
                                                         unsigned long long *stats)
{
    __shared__ uint2 lds[KLDS * BLOCK];
    RayStack<BLOCK, KLDS> stk;
    stk.lds = &lds[threadIdx.x];
    constexpr int TW = Tile<BLOCK>::W, TH = Tile<BLOCK>::H;
    const uint32_t tiles_x = (wp.width + TW - 1) / TW;
    const uint32_t tile = wp.order ? wp.order[blockIdx.x] : blockIdx.x;
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    uint32_t ox_, oy_;
    Tile<BLOCK>::offset(threadIdx.x, ox_, oy_);
    const uint32_t px = tx * TW + ox_;
    const uint32_t py = ty * TH + oy_;
    int32_t root_ref;
    float root_box[6];
    load_root(sv.hdr, root_ref, root_box);
    WCount cnt;
    WRay r;
    bool live = false;
    if (px < wp.width && py < wp.height) {
        r = whitted_primary(wp, px, py);
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        bool white = false;
        live = whitted_pass<FORM>(sv, sh, wp, root_ref, root_box, r, acc, white, keep ? &hits[r.pix] : nullptr,
                                  keep ? &occ[r.pix] : nullptr, stk, cnt);
        live = live && wp.bounces > 0;
        rgba[r.pix] = white ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : make_float4(acc[0], acc[1], acc[2], acc[3]);
    }
    const uint32_t slot = wave_append(qcount, live);
    if (live) put_ray(q, slot, r);
    if (count) {
        count_add(stats, 0, cnt.boxes);
        count_add(stats, 1, cnt.ntris);
    }
    count_add(stats, 2, cnt.nshadow);
    if (wp.cost) tile_cost_max(wp.cost, tile, cnt.boxes + cnt.ntris);
}

// Wavefront form, a bounce pass: the *in_count rays of queue in (at most one
// per pixel), their colour added into rgba, live ones to queue out.
template <int BLOCK, int KLDS, int FORM>
__global__ __launch_bounds__(BLOCK) void k_whitted_next(SceneView sv, ShadeView sh, WhittedParams wp,
                                                        float4 *__restrict__ rgba, const RayRec *__restrict__ qin,
                                                        const uint32_t *in_count, RayRec *__restrict__ qout,
                                                        uint32_t *out_count, int last, int count,
                                                        unsigned long long *stats)
{
    __shared__ uint2 lds[KLDS * BLOCK];
    RayStack<BLOCK, KLDS> stk;
    stk.lds = &lds[threadIdx.x];
    int32_t root_ref;
    float root_box[6];
    load_root(sv.hdr, root_ref, root_box);
    WCount cnt;
    uint32_t nbounce = 0;
    const uint32_t n = *in_count;
    for (uint32_t base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
        const uint32_t i = base + threadIdx.x;
        WRay r;
        bool live = false;
        if (i < n) {
            r = get_ray(qin, i);
            const float4 a = rgba[r.pix];
            float acc[4] = {a.x, a.y, a.z, a.w};
            bool white = false;
            nbounce += 1;
            live = whitted_pass<FORM>(sv, sh, wp, root_ref, root_box, r, acc, white, nullptr, nullptr, stk, cnt);
            live = live && !last;
            rgba[r.pix] = white ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : make_float4(acc[0], acc[1], acc[2], acc[3]);
        }
        const uint32_t slot = wave_append(out_count, live);
        if (live) put_ray(qout, slot, r);
    }
    if (count) {
        count_add(stats, 0, cnt.boxes);
        count_add(stats, 1, cnt.ntris);
    }
    count_add(stats, 2, cnt.nshadow);
    count_add(stats, 3, nbounce);
}

// Split form, after k_whitted_first: each live path of queue qin runs all of
// its remaining passes in one lane.  The bounce passes of the wavefront form
// each hold few rays (a fraction of a GPU's waves), so each lasts as long as
// its slowest ray; one launch for every bounce pays that tail once.  Same
// passes in the same order per pixel as the other forms, so the same frame.
template <int BLOCK, int KLDS, int FORM>
__global__ __launch_bounds__(BLOCK) void k_whitted_rest(SceneView sv, ShadeView sh, WhittedParams wp,
                                                        float4 *__restrict__ rgba, const RayRec *__restrict__ qin,
                                                        const uint32_t *in_count, int count,
                                                        unsigned long long *stats)
{
    __shared__ uint2 lds[KLDS * BLOCK];
    RayStack<BLOCK, KLDS> stk;
    stk.lds = &lds[threadIdx.x];
    int32_t root_ref;
    float root_box[6];
    load_root(sv.hdr, root_ref, root_box);
    WCount cnt;
    uint32_t nbounce = 0;
    const uint32_t n = *in_count;
    for (uint32_t base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
        const uint32_t i = base + threadIdx.x;
        if (i < n) {
            WRay r = get_ray(qin, i);
            const float4 a = rgba[r.pix];
            float acc[4] = {a.x, a.y, a.z, a.w};
            bool white = false;
            for (uint32_t pass = 1; pass <= wp.bounces; ++pass) {
                nbounce += 1;
                if (!whitted_pass<FORM>(sv, sh, wp, root_ref, root_box, r, acc, white, nullptr, nullptr, stk, cnt))
                    break;
            }
            rgba[r.pix] = white ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : make_float4(acc[0], acc[1], acc[2], acc[3]);
        }
    }
    if (count) {
        count_add(stats, 0, cnt.boxes);
        count_add(stats, 1, cnt.ntris);
    }
    count_add(stats, 2, cnt.nshadow);
    count_add(stats, 3, nbounce);
}

// The passes first_pass .. wp.bounces of the paths a wave holds, one per lane
// (alive: this lane holds a live path r with accumulator acc): one lane per
// path while more than 32 of the wave's paths live, then the m live paths
// shared out to aligned groups of 64 / m lanes (a power of two, at most 16)
// that take the path's state by shuffles and run the pass as a quad does.
// Every lane of the wave calls it.  A path's colour is written when it ends;
// on return alive says the path went through the last pass (the caller
// writes it).
template <int FORM, class Stack>
__device__ __forceinline__ void whitted_group_passes(const SceneView &sv, const ShadeView &sh, const WhittedParams &wp,
                                                     int32_t root_ref, const float *root_box, uint32_t first_pass,
                                                     WRay &r, float *acc, bool &white, bool &alive,
                                                     float4 *__restrict__ rgba, Stack &stk, WCount &cnt,
                                                     uint32_t &nbounce)
{
    const uint32_t lane = __lane_id();
    for (uint32_t pass = first_pass; pass <= wp.bounces; ++pass) {
        const uint64_t live = __ballot(alive);
        if (live == 0) break;
        const uint32_t m = (uint32_t)__popcll(live);
        if (m > 32) {  // one lane per path
            if (alive) {
                nbounce += pass > 0 ? 1u : 0u;
                alive = whitted_pass<FORM>(sv, sh, wp, root_ref, root_box, r, acc, white, nullptr, nullptr, stk,
                                           cnt);
                if (!alive)
                    rgba[r.pix] = white ? make_float4(1.0f, 1.0f, 1.0f, 1.0f)
                                        : make_float4(acc[0], acc[1], acc[2], acc[3]);
            }
            continue;
        }
        const uint32_t lk = m > 16 ? 1u : m > 8 ? 2u : m > 4 ? 3u : 4u;  // log2 k
        const uint32_t k = 1u << lk, g = lane >> lk, s = lane & (k - 1u);
        // group g takes the g-th live path
        uint64_t rest = live;
        uint32_t src = 0;
        for (uint32_t j = 0; j < m; ++j) {
            const uint32_t b = (uint32_t)__builtin_ctzll(rest);
            rest &= rest - 1ull;
            if (j == g) src = b;
        }
        r.ox = __shfl(r.ox, src);
        r.oy = __shfl(r.oy, src);
        r.oz = __shfl(r.oz, src);
        r.dx = __shfl(r.dx, src);
        r.dy = __shfl(r.dy, src);
        r.dz = __shfl(r.dz, src);
        r.fx = __shfl(r.fx, src);
        r.fy = __shfl(r.fy, src);
        r.fz = __shfl(r.fz, src);
        r.self = __shfl(r.self, src);
        r.pix = __shfl(r.pix, src);
        r.rng.state = ((uint64_t)(uint32_t)__shfl((uint32_t)(r.rng.state >> 32), src) << 32) |
                      (uint32_t)__shfl((uint32_t)r.rng.state, src);
        r.rng.inc = ((uint64_t)(uint32_t)__shfl((uint32_t)(r.rng.inc >> 32), src) << 32) |
                    (uint32_t)__shfl((uint32_t)r.rng.inc, src);
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = __shfl(acc[c], src);
        white = __shfl(white ? 1 : 0, src) != 0;
        bool cont = false;
        if (g < m) {  // the whole group together
            if (s == 0 && pass > 0) nbounce += 1;
            cont = whitted_pass<FORM, true>(sv, sh, wp, root_ref, root_box, r, acc, white, nullptr, nullptr,
                                            stk, cnt, s, k);
            if (!cont && s == 0)
                rgba[r.pix] = white ? make_float4(1.0f, 1.0f, 1.0f, 1.0f)
                                    : make_float4(acc[0], acc[1], acc[2], acc[3]);
        }
        alive = g < m && s == 0 && cont;  // the group's first lane owns the path from here
    }
}

// k_whitted_rest for uncounted frames with the live paths of a wave on lane
// groups: the wave takes 64 paths, one per lane; before each pass the m
// paths still alive are shared out to aligned groups of k = 64 / m lanes (a
// power of two, at most 16; one lane each above 32 paths), whose lanes take
// the path's state by shuffles and run the pass together as a quad does
// (whitted_pass with QUAD and group size k).  A wave's last passes hold few
// paths, and those are the chains the launch waits for.  Same frame.
// (Four waves per SIMD requested: the one-lane and the group pass together
// need about 147 VGPRs, three waves; at 128 with 20 spilled it runs 25 %
// faster on the soup.)
template <int BLOCK, int KLDS, int FORM>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_whitted_rest_group(SceneView sv, ShadeView sh, WhittedParams wp,
                                                              float4 *__restrict__ rgba,
                                                              const RayRec *__restrict__ qin,
                                                              const uint32_t *in_count, unsigned long long *stats)
{
    __shared__ uint2 lds[KLDS * BLOCK];
    RayStack<BLOCK, KLDS> stk;
    stk.lds = &lds[threadIdx.x];
    int32_t root_ref;
    float root_box[6];
    load_root(sv.hdr, root_ref, root_box);
    WCount cnt;
    uint32_t nbounce = 0;
    const uint32_t n = *in_count;
    const uint32_t lane = __lane_id();
    const uint32_t first = (blockIdx.x * BLOCK + threadIdx.x) & ~63u;  // the wave's first path
    for (uint32_t base = first; base < n; base += gridDim.x * BLOCK) {
        const uint32_t i = base + lane;
        bool alive = i < n;  // this lane owns a live path
        WRay r;
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        bool white = false;
        if (alive) {
            r = get_ray(qin, i);
            const float4 a = rgba[r.pix];
            acc[0] = a.x;
            acc[1] = a.y;
            acc[2] = a.z;
            acc[3] = a.w;
        }
        whitted_group_passes<FORM>(sv, sh, wp, root_ref, root_box, 1u, r, acc, white, alive, rgba, stk, cnt,
                                   nbounce);
        if (alive)  // through the last pass
            rgba[r.pix] = white ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : make_float4(acc[0], acc[1], acc[2], acc[3]);
    }
    count_add(stats, 2, cnt.nshadow);
    count_add(stats, 3, nbounce);
}

// k_whitted_rest with a quad per path: the four lanes of a quad share each
// pass's closest-hit and shadow traversals (the heaviest paths' chains of
// dependent iterations are what the launch waits for).  Same frame.
template <int BLOCK, int KLDS, int FORM>
__global__ __launch_bounds__(BLOCK) void k_whitted_rest_quad(SceneView sv, ShadeView sh, WhittedParams wp,
                                                             float4 *__restrict__ rgba,
                                                             const RayRec *__restrict__ qin,
                                                             const uint32_t *in_count, unsigned long long *stats)
{
    __shared__ uint2 lds[KLDS * BLOCK];
    RayStack<BLOCK, KLDS> stk;
    stk.lds = &lds[threadIdx.x];
    int32_t root_ref;
    float root_box[6];
    load_root(sv.hdr, root_ref, root_box);
    WCount cnt;
    uint32_t nbounce = 0;
    const uint32_t n = *in_count;
    const uint32_t s = threadIdx.x & 3u;
    constexpr uint32_t kPaths = BLOCK / 4;  // paths per block and round
    for (uint32_t base = blockIdx.x * kPaths; base < n; base += gridDim.x * kPaths) {
        const uint32_t i = base + (threadIdx.x >> 2);
        if (i < n) {  // the same for the four lanes of a quad
            WRay r = get_ray(qin, i);
            const float4 a = rgba[r.pix];
            float acc[4] = {a.x, a.y, a.z, a.w};
            bool white = false;
            for (uint32_t pass = 1; pass <= wp.bounces; ++pass) {
                nbounce += s == 0 ? 1u : 0u;
                if (!whitted_pass<FORM, true>(sv, sh, wp, root_ref, root_box, r, acc, white, nullptr, nullptr, stk,
                                              cnt, s))
                    break;
            }
            if (s == 0)
                rgba[r.pix] =
                    white ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : make_float4(acc[0], acc[1], acc[2], acc[3]);
        }
    }
    count_add(stats, 2, cnt.nshadow);
    count_add(stats, 3, nbounce);
}

__global__ void k_zero_words(uint32_t *p, uint32_t n)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

struct RenderJob {
    const TravNode *nodes;
    const TravTri *tris;
    const SceneHdr *hdr;
    float4 *accum;
    const uint32_t *order;  // tile at each position of the job's queue; null: row-major
    rt_camera cam;
    uint32_t width, height, tiles_x, tile_off;
    uint32_t local_base, mode, spp, max_bounces, nlights;
    uint32_t lbvh;  // the nodes carry the LBVH's heap indices (lane groups may split a ray)
    float albedo, sky, ambient, pad;
    uint64_t seed;
    uint32_t frame, pad2;
    float lights[24];
};

template <int BLOCK, int KLDS, bool PERSISTENT, int FORM>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_render_jobs(const RenderJob *__restrict__ jobs,
                                                       uint32_t njobs, uint32_t total_tiles,
                                                       uint32_t tile_base, uint32_t *tile_counter,
                                                       const uint32_t *frame_counter,
                                                       unsigned long long *stats, uint32_t count)
{
    __shared__ uint2 lds[KLDS * BLOCK];
    __shared__ uint32_t s_tile;
    RayStack<BLOCK, KLDS> stk;
    stk.lds = &lds[threadIdx.x];
    constexpr int TW = Tile<BLOCK>::W, TH = Tile<BLOCK>::H;
    uint32_t boxes = 0, ntris = 0, nshadow = 0, nbounce = 0, nprimary = 0;
    const uint32_t frame_add = frame_counter ? *frame_counter : 0u;
    uint32_t tile = tile_base + blockIdx.x;
    while (true) {
        if (PERSISTENT) {
            if (threadIdx.x == 0) s_tile = atomicAdd(tile_counter, 1u);
            __syncthreads();
            tile = s_tile;
            __syncthreads();
        }
        if (tile >= total_tiles) break;
        uint32_t lo = 0, hi = njobs;  // last job with tile_off <= tile
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (jobs[mid].tile_off <= tile) lo = mid;
            else hi = mid;
        }
        const RenderJob &J = jobs[lo];
        const uint32_t rank = tile - J.tile_off + J.local_base;
        const uint32_t lt = J.order ? J.order[rank] : rank;
        uint32_t ox_, oy_;
        Tile<BLOCK>::offset(threadIdx.x, ox_, oy_);
        const uint32_t px = (lt % J.tiles_x) * TW + ox_;
        const uint32_t py = (lt / J.tiles_x) * TH + oy_;
        const bool valid = px < J.width && py < J.height;
        if (!count && J.lbvh) {
            // Uncounted LBVH jobs: the lanes of a wave together.  Direct
            // samples shade with shade_wave (shadow rays on lane groups); GI
            // samples send each bounce's rays out to lane groups once at most
            // 32 paths of the wave live (the group takes the ray from its
            // path's lane, traces it as a quad does, and the path's lane
            // takes the closest hit back).  Same images; box and triangle
            // counts are not kept.
            const SceneView sv{J.nodes, J.tris, J.hdr};
            int32_t root_ref;
            float root_box[6];
            load_root(J.hdr, root_ref, root_box);
            const uint32_t pix = valid ? py * J.width + px : 0u;
            const float fw = (float)J.width, fh = (float)J.height;
            const uint32_t lane = __lane_id();
            const uint64_t below = (1ull << lane) - 1ull;
            uint32_t gboxes = 0, gtris = 0;  // not reported
            float sum = 0.0f;
            for (uint32_t s = 0; s < J.spp; ++s) {
                Pcg32 rng{0, 0};
                float dx = 0.0f, dy = 0.0f, dz = 0.0f;
                float ox = J.cam.pos[0], oy = J.cam.pos[1], oz = J.cam.pos[2];
                if (valid) {
                    rng = pcg_seed(sample_seed(J.seed, J.frame + frame_add, s), pix);
                    const float jx = 2.0f * pcg_unit(rng) - 1.0f;
                    const float jy = 2.0f * pcg_unit(rng) - 1.0f;
                    primary_ray(J.cam, px, py, fw, fh, jx, jy, dx, dy, dz);
                }
                if (J.mode == 0) {
                    TraceOut h;
                    h.tri = -1;
                    if (valid) {
                        h = traverse_form<FORM, false>(sv.nodes, sv.tris, root_ref, root_box, ox, oy, oz, dx, dy,
                                                       dz, 10000000000000000.0f, -1, stk);
                        nprimary += 1;
                    }
                    const bool lit = valid && h.tri >= 0;
                    const float shaded = shade_wave<FORM, false>(sv, root_ref, root_box, h, lit, ox, oy, oz, dx, dy,
                                                                 dz, J.lights, J.nlights, J.ambient, J.albedo, stk,
                                                                 gboxes, gtris, nshadow, nullptr);
                    if (valid) sum += lit ? shaded : J.sky;
                    continue;
                }
                float thr = 1.0f, L = 0.0f;
                int32_t self = -1;
                bool alive = valid;
                for (uint32_t b = 0; b <= J.max_bounces; ++b) {
                    const uint64_t live = __ballot(alive);
                    if (live == 0) break;
                    const uint32_t m = (uint32_t)__popcll(live);
                    TraceOut h;
                    h.tri = -1;
                    h.t = 0.0f;
                    h.pos = 0;
                    if (m > 32) {
                        if (alive)
                            h = traverse_form<FORM, false>(sv.nodes, sv.tris, root_ref, root_box, ox, oy, oz, dx, dy,
                                                           dz, 10000000000000000.0f, self, stk);
                    } else {
                        const uint32_t lk = m > 16 ? 1u : m > 8 ? 2u : m > 4 ? 3u : 4u;  // log2 k
                        const uint32_t k = 1u << lk, g = lane >> lk;
                        uint64_t rest = live;
                        uint32_t src = 0;
                        for (uint32_t j = 0; j < m; ++j) {
                            const uint32_t bit = (uint32_t)__builtin_ctzll(rest);
                            rest &= rest - 1ull;
                            if (j == g) src = bit;
                        }
                        const float gox = __shfl(ox, src), goy = __shfl(oy, src), goz = __shfl(oz, src);
                        const float gdx = __shfl(dx, src), gdy = __shfl(dy, src), gdz = __shfl(dz, src);
                        const int32_t gself = __shfl(self, src);
                        TraceOut gh;
                        gh.tri = -1;
                        gh.t = 0.0f;
                        gh.pos = 0;
                        if (g < m) {  // the whole group together
                            uint32_t rank;
                            gh = traverse_quad<false>(sv.nodes, sv.tris, root_ref, root_box, lane & (k - 1u),
                                                      kQuadDepth, gox, goy, goz, gdx, gdy, gdz,
                                                      10000000000000000.0f, gself, stk, rank, k - 1u);
                            group_closest(gh, rank, k);
                        }
                        // the path of rank r takes its hit from group r's first lane
                        const uint32_t from = alive ? (uint32_t)__popcll(live & below) * k : 0u;
                        const int32_t tri = __shfl(gh.tri, from);
                        const float t = __shfl(gh.t, from);
                        const int32_t pos = __shfl(gh.pos, from);
                        h.tri = tri;
                        h.t = t;
                        h.pos = pos;
                    }
                    if (alive) {
                        if (b == 0) nprimary += 1;
                        else nbounce += 1;
                        if (h.tri < 0) {
                            L = L + thr * J.sky;
                            alive = false;
                        } else if (b == J.max_bounces) {
                            alive = false;
                        } else {
                            float nx, ny, nz;
                            flat_normal(sv.tris[h.pos], nx, ny, nz);
                            if (((nx * dx + ny * dy) + nz * dz) > 0.0f) {
                                nx = -nx;
                                ny = -ny;
                                nz = -nz;
                            }
                            ox = ox + h.t * dx;
                            oy = oy + h.t * dy;
                            oz = oz + h.t * dz;
                            const float r1 = pcg_unit(rng), r2 = pcg_unit(rng);
                            cosine_dir(nx, ny, nz, r1, r2, dx, dy, dz);
                            thr = thr * J.albedo;
                            self = h.tri;
                        }
                    }
                }
                if (valid) sum += L;
            }
            if (valid) {
                float4 a = J.accum[pix];
                a.x += sum;
                a.y += sum;
                a.z += sum;
                a.w += (float)J.spp;
                J.accum[pix] = a;
            }
        } else if (valid) {
            const SceneView sv{J.nodes, J.tris, J.hdr};
            int32_t root_ref;
            float root_box[6];
            load_root(J.hdr, root_ref, root_box);
            const uint32_t pix = py * J.width + px;
            const float fw = (float)J.width, fh = (float)J.height;
            float sum = 0.0f;
            for (uint32_t s = 0; s < J.spp; ++s) {
                Pcg32 rng = pcg_seed(sample_seed(J.seed, J.frame + frame_add, s), pix);
                const float jx = 2.0f * pcg_unit(rng) - 1.0f;
                const float jy = 2.0f * pcg_unit(rng) - 1.0f;
                float dx, dy, dz;
                primary_ray(J.cam, px, py, fw, fh, jx, jy, dx, dy, dz);
                float ox = J.cam.pos[0], oy = J.cam.pos[1], oz = J.cam.pos[2];
                if (J.mode == 0) {
                    const TraceOut h = traverse_form<FORM, false>(sv.nodes, sv.tris, root_ref, root_box, ox,
                                                                  oy, oz, dx, dy, dz, 10000000000000000.0f, -1,
                                                                  stk);
                    nprimary += 1;
                    boxes += h.boxes;
                    ntris += h.tris;
                    sum += h.tri < 0 ? J.sky
                                     : direct_term<FORM>(sv, root_ref, root_box, h, ox, oy, oz, dx, dy, dz,
                                                   J.lights, J.nlights, J.ambient, J.albedo, stk,
                                                   boxes, ntris, nshadow, nullptr);
                    continue;
                }
                float thr = 1.0f, L = 0.0f;
                int32_t self = -1;
                for (uint32_t b = 0; b <= J.max_bounces; ++b) {
                    const TraceOut h = traverse_form<FORM, false>(sv.nodes, sv.tris, root_ref, root_box, ox,
                                                                  oy, oz, dx, dy, dz, 10000000000000000.0f, self,
                                                                  stk);
                    if (b == 0) nprimary += 1;
                    else nbounce += 1;
                    boxes += h.boxes;
                    ntris += h.tris;
                    if (h.tri < 0) {
                        L = L + thr * J.sky;
                        break;
                    }
                    if (b == J.max_bounces) break;
                    float nx, ny, nz;
                    flat_normal(sv.tris[h.pos], nx, ny, nz);
                    if (((nx * dx + ny * dy) + nz * dz) > 0.0f) {
                        nx = -nx;
                        ny = -ny;
                        nz = -nz;
                    }
                    ox = ox + h.t * dx;
                    oy = oy + h.t * dy;
                    oz = oz + h.t * dz;
                    const float r1 = pcg_unit(rng), r2 = pcg_unit(rng);
                    cosine_dir(nx, ny, nz, r1, r2, dx, dy, dz);
                    thr = thr * J.albedo;
                    self = h.tri;
                }
                sum += L;
            }
            float4 a = J.accum[pix];
            a.x += sum;
            a.y += sum;
            a.z += sum;
            a.w += (float)J.spp;
            J.accum[pix] = a;
        }
        if (!PERSISTENT) break;
    }
    count_add(stats, 0, boxes);
    count_add(stats, 1, ntris);
    count_add(stats, 2, nshadow);
    count_add(stats, 3, nbounce);
    count_add(stats, 4, nprimary);
}

__global__ void k_reset_counter(uint32_t *c)
{
    if (threadIdx.x == 0) *c = 0u;
}

__global__ void k_bump_counter(uint32_t *c)
{
    if (threadIdx.x == 0) *c += 1u;
}

// ---------------------------------------------------------------------------
// Host launchers.  The tuning picks one of the instantiated (block, LDS stack
// entries) pairs.
// ---------------------------------------------------------------------------
static SceneView view_of(const Batch &b, uint32_t scene)
{
    if (uses_bvh(b, scene)) {
        const Batch::Bvh &v = b.bvh[scene];
        return SceneView{v.trav, v.tris, v.hdr};
    }
    return SceneView{b.trav, b.tris, b.hdr + scene};
}

#define RT_SHAPES(MACRO, F)                              \
    switch (launch_shape(tune, max_depth)) {             \
        case 0: MACRO(64, 16, F); break;                 \
        case 1: MACRO(64, 20, F); break;                 \
        case 2: MACRO(128, 12, F); break;                \
        case 3: MACRO(128, 16, F); break;                \
        case 4: MACRO(128, 20, F); break;                \
        case 5: MACRO(128, 22, F); break;                \
        case 6: MACRO(128, 26, F); break;                \
        default: MACRO(256, 12, F); break;               \
    }
#define RT_TUNED(MACRO)                                  \
    if (tune.traversal == 1) {                           \
        RT_SHAPES(MACRO, 1)                              \
    } else if (tune.traversal == 2) {                    \
        RT_SHAPES(MACRO, 2)                              \
    } else {                                             \
        RT_SHAPES(MACRO, 0)                              \
    }

// Instantiated (block, LDS stack entries): a stack of depth - 1 entries never
// spills, so 20 / 22 / 26 cover heaps of depth 21 / 23 / 27 entirely in LDS,
// while smaller stacks trade spills to private memory for occupancy.
int tune_index(const rt_tuning &t)
{
    if (t.block == 64) return t.lds_stack <= 16 ? 0 : 1;
    if (t.block == 256) return 7;
    if (t.lds_stack <= 12) return 2;
    if (t.lds_stack <= 16) return 3;
    if (t.lds_stack <= 20) return 4;
    if (t.lds_stack <= 22) return 5;
    return 6;
}

// The shape actually launched for heaps up to max_depth: a no-spill stack
// too small for the deepest heap is replaced by the smallest one that fits
// (block 128), or by the spilling stack of the same block size (block 64).
static int launch_shape(const rt_tuning &t, uint32_t max_depth)
{
    static const int kEntries[8] = {16, 20, 12, 16, 20, 22, 26, 12};
    const int need = max_depth > 1 ? (int)max_depth - 1 : 0;
    int i = tune_index(t);
    if (stack_spills(kEntries[i]) || kEntries[i] >= need) return i;
    if (i == 1) return 0;
    if (need <= 22) return 5;
    return 6;
}

hipError_t launch_trace(const Batch &b, uint32_t scene, const rt_ray48 *rays, uint32_t n,
                        rt_hit24 *hits, uint32_t flags, unsigned long long *stats,
                        const rt_tuning &tune, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const SceneView sv = view_of(b, scene);
    const uint32_t max_depth = trav_depth(b, scene);
    const bool any = flags & RT_FLAG_ANYHIT, cnt = flags & RT_FLAG_COUNT;
    const int from_root = (flags & RT_FLAG_START_AT_IROOT) ? 1 : 0;
#define RT_L(B, K, F)                                                                               \
    {                                                                                              \
        const dim3 grid((n + B - 1) / B), block(B);                                                \
        if (any && cnt) hipLaunchKernelGGL((k_trace<true, true, B, K, F>), grid, block, 0, st, sv, rays, n, hits, from_root, stats); \
        else if (any) hipLaunchKernelGGL((k_trace<true, false, B, K, F>), grid, block, 0, st, sv, rays, n, hits, from_root, stats); \
        else if (cnt) hipLaunchKernelGGL((k_trace<false, true, B, K, F>), grid, block, 0, st, sv, rays, n, hits, from_root, stats); \
        else hipLaunchKernelGGL((k_trace<false, false, B, K, F>), grid, block, 0, st, sv, rays, n, hits, from_root, stats); \
    }
    RT_TUNED(RT_L)
#undef RT_L
    return hipGetLastError();
}

// The order slot of a view: the one holding (key, block, tiles), else the
// least recently used one, taken over with its costs cleared and no order
// yet (the view's first frame runs in block order).  Null if clearing failed.
static FrameBuffers::OrderSlot *order_slot(FrameBuffers &fb, uint64_t key, int block, uint32_t tiles,
                                           hipStream_t st)
{
    FrameBuffers::OrderSlot *lru = &fb.slots[0];
    for (FrameBuffers::OrderSlot &o : fb.slots) {
        if (o.used && o.key == key && o.block == (uint32_t)block && o.tiles == tiles) {
            o.tick = ++fb.order_tick;
            return &o;
        }
        if (o.tick < lru->tick) lru = &o;
    }
    lru->used = true;
    lru->valid = false;
    lru->key = key;
    lru->block = (uint32_t)block;
    lru->tiles = tiles;
    lru->tick = ++fb.order_tick;
    return hipMemsetAsync(lru->cost, 0, sizeof(uint32_t) * tiles, st) == hipSuccess ? lru : nullptr;
}

bool camera_ok(const rt_camera &c) { return c.kind == RT_CAMERA_BASIS || c.kind == RT_CAMERA_WORLD; }

rt_camera device_camera(const rt_camera &c)
{
    rt_camera d = c;
    if (c.kind == RT_CAMERA_WORLD) {
        const float *m = c.world;
        const float w = ((m[12] * 0.0f + m[13] * 0.0f) + m[14] * 0.0f) + m[15];
        for (int k = 0; k < 3; ++k) d.pos[k] = (((m[4 * k] * 0.0f + m[4 * k + 1] * 0.0f) + m[4 * k + 2] * 0.0f) + m[4 * k + 3]) / w;
    }
    return d;
}

// A view's key: the scene, the kernel (tag) and the camera's bytes (FNV-1a).
static uint64_t view_key(const Batch &b, uint32_t scene, uint32_t tag, const rt_camera &cam)
{
    uint64_t h = 1469598103934665603ull;
    const unsigned char *p = reinterpret_cast<const unsigned char *>(&cam);
    for (size_t k = 0; k < sizeof(cam); ++k) h = (h ^ p[k]) * 1099511628211ull;
    return ((((uint64_t)(uintptr_t)&b << 16) ^ scene) + tag) ^ h;
}

hipError_t launch_render_direct(const Batch &b, uint32_t scene, const rt_render_desc &d,
                                FrameBuffers &fb, uint32_t flags, unsigned long long *stats,
                                const rt_tuning &tune, hipStream_t st)
{
    const uint64_t key = view_key(b, scene, 1u, d.cam);
    const uint32_t max_depth = trav_depth(b, scene);
    RenderParams rp;
    rp.cam = device_camera(d.cam);
    rp.width = d.width;
    rp.height = d.height;
    rp.nlights = d.nlights;
    rp.ambient = d.ambient;
    rp.albedo = d.albedo;
    rp.env = d.env;
    for (uint32_t i = 0; i < 24; ++i) rp.lights[i] = (i < 3 * d.nlights) ? d.lights[i] : 0.0f;
    const SceneView sv = view_of(b, scene);
    const bool keep = flags & RT_FLAG_KEEP_HITS, cnt = flags & RT_FLAG_COUNT;
    // quads need heap indices in the nodes (the LBVH's), and frames that keep
    // hits or count keep one lane per ray
    const uint32_t quad_tiles = (keep || cnt || uses_bvh(b, scene)) ? 0u : (uint32_t)tune.quad_tiles;
    rp.quad_depth = kQuadDepth;
    // shadow rays on lane groups (the shared-fetch form) need the LBVH's
    // heap indices; counted frames keep one lane per ray
    rp.group = (!cnt && !uses_bvh(b, scene)) ? 1u : 0u;
    uint32_t ntiles = 0;
    FrameBuffers::OrderSlot *os = nullptr;
#define RT_L(B, K, F)                                                                               \
    {                                                                                              \
        constexpr int TW = Tile<B>::W, TH = Tile<B>::H;                                            \
        const uint32_t tiles = ((d.width + TW - 1) / TW) * ((d.height + TH - 1) / TH);             \
        const dim3 grid(tiles), block(B);                                                          \
        ntiles = tiles;                                                                            \
        os = order_slot(fb, key, B, tiles, st);                                                    \
        if (!os) return hipGetLastError();                                                         \
        rp.order = os->valid ? os->order : nullptr;                                                \
        rp.cost = os->cost;                                                                        \
        rp.nsplit = rp.order ? (quad_tiles < tiles ? quad_tiles : tiles) : 0u;                     \
        if (rp.nsplit) hipLaunchKernelGGL((k_render_direct_quad<B, K, F>), dim3(tiles + 3u * rp.nsplit), block, 0, st, sv, rp, fb.rgba, stats); \
        else if (keep && cnt) hipLaunchKernelGGL((k_render_direct<true, true, B, K, F>), grid, block, 0, st, sv, rp, fb.rgba, fb.hits, fb.occ, stats); \
        else if (keep) hipLaunchKernelGGL((k_render_direct<true, false, B, K, F>), grid, block, 0, st, sv, rp, fb.rgba, fb.hits, fb.occ, stats); \
        else if (cnt) hipLaunchKernelGGL((k_render_direct<false, true, B, K, F>), grid, block, 0, st, sv, rp, fb.rgba, fb.hits, fb.occ, stats); \
        else hipLaunchKernelGGL((k_render_direct<false, false, B, K, F>), grid, block, 0, st, sv, rp, fb.rgba, fb.hits, fb.occ, stats); \
    }
    RT_TUNED(RT_L)
#undef RT_L
    hipLaunchKernelGGL(k_tile_order, dim3(tile_order_blocks(ntiles)), dim3(1024), 0, st, os->cost, ntiles,
                       os->order);
    os->valid = true;
    return hipGetLastError();
}

hipError_t launch_render_whitted(const Batch &b, uint32_t scene, const rt_whitted_desc &d,
                                 FrameBuffers &fb, uint32_t flags, unsigned long long *stats,
                                 const rt_tuning &tune, void *queue, uint32_t *qcount, int cu_count,
                                 hipStream_t st)
{
    const uint32_t max_depth = trav_depth(b, scene);
    const SceneDesc &sd = b.desc[scene];
    WhittedParams wp;
    memset(&wp, 0, sizeof(wp));
    wp.cam = device_camera(d.cam);
    wp.width = d.width;
    wp.height = d.height;
    wp.bounces = d.bounces;
    wp.switches = d.switches;
    for (int k = 0; k < 3; ++k) {
        wp.light[k] = d.light[k];
        wp.env[k] = d.env[k];
    }
    wp.env_flag = d.env_flag;
    wp.frame = d.frame;
    wp.seed = d.seed;
    // MaterialData() (Common/Scenes/MaterialData.cpp:17-24)
    wp.defmat.diffuse[0] = wp.defmat.diffuse[1] = wp.defmat.diffuse[2] = 0.5f;
    wp.defmat.diffuse[3] = 0.8f;
    wp.defmat.ambient[3] = 1.0f;
    wp.defmat.specular[0] = wp.defmat.specular[1] = wp.defmat.specular[2] = 0.8f;
    wp.defmat.specular[3] = 1.0f;
    ShadeView sh{nullptr, b.idx + 3 * (size_t)sd.in_prim_off, nullptr, nullptr};
    if (scene < b.shading.size()) {
        const Batch::Shading &S = b.shading[scene];
        if (S.normals) sh.nrm = b.nrm + 3 * (size_t)sd.vert_off;
        if (S.ids) sh.mat = b.mat + sd.in_prim_off;
        sh.mats = S.mats;
    }
    const SceneView sv = view_of(b, scene);
    const int keep = (flags & RT_FLAG_KEEP_HITS) ? 1 : 0, cnt = (flags & RT_FLAG_COUNT) ? 1 : 0;
    const bool one_pass = (d.switches & RT_WHITTED_ONE_PASS) != 0 || d.bounces == 0;
    const bool split = (d.switches & RT_WHITTED_SPLIT) != 0;
    // quads need heap indices in the nodes (the LBVH's); counted frames keep
    // one lane per ray (the same counts as the other forms)
    const bool quads = (d.switches & RT_WHITTED_QUADS) != 0 && !cnt && !uses_bvh(b, scene);
    // the split form's remaining passes with live paths on lane groups: the
    // same heap-index requirement, and uncounted frames only
    const bool groups = !cnt && !uses_bvh(b, scene);
    wp.groups = groups ? 1u : 0u;
    const uint64_t px = (uint64_t)d.width * d.height;
    RayRec *q[2] = {static_cast<RayRec *>(queue), static_cast<RayRec *>(queue) + px};
    if (!one_pass) hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, qcount, d.bounces + 1);
    // the tile order of the one-pass kernel or of the wavefront's first pass
    const uint64_t order_key = view_key(b, scene, one_pass ? 2u : 3u, d.cam);
    uint32_t ntiles = 0;
    FrameBuffers::OrderSlot *os = nullptr;
#define RT_L(B, K, F)                                                                                      \
    {                                                                                                     \
        constexpr int TW = Tile<B>::W, TH = Tile<B>::H;                                                   \
        const uint32_t tiles = ((d.width + TW - 1) / TW) * ((d.height + TH - 1) / TH);                    \
        ntiles = tiles;                                                                                   \
        os = order_slot(fb, order_key, B, tiles, st);                                                     \
        if (!os) return hipGetLastError();                                                                \
        wp.order = os->valid ? os->order : nullptr;                                                       \
        wp.cost = os->cost;                                                                               \
        if (one_pass) {                                                                                   \
            hipLaunchKernelGGL((k_render_whitted<B, K, F>), dim3(tiles), dim3(B), 0, st, sv, sh, wp,      \
                               fb.rgba, fb.hits, fb.occ, keep, cnt, stats);                               \
        } else {                                                                                          \
            hipLaunchKernelGGL((k_whitted_first<B, K, F>), dim3(tiles), dim3(B), 0, st, sv, sh, wp,       \
                               fb.rgba, fb.hits, fb.occ, keep, cnt, q[0], qcount, stats);                 \
            uint32_t blocks = (uint32_t)((px + B - 1) / B);                                               \
            const uint32_t cap = (uint32_t)cu_count * 8u;                                                 \
            if (blocks > cap) blocks = cap;                                                               \
            if (split && quads)                                                                           \
                hipLaunchKernelGGL((k_whitted_rest_quad<B, K, F>), dim3(blocks), dim3(B), 0, st, sv, sh,  \
                                   wp, fb.rgba, q[0], qcount, stats);                                     \
            else if (split && groups)                                                                     \
                hipLaunchKernelGGL((k_whitted_rest_group<B, K, F>), dim3(blocks), dim3(B), 0, st, sv, sh, \
                                   wp, fb.rgba, q[0], qcount, stats);                                     \
            else if (split)                                                                               \
                hipLaunchKernelGGL((k_whitted_rest<B, K, F>), dim3(blocks), dim3(B), 0, st, sv, sh, wp,   \
                                   fb.rgba, q[0], qcount, cnt, stats);                                    \
            else                                                                                          \
                for (uint32_t p = 1; p <= d.bounces; ++p)                                                 \
                    hipLaunchKernelGGL((k_whitted_next<B, K, F>), dim3(blocks), dim3(B), 0, st, sv, sh,   \
                                       wp, fb.rgba, q[(p - 1) & 1], qcount + (p - 1), q[p & 1],           \
                                       qcount + p, p == d.bounces ? 1 : 0, cnt, stats);                   \
        }                                                                                                 \
    }
    RT_TUNED(RT_L)
#undef RT_L
    hipLaunchKernelGGL(k_tile_order, dim3(tile_order_blocks(ntiles)), dim3(1024), 0, st, os->cost, ntiles,
                       os->order);
    os->valid = true;
    return hipGetLastError();
}

size_t whitted_queue_bytes(uint64_t pixels) { return 2 * sizeof(RayRec) * (size_t)pixels; }

uint32_t job_tile_pixels(const rt_tuning &t) { return (uint32_t)t.block; }

void fill_render_job(void *dst, const Batch &b, uint32_t scene, const rt_job_desc &d,
                     const rt_tuning &t, uint32_t tile_off, uint32_t *tiles, const uint32_t *order)
{
    RenderJob *j = static_cast<RenderJob *>(dst);
    memset(j, 0, sizeof(*j));
    const uint32_t TW = 16, TH = (uint32_t)t.block / 16;
    const SceneView sv = view_of(b, scene);
    j->nodes = sv.nodes;
    j->tris = sv.tris;
    j->hdr = sv.hdr;
    j->accum = reinterpret_cast<float4 *>(d.accum);
    j->order = order;
    j->cam = device_camera(d.cam);
    j->width = d.width;
    j->height = d.height;
    j->tiles_x = (d.width + TW - 1) / TW;
    j->tile_off = tile_off;
    j->mode = d.mode == RT_MODE_GI ? 1u : 0u;
    j->lbvh = uses_bvh(b, scene) ? 0u : 1u;
    const uint32_t all = j->tiles_x * ((d.height + TH - 1) / TH);
    j->local_base = d.tile_count ? (d.tile_begin < all ? d.tile_begin : all) : 0u;
    j->spp = d.spp;
    j->max_bounces = d.max_bounces;
    j->nlights = d.nlights;
    j->albedo = d.albedo;
    j->sky = d.sky;
    j->ambient = d.ambient;
    j->seed = d.seed;
    j->frame = d.frame;
    for (uint32_t i = 0; i < 24; ++i) j->lights[i] = (i < 3 * d.nlights) ? d.lights[i] : 0.0f;
    *tiles = d.tile_count ? (d.tile_count < all - j->local_base ? d.tile_count : all - j->local_base) : all;
}

static_assert(sizeof(RenderJob) <= kRenderJobBytesMax, "RenderJob outgrew the no-allocation job table");

void centre_tile_order(uint32_t width, uint32_t height, uint32_t tile_h, std::vector<uint32_t> &order)
{
    const uint32_t tx = (width + 15) / 16, ty = (height + tile_h - 1) / tile_h;
    std::vector<std::pair<uint64_t, uint32_t>> key(tx * ty);
    for (uint32_t t = 0; t < tx * ty; ++t) {
        // doubled pixel coordinates keep the tile and image centres integral
        const int64_t cx = 2 * (int64_t)((t % tx) * 16) + 16 - (int64_t)width;
        const int64_t cy = 2 * (int64_t)((t / tx) * tile_h) + tile_h - (int64_t)height;
        key[t] = {(uint64_t)(cx * cx + cy * cy), t};
    }
    std::sort(key.begin(), key.end());
    order.resize(key.size());
    for (size_t k = 0; k < key.size(); ++k) order[k] = key[k].second;
}
size_t render_job_bytes() { return sizeof(RenderJob); }

hipError_t launch_render_jobs(const void *d_jobs, uint32_t njobs, uint32_t total_tiles,
                              uint32_t first_tile, uint32_t tile_count, bool persistent,
                              uint32_t *tile_counter, const uint32_t *frame_counter,
                              unsigned long long *stats, const rt_tuning &tune, uint32_t max_depth,
                              int cu_count, hipStream_t st, bool count)
{
    if (total_tiles == 0 || njobs == 0 || (!persistent && tile_count == 0)) return hipSuccess;
    const RenderJob *jobs = static_cast<const RenderJob *>(d_jobs);
#define RT_L(B, K, F)                                                                                \
    {                                                                                               \
        if (persistent) {                                                                           \
            hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, tile_counter);            \
            uint32_t blocks = (uint32_t)cu_count * (uint32_t)(tune.blocks_per_cu > 0 ? tune.blocks_per_cu : 4); \
            if (blocks > total_tiles) blocks = total_tiles;                                         \
            hipLaunchKernelGGL((k_render_jobs<B, K, true, F>), dim3(blocks), dim3(B), 0, st, jobs, njobs, \
                               total_tiles, 0u, tile_counter, frame_counter, stats, count ? 1u : 0u); \
        } else {                                                                                    \
            hipLaunchKernelGGL((k_render_jobs<B, K, false, F>), dim3(tile_count), dim3(B), 0, st, jobs, \
                               njobs, first_tile + tile_count, first_tile, tile_counter,            \
                               frame_counter, stats, count ? 1u : 0u);                              \
        }                                                                                           \
    }
    RT_TUNED(RT_L)
#undef RT_L
    return hipGetLastError();
}

hipError_t launch_bump_counter(uint32_t *c, hipStream_t st)
{
    hipLaunchKernelGGL(k_bump_counter, dim3(1), dim3(64), 0, st, c);
    return hipGetLastError();
}

hipError_t launch_zero_words(uint32_t *p, uint32_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace rt
