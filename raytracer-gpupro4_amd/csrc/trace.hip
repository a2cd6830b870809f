// Ray traversal on gfx950: closest-hit and any-hit over the LBVH, primary-ray
// generation and direct shading with shadow rays.
//
// The visit order is the reference's stackless trail traversal
// (LBVH_IntersectP, LBVH.hlsl:363-440): depth first, at an internal node the
// child on side dirIsNeg[axis] first, a box culled when T0 > best t or T1 < 0,
// leaf triangles tested first to last, a hit taken only with t < best t.
// Here a node holds both child boxes, so both are tested when the node is
// entered; the second child is kept on a per-lane LDS stack with its T0 and
// culled on pop against the then-current best t.  That is the same decision,
// at the same point of the sequence, as the trail algorithm makes when it
// reaches the sibling, so the hits, the tie-breaking and the visit counters
// (+1 per box, +65536 per triangle) equal the reference's.
// Float arithmetic is IEEE single with no contraction (-ffp-contract=off):
// IntersectBox and getIntersection (Intersection.hlsl:18-64) in source order.
#include "rt_internal.h"

namespace rt {

constexpr int kTraceBlock = 128;
constexpr int kStack = 26;  // >= depth - 1 for depth <= 27

struct TraceOut {
    int32_t tri;
    float u, v, t;
    uint32_t boxes, tris;
    int32_t leaf_heap;
    int32_t pos;  // sorted slot of the hit triangle, -1 when none
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

// Traverses one ray.  stk points at this lane's first LDS stack slot; slot j
// is stk[j * kTraceBlock].
template <bool ANYHIT>
__device__ __forceinline__ TraceOut traverse(const TravNode *__restrict__ nodes,
                                             const TravTri *__restrict__ tris, int32_t root_ref,
                                             const float *root_box, float ox, float oy, float oz,
                                             float dx, float dy, float dz, float tmax, int32_t self,
                                             uint2 *stk)
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
    const int negx = ix < 0.0f, negy = iy < 0.0f, negz = iz < 0.0f;
    float t0, t1;
    box_test(root_box, ox, oy, oz, ix, iy, iz, t0, t1);
    r.boxes = 1;
    if (t0 > r.t || t1 < 0.0f) return r;
    int32_t cur = root_ref;
    int sp = 0;
    while (true) {
        if (cur >= 0) {
            const TravNode *n = &nodes[cur];
            const float4 q0 = *reinterpret_cast<const float4 *>(&n->box[0]);
            const float4 q1 = *reinterpret_cast<const float4 *>(&n->box[4]);
            const float4 q2 = *reinterpret_cast<const float4 *>(&n->box[8]);
            const int4 q3 = *reinterpret_cast<const int4 *>(&n->ref[0]);
            const float ba[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
            const float bb[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
            float a0, a1, b0, b1;
            box_test(ba, ox, oy, oz, ix, iy, iz, a0, a1);
            box_test(bb, ox, oy, oz, ix, iy, iz, b0, b1);
            r.boxes += 2;
            const int axis = q3.z;
            const int first_is_b = axis == 0 ? negx : (axis == 1 ? negy : negz);
            const float f0 = first_is_b ? b0 : a0, f1 = first_is_b ? b1 : a1;
            const float s0 = first_is_b ? a0 : b0, s1 = first_is_b ? a1 : b1;
            const int32_t fref = first_is_b ? q3.y : q3.x;
            const int32_t sref = first_is_b ? q3.x : q3.y;
            const bool f_ok = !(f0 > r.t || f1 < 0.0f);
            const bool s_live = !(s1 < 0.0f);
            if (f_ok) {
                if (s_live) {
                    stk[sp * kTraceBlock] = make_uint2((uint32_t)sref, __float_as_uint(s0));
                    ++sp;
                }
                cur = fref;
                continue;
            }
            if (s_live && !(s0 > r.t)) {
                cur = sref;
                continue;
            }
        } else {
            const uint32_t pos = (uint32_t)cur & ~kLeafBit;
            const TravTri *tp = &tris[pos];
            int32_t cnt = 1;
            for (int32_t j = 0; j < cnt; ++j) {
                const float4 a = tp[j].a;
                const float4 e1 = tp[j].e1;
                const float4 e2 = tp[j].e2;
                if (j == 0) cnt = __float_as_int(e1.w);
                r.tris += 1;
                const int32_t id = __float_as_int(a.w);
                // getIntersection (Intersection.hlsl:47-64)
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
                if (id != self && u >= 0.0f && v >= 0.0f && (u + v) <= 1.0f && t >= 0.0f &&
                    t < r.t) {
                    r.tri = id;
                    r.u = u;
                    r.v = v;
                    r.t = t;
                    r.leaf_heap = __float_as_int(tp[0].e2.w);
                    r.pos = (int32_t)pos + j;
                    if (ANYHIT) return r;
                }
            }
        }
        // pop the next live entry
        bool found = false;
        while (sp > 0) {
            --sp;
            const uint2 e = stk[sp * kTraceBlock];
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

// rt_trace: one lane per ray.
template <bool ANYHIT, bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_trace(SceneView sv, const rt_ray48 *__restrict__ rays,
                                                       uint32_t n, rt_hit24 *__restrict__ hits,
                                                       unsigned long long *stats)
{
    __shared__ uint2 stack[kStack * kTraceBlock];
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
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
        rt_hit24 h;
        if (self <= -2) {
            h.tri = -2;
            h.u = -1.0f;
            h.v = -1.0f;
            h.t = r2.z;
            h.visited = 0;
            h.root = 1;
        } else {
            const TraceOut o = traverse<ANYHIT>(sv.nodes, sv.tris, root_ref, root_box, r0.x, r0.y,
                                                r0.z, r0.w, r1.x, r1.y, r2.z, self,
                                                &stack[threadIdx.x]);
            h.tri = o.tri;
            h.u = o.u;
            h.v = o.v;
            h.t = o.t;
            h.visited = (int32_t)(o.boxes + 65536u * o.tris);
            h.root = o.leaf_heap;
            boxes = o.boxes;
            ntris = o.tris;
        }
        int2 *hp = reinterpret_cast<int2 *>(&hits[i]);
        hp[0] = make_int2(h.tri, __float_as_int(h.u));
        hp[1] = make_int2(__float_as_int(h.v), __float_as_int(h.t));
        hp[2] = make_int2(h.visited, h.root);
    }
    if (COUNT) {
        count_add(stats, 0, boxes);
        count_add(stats, 1, ntris);
    }
}

// Pinhole primary ray: pixel-centre rule of CSPrimaryRays.hlsl:37-40 with an
// explicit camera basis, maxT 1e16 and self id -1 as at :49-53.
__device__ __forceinline__ void primary_ray(const rt_camera &c, uint32_t px, uint32_t py, float fw,
                                            float fh, float &dx, float &dy, float &dz)
{
    const float x = ((float)(2u * px + 1u) - fw) / fw;
    const float y = -(((float)(2u * py + 1u) - fh) / fh);
    const float sx = x * c.tan_x, sy = y * c.tan_y;
    const float ax = c.fwd[0] + sx * c.right[0];
    const float ay = c.fwd[1] + sx * c.right[1];
    const float az = c.fwd[2] + sx * c.right[2];
    const float ddx = ax + sy * c.up[0];
    const float ddy = ay + sy * c.up[1];
    const float ddz = az + sy * c.up[2];
    const float len = sqrtf((ddx * ddx + ddy * ddy) + ddz * ddz);
    dx = ddx / len;
    dy = ddy / len;
    dz = ddz / len;
}

struct RenderParams {
    rt_camera cam;
    uint32_t width, height, nlights;
    float ambient, albedo, env;
    float lights[24];  // up to 8 point lights
};

constexpr int kTileW = 16, kTileH = 8;  // 128 pixels per block

// Primary + shadow rays + direct term (CSComputeColorRT.hlsl:54-198 with a
// white material of alpha 1, flat normals, no textures).  Shadow rays are
// any-hit from the root: occlusion is the same predicate the reference's
// closest-hit shadow ray from the hit leaf evaluates (any accepted triangle
// with t < |hit - L|), without the search for the closest one.
template <bool KEEP, bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_render_direct(SceneView sv, RenderParams rp,
                                                               float4 *__restrict__ rgba,
                                                               rt_hit24 *__restrict__ hits,
                                                               uint8_t *__restrict__ occ,
                                                               unsigned long long *stats)
{
    __shared__ uint2 stack[kStack * kTraceBlock];
    const uint32_t tiles_x = (rp.width + kTileW - 1) / kTileW;
    const uint32_t tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const uint32_t px = tx * kTileW + (threadIdx.x % kTileW);
    const uint32_t py = ty * kTileH + (threadIdx.x / kTileW);
    const bool valid = px < rp.width && py < rp.height;
    int32_t root_ref;
    float root_box[6];
    load_root(sv.hdr, root_ref, root_box);
    uint2 *stk = &stack[threadIdx.x];
    uint32_t boxes = 0, ntris = 0, nshadow = 0;
    if (valid) {
        const uint32_t pix = py * rp.width + px;
        float dx, dy, dz;
        primary_ray(rp.cam, px, py, (float)rp.width, (float)rp.height, dx, dy, dz);
        const float ox = rp.cam.pos[0], oy = rp.cam.pos[1], oz = rp.cam.pos[2];
        const TraceOut h = traverse<false>(sv.nodes, sv.tris, root_ref, root_box, ox, oy, oz, dx,
                                           dy, dz, 10000000000000000.0f, -1, stk);
        boxes += h.boxes;
        ntris += h.tris;
        if (KEEP) {
            int2 *hp = reinterpret_cast<int2 *>(&hits[pix]);
            hp[0] = make_int2(h.tri, __float_as_int(h.u));
            hp[1] = make_int2(__float_as_int(h.v), __float_as_int(h.t));
            hp[2] = make_int2((int32_t)(h.boxes + 65536u * h.tris), h.leaf_heap);
        }
        float color;
        if (h.tri < 0) {
            color = rp.env;
            if (KEEP)
                for (uint32_t l = 0; l < rp.nlights; ++l) occ[(size_t)pix * rp.nlights + l] = 2;
        } else {
            const TravTri &T = sv.tris[h.pos];
            const float4 e1 = T.e1, e2 = T.e2;
            // NorT = normalize(cross(Edge1, Edge2)) (CSComputeColorRT.hlsl:61-77)
            float nx = (e1.y * e2.z) - (e1.z * e2.y);
            float ny = (e1.z * e2.x) - (e1.x * e2.z);
            float nz = (e1.x * e2.y) - (e1.y * e2.x);
            const float nl = sqrtf((nx * nx + ny * ny) + nz * nz);
            nx = nx / nl;
            ny = ny / nl;
            nz = nz / nl;
            float mx = nx, my = ny, mz = nz;  // normal facing the ray (:125-131)
            if (((mx * dx + my * dy) + mz * dz) > 0.0f) {
                mx = -mx;
                my = -my;
                mz = -mz;
            }
            const float hx = ox + h.t * dx, hy = oy + h.t * dy, hz = oz + h.t * dz;
            float diffuse = 0.0f;
            for (uint32_t l = 0; l < rp.nlights; ++l) {
                float lx = hx - rp.lights[3 * l + 0];
                float ly = hy - rp.lights[3 * l + 1];
                float lz = hz - rp.lights[3 * l + 2];
                const float dist = sqrtf((lx * lx + ly * ly) + lz * lz);
                lx = lx / dist;
                ly = ly / dist;
                lz = lz / dist;
                const float ls = (nx * lx + ny * ly) + nz * lz;
                const float rs = (nx * dx + ny * dy) + nz * dz;
                uint8_t o = 2;
                if ((ls * rs) > 0.0f) {
                    const TraceOut s = traverse<true>(sv.nodes, sv.tris, root_ref, root_box, hx, hy,
                                                      hz, -lx, -ly, -lz, dist, h.tri, stk);
                    boxes += s.boxes;
                    ntris += s.tris;
                    nshadow += 1;
                    o = s.tri < 0 ? 0 : 1;
                    if (s.tri < 0) {
                        const float dd = -((mx * lx + my * ly) + mz * lz);
                        diffuse = diffuse + (dd > 0.0f ? dd : 0.0f);
                    }
                }
                if (KEEP) occ[(size_t)pix * rp.nlights + l] = o;
            }
            color = (rp.ambient + diffuse) * rp.albedo;
        }
        rgba[pix] = make_float4(color, color, color, 1.0f);
    }
    if (COUNT) {
        count_add(stats, 0, boxes);
        count_add(stats, 1, ntris);
    }
    count_add(stats, 2, nshadow);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static SceneView view_of(const Batch &b, uint32_t scene)
{
    return SceneView{b.trav, b.tris, b.hdr + scene};
}

hipError_t launch_trace(const Batch &b, uint32_t scene, const rt_ray48 *rays, uint32_t n,
                        rt_hit24 *hits, uint32_t flags, unsigned long long *stats, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const dim3 grid((n + kTraceBlock - 1) / kTraceBlock), block(kTraceBlock);
    const SceneView sv = view_of(b, scene);
    const bool any = flags & RT_FLAG_ANYHIT, cnt = flags & RT_FLAG_COUNT;
    if (any && cnt) hipLaunchKernelGGL((k_trace<true, true>), grid, block, 0, st, sv, rays, n, hits, stats);
    else if (any) hipLaunchKernelGGL((k_trace<true, false>), grid, block, 0, st, sv, rays, n, hits, stats);
    else if (cnt) hipLaunchKernelGGL((k_trace<false, true>), grid, block, 0, st, sv, rays, n, hits, stats);
    else hipLaunchKernelGGL((k_trace<false, false>), grid, block, 0, st, sv, rays, n, hits, stats);
    return hipGetLastError();
}

hipError_t launch_render_direct(const Batch &b, uint32_t scene, const rt_render_desc &d,
                                FrameBuffers &fb, uint32_t flags, unsigned long long *stats,
                                hipStream_t st)
{
    RenderParams rp;
    rp.cam = d.cam;
    rp.width = d.width;
    rp.height = d.height;
    rp.nlights = d.nlights;
    rp.ambient = d.ambient;
    rp.albedo = d.albedo;
    rp.env = d.env;
    for (uint32_t i = 0; i < 24; ++i) rp.lights[i] = (i < 3 * d.nlights) ? d.lights[i] : 0.0f;
    const uint32_t tiles = ((d.width + kTileW - 1) / kTileW) * ((d.height + kTileH - 1) / kTileH);
    if (tiles == 0) return hipSuccess;
    const SceneView sv = view_of(b, scene);
    const bool keep = flags & RT_FLAG_KEEP_HITS, cnt = flags & RT_FLAG_COUNT;
    const dim3 grid(tiles), block(kTraceBlock);
    if (keep && cnt)
        hipLaunchKernelGGL((k_render_direct<true, true>), grid, block, 0, st, sv, rp, fb.rgba, fb.hits, fb.occ, stats);
    else if (keep)
        hipLaunchKernelGGL((k_render_direct<true, false>), grid, block, 0, st, sv, rp, fb.rgba, fb.hits, fb.occ, stats);
    else if (cnt)
        hipLaunchKernelGGL((k_render_direct<false, true>), grid, block, 0, st, sv, rp, fb.rgba, fb.hits, fb.occ, stats);
    else
        hipLaunchKernelGGL((k_render_direct<false, false>), grid, block, 0, st, sv, rp, fb.rgba, fb.hits, fb.occ, stats);
    return hipGetLastError();
}

}  // namespace rt
