/*
 * CPU oracle for the LBVH build + trace path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker / the CPU baseline.
 *
 * It restates, scalar and in source order, the reference's own CPU code and
 * HLSL kernels for this path (citations are into elfrank/raytracer-gpupro4):
 *   - orc_prim_bounds / orc_morton_codes : AccelerationStructures/LBVH/LBVH.cpp:495-553
 *                                          (same arithmetic as LBVH.hlsl:26-114)
 *   - orc_lbvh_build                     : LBVH.cpp:190-314 (build_recursive, in-place
 *                                          MSD partition) + LBVH.cpp:577-629
 *   - orc_sort_pairs                     : the (code, id) order of cmpMorton, LBVH.cpp:21-38
 *   - orc_trace                          : LBVH.hlsl:363-440 (stackless trail traversal)
 *                                          + Shaders/Core/Intersection.hlsl:18-78
 *   - orc_trace_bvh                      : BVH.hlsl:124-217 (stack traversal of the SAH BVH)
 *   - orc_raygen_pinhole                 : pixel-centre rule of CSPrimaryRays.hlsl:37-40
 *   - orc_raygen_pinhole_jitter,         : jittered primary rays (CSPrimaryRays.hlsl:25-36)
 *     orc_pcg_jitter                       with PCG32 streams (pcg32.h)
 *   - orc_raygen_world                   : the camera-matrix rule of CSPrimaryRays.hlsl:33-56
 *   - orc_render_direct                  : shadow-ray construction + direct term of
 *                                          CSComputeColorRT.hlsl:54-177
 *   - orc_render_whitted                 : the ray-tracing frame, RayTracerCS.cpp:286-295 +
 *                                          CSComputeColorRT.hlsl:18-262 (whitted_oracle.c)
 *
 * Floating point: every float expression is evaluated in IEEE single
 * precision, in the order written, with no contraction into FMA (the build
 * recipe passes -ffp-contract=off and never -ffast-math).
 *
 * How each function is pinned is listed in tests/test_oracle.py.
 */
#ifndef LBVH_ORACLE_H
#define LBVH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* LBVHNode (LBVH.h:25-28) == HLSL Node (Structures.hlsl:18-26), 32 bytes. */
typedef struct {
    float mn[3];
    float mx[3];
    int32_t prim_count; /* (count << 3) | leaf(4) | axis(0..2) */
    int32_t prim_pos;
} orc_node;

/* HLSL Ray (Structures.hlsl:28-35), 48 bytes. */
typedef struct {
    float o[3];
    float d[3];
    float refl[3];
    float tmin, tmax;
    int32_t tri; /* triangle to skip (self id); -1 = none */
} orc_ray;

/* HLSL Intersection (Structures.hlsl:44-50), 24 bytes. */
typedef struct {
    int32_t tri;
    float u, v, t;
    int32_t visited; /* +1 per box test, +65536 per triangle test (LBVH.hlsl:389,404) */
    int32_t root;    /* heap index of the leaf that holds the hit; 1 when no hit */
} orc_hit;

/* Marker written into heap slots the builder never reaches. */
#define ORC_INVALID (-2147483647)

/* Per-triangle AABB, 3x centroid and scene box (LBVH.cpp:495-516). */
void orc_prim_bounds(const float *xyz, const uint32_t *idx, uint32_t nt,
                     float *bmin, float *bmax, float *cent, float scene[6]);

/* 30-bit Morton codes of the 3x centroids (LBVH.cpp:520-553). */
void orc_morton_codes(const float *cent, uint32_t nt, const float scene[6],
                      uint32_t *codes);

/* Closed-form interleave of Part1By2 (MortonCode.hlsl:15-27) applied to
 * q & 1023; used only as an independent cross-check of the loop form. */
uint32_t orc_morton_closed_form(uint32_t qx, uint32_t qy, uint32_t qz);

/* build_recursive(0, nt, 1<<30, 1, 0) over (codes, ids), permuting them in
 * place exactly as the reference's in-place partition does.  nodes must hold
 * 1 << depth entries; unreachable slots get ORC_INVALID in both int fields and
 * zero boxes.  Returns the number of reachable nodes. */
int orc_lbvh_build(uint32_t *codes, int32_t *ids, uint32_t nt, uint32_t depth,
                   const float *bmin, const float *bmax, orc_node *nodes);

/* Sort (code, id) pairs ascending by code, then id (cmpMorton order). */
void orc_sort_pairs(uint32_t *codes, int32_t *ids, uint32_t n);

/* LBVH_IntersectP for each ray.  start[i] is the heap node to start from
 * (NULL = root 1).  anyhit != 0 stops at the first accepted triangle (same
 * visit order).  nthreads > 1 splits rays statically over pthreads. */
void orc_trace(const orc_node *nodes, const int32_t *ids, const float *xyz,
               const uint32_t *idx, const orc_ray *rays, const int32_t *start,
               uint32_t n, int anyhit, orc_hit *hits, int nthreads);

/* BVH_IntersectP (BVH.hlsl:124-217) over a flattened SAH BVH in the
 * reference's LinearBVHNode layout (orc_node: prim_count > 0 is a leaf of
 * prim_count primitives from prim_pos of ordered; otherwise -axis, first
 * child next, second child at prim_pos).  ref_leaf_loop = 1 runs the
 * reference's leaf loop, prim_count + 1 tests all of the leaf's first
 * primitive (BVH.hlsl:165-169); 0 tests each primitive once.  hit.root is
 * the leaf's node index, 0 without a hit. */
void orc_trace_bvh(const orc_node *nodes, const int32_t *ordered, const float *xyz, const uint32_t *idx,
                   const orc_ray *rays, uint32_t n, int anyhit, int ref_leaf_loop, orc_hit *hits,
                   int nthreads);
/* Closest hit by testing every triangle in index order (no structure). */
void orc_trace_brute(const float *xyz, const uint32_t *idx, uint32_t nt, const orc_ray *rays, uint32_t n,
                     orc_hit *hits);
/* Pinhole camera.  cam = {pos[3], fwd[3], right[3], up[3], tan_x, tan_y}
 * (unit basis vectors and the half-angle tangents, already rounded to float).
 * Pixel (px, py): sx = ((2px+1) - W)/W * tan_x, sy = -((2py+1) - H)/H * tan_y,
 * d = normalize((fwd + sx*right) + sy*up). */
void orc_raygen_pinhole(const float cam[14], uint32_t w, uint32_t h, orc_ray *rays);

/* The same with a sub-pixel offset per pixel (the jittered primary rays of
 * CSPrimaryRays.hlsl:25-36): x = (((2px+1) + jx) - W)/W,
 * y = -((((2py+1) + jy) - H)/H), jxy = {jx, jy} per pixel (NULL: 0). */
void orc_raygen_pinhole_jitter(const float cam[14], uint32_t w, uint32_t h, const float *jxy,
                               orc_ray *rays);

/* n outputs of the PCG32 stream pcg32_srandom_r(initstate, initseq) (pcg32.h). */
void orc_pcg32_draws(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t *out);

/* The render jobs' pixel jitter: per pixel p, the PCG32 stream
 * (initstate = seed + 0x9E3779B97F4A7C15 * (frame * 65536 + sample + 1),
 * initseq = p) draws u0, u1 = (next >> 8) / 2^24 and jx = 2 u0 - 1,
 * jy = 2 u1 - 1 (the device's replacement of the random texture of
 * CSPrimaryRays.hlsl:26-32). */
void orc_pcg_jitter(uint64_t seed, uint32_t frame, uint32_t sample, uint32_t w, uint32_t h, float *jxy);

/* The reference's camera-matrix rule (CSPrimaryRays.hlsl:39-56, GI jitter
 * :33-37): world = cbCamera.g_mfWorld as the host's Matrix4 holds it
 * (Camera::GetInverseMatrix, row-major, translation in column 3;
 * RayTracerCS/ConstantBuffers.h:23-31), inverse = 1/H, x = (((2px + jx) + 1)
 * - W) * inverse, y = -((((2py + jy) + 1) - H) * inverse), z = 2,
 * scale = 0.008 / |0.035 z|, origin = (M (0,0,0,1)).xyz / w,
 * d = normalize((M (scale x, scale y, scale z, 1)).xyz / w - origin).  Each
 * matrix row is summed left to right.  For a square image (the reference's
 * only case) inverse = 1/N; wider images keep the pixel pitch. */
void orc_raygen_world(const float world[16], uint32_t w, uint32_t h, const float *jxy, orc_ray *rays);

/* Direct lighting of one primary-hit buffer: per pixel, the flat normal,
 * and for each light a shadow ray {o = hit, d = -Ldir, maxT = |hit - L|,
 * self = hit triangle} traced with closest-hit from the primary hit's leaf
 * (reference semantics, CSComputeColorRT.hlsl:126-160).  shade = {ambient,
 * albedo, env} with a white material of alpha 1 (no specular term).
 * rgba: 4 floats per pixel.  occluded (optional): one byte per pixel per
 * light: 0 = lit, 1 = occluded, 2 = no shadow ray cast. */
void orc_render_direct(const orc_node *nodes, const int32_t *ids,
                       const float *xyz, const uint32_t *idx,
                       const orc_ray *rays, const orc_hit *hits, uint32_t n,
                       const float *lights, uint32_t nlights,
                       const float shade[3], float *rgba, uint8_t *occluded,
                       int nthreads);

/* FP64 diffuse path tracer (gi_oracle.c): sums over samples
 * first_sample .. first_sample + spp - 1 per pixel into out[w*h]; returns the
 * number of rays traced. */
uint64_t orc_render_gi(const orc_node *nodes, const int32_t *ids, const float *xyz,
                       const uint32_t *idx, const float cam[14], uint32_t w, uint32_t h,
                       uint32_t spp, uint32_t max_bounces, double albedo, double sky,
                       uint64_t seed, uint32_t frame, uint32_t first_sample, double *out,
                       int nthreads);

/* Material (Structures.hlsl:69-78), 64 bytes. */
typedef struct {
    float diffuse[4];
    float ambient[4];
    float specular[4];
    float shininess, alpha;
    uint32_t id, pad;
} orc_material;

/* Whitted frame parameters (whitted_oracle.c).  switches: 1 shadows,
 * 2 Phong normals, 4 gloss; env_flag -1 or 0; ref_order 1 starts each ray at
 * the previous hit's leaf as the reference does, 0 at the root. */
typedef struct {
    float light[3];
    uint32_t bounces, switches;
    int32_t env_flag;
    float env[3];
    uint32_t frame;
    uint64_t seed;
    int32_t ref_order;
} orc_whitted_params;

/* Whitted frame of w x h pixels into rgba (4 floats per pixel); normals,
 * mat_ids and mats may be NULL (flat only / material 0 / the default
 * MaterialData).  hits0 / occ0 (optional) receive the first pass's hit and
 * shadow result (0 lit, 1 occluded, 2 none).  Returns the rays traced. */
uint64_t orc_render_whitted(const orc_node *nodes, const int32_t *ids, const float *xyz,
                            const uint32_t *idx, const float *normals, const uint32_t *mat_ids,
                            const orc_material *mats, const float cam[14], uint32_t w, uint32_t h,
                            const orc_whitted_params *p, float *rgba, orc_hit *hits0, uint8_t *occ0,
                            int nthreads);

#ifdef __cplusplus
}
#endif

#endif
