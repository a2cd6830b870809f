/*
 * CPU oracle for the LBVH build + trace path -- TEST INFRASTRUCTURE ONLY.
 * See lbvh_oracle.h for what each function restates and where.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no -ffast-math).
 */
#include "lbvh_oracle.h"
#include "pcg32.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* The reference's scalar min/max (std::min / std::max semantics): exact for
 * the finite inputs used here. */
static inline float rmin(float a, float b) { return (b < a) ? b : a; }
static inline float rmax(float a, float b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------ */
/* LBVH.cpp:495-516: per-primitive box, 3x centroid and the scene box.       */
/* ------------------------------------------------------------------------ */
void orc_prim_bounds(const float *xyz, const uint32_t *idx, uint32_t nt,
                     float *bmin, float *bmax, float *cent, float scene[6])
{
    float smin[3] = {0, 0, 0}, smax[3] = {0, 0, 0};
    if (nt > 0) {
        const float *p0 = xyz + 3 * (size_t)idx[0];
        for (int k = 0; k < 3; ++k) smin[k] = smax[k] = p0[k];
    }
    for (uint32_t i = 0; i < nt; ++i) {
        const float *a = xyz + 3 * (size_t)idx[3 * (size_t)i + 0];
        const float *b = xyz + 3 * (size_t)idx[3 * (size_t)i + 1];
        const float *c = xyz + 3 * (size_t)idx[3 * (size_t)i + 2];
        for (int k = 0; k < 3; ++k) {
            float lo = rmin(a[k], rmin(b[k], c[k]));
            float hi = rmax(a[k], rmax(b[k], c[k]));
            bmin[3 * (size_t)i + k] = lo;
            bmax[3 * (size_t)i + k] = hi;
            /* gCentroid = A + B + C, left to right (LBVH.cpp:512) */
            float s = a[k] + b[k];
            cent[3 * (size_t)i + k] = s + c[k];
            smin[k] = rmin(smin[k], lo);
            smax[k] = rmax(smax[k], hi);
        }
    }
    for (int k = 0; k < 3; ++k) {
        scene[k] = smin[k];
        scene[3 + k] = smax[k];
    }
}

/* ------------------------------------------------------------------------ */
/* LBVH.cpp:520-553: quantise (c - 3*min) * 1024/(3*max - 3*min) to int and  */
/* interleave 10 bits per axis, x -> bit 3k+2, y -> 3k+1, z -> 3k.           */
/* ------------------------------------------------------------------------ */
static inline uint32_t quantise(float c, float m, float f)
{
    float s = (c - m) * f;
    /* A flat axis gives f = inf and s = NaN; the reference's int() of NaN is
     * INT_MIN on x86 and 0 in HLSL -- both have zero low 10 bits. */
    if (!(s == s)) return 0u;
    return (uint32_t)(int)s;
}

void orc_morton_codes(const float *cent, uint32_t nt, const float scene[6],
                      uint32_t *codes)
{
    float m[3], f[3];
    for (int k = 0; k < 3; ++k) {
        float hi = 3.0f * scene[3 + k];
        float lo = 3.0f * scene[k];
        float ext = hi - lo;
        m[k] = lo;
        f[k] = 1024.0f / ext;
    }
    for (uint32_t i = 0; i < nt; ++i) {
        uint32_t inx = quantise(cent[3 * (size_t)i + 0], m[0], f[0]);
        uint32_t iny = quantise(cent[3 * (size_t)i + 1], m[1], f[1]);
        uint32_t inz = quantise(cent[3 * (size_t)i + 2], m[2], f[2]);
        uint32_t code = (inz & 1u) | ((iny & 1u) << 1) | ((inx & 1u) << 2);
        uint32_t shift3 = 2, shift = 2;
        for (int j = 1; j < 10; ++j) {
            code |= (inz & shift) << (shift3++);
            code |= (iny & shift) << (shift3++);
            code |= (inx & shift) << (shift3);
            shift <<= 1;
        }
        codes[i] = code;
    }
}

static uint32_t part1by2(uint32_t n)
{
    n &= 1023u;
    n = (n ^ (n << 16)) & 0xff0000ffu;
    n = (n ^ (n << 8)) & 0x0300f00fu;
    n = (n ^ (n << 4)) & 0x030c30c3u;
    n = (n ^ (n << 2)) & 0x09249249u;
    return n;
}

uint32_t orc_morton_closed_form(uint32_t qx, uint32_t qy, uint32_t qz)
{
    /* ComputeMortonCode (MortonCode.hlsl:24-27) puts z high; the live path
     * (LBVH.hlsl:98-111) puts x high, so the axes are swapped here. */
    return (part1by2(qx) << 2) | (part1by2(qy) << 1) | part1by2(qz);
}

/* ------------------------------------------------------------------------ */
/* LBVH.cpp:190-314: recursive MSD partition + heap emit + boxes.            */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t *codes;
    int32_t *ids;
    const float *bmin, *bmax;
    orc_node *nodes;
    uint32_t num_nodes;
    int reached;
} build_ctx;

static void swap_pair(build_ctx *b, int i, int j)
{
    uint32_t c = b->codes[i];
    int32_t d = b->ids[i];
    b->codes[i] = b->codes[j];
    b->ids[i] = b->ids[j];
    b->codes[j] = c;
    b->ids[j] = d;
}

static void build_rec(build_ctx *b, int position, int count, uint32_t bit,
                      int node, int axis)
{
    uint32_t *code = b->codes + position;
    if (bit == 0 || count <= 1 || ((uint32_t)node << 1) >= b->num_nodes) {
        orc_node *n = &b->nodes[node];
        n->prim_count = (count << 3) | 4 | axis;
        n->prim_pos = position;
        int id0 = b->ids[position];
        for (int k = 0; k < 3; ++k) {
            n->mx[k] = b->bmax[3 * (size_t)id0 + k];
            n->mn[k] = b->bmin[3 * (size_t)id0 + k];
        }
        for (int j = 1; j < count; ++j) {
            int id = b->ids[position + j];
            for (int k = 0; k < 3; ++k) {
                n->mx[k] = rmax(n->mx[k], b->bmax[3 * (size_t)id + k]);
                n->mn[k] = rmin(n->mn[k], b->bmin[3 * (size_t)id + k]);
            }
        }
        b->reached++;
        return;
    }
    int cero = 0;
    while (cero < count) {
        if ((code[cero] & bit) != 0) break;
        ++cero;
    }
    if (cero == count) { /* every code has a 0 at this bit */
        build_rec(b, position, count, bit >> 1, node, (axis + 1) % 3);
        return;
    }
    int one = count - 1;
    while (cero < one) {
        while (((code[cero] & bit) == 0) && (cero != one)) ++cero;
        while (((code[one] & bit) != 0) && (cero != one)) --one;
        if (cero != one) swap_pair(b, position + cero, position + one);
    }
    if ((code[0] & bit) != 0) { /* every code has a 1 at this bit */
        build_rec(b, position, count, bit >> 1, node, (axis + 1) % 3);
        return;
    }
    b->reached++;
    build_rec(b, position, one, bit >> 1, node << 1, (axis + 1) % 3);
    build_rec(b, position + one, count - one, bit >> 1, (node << 1) | 1,
              (axis + 1) % 3);
    orc_node *n = &b->nodes[node];
    const orc_node *l = &b->nodes[node << 1];
    const orc_node *r = &b->nodes[(node << 1) | 1];
    n->prim_count = (count << 3) | axis;
    n->prim_pos = position;
    for (int k = 0; k < 3; ++k) {
        n->mx[k] = rmax(l->mx[k], r->mx[k]);
        n->mn[k] = rmin(l->mn[k], r->mn[k]);
    }
}

int orc_lbvh_build(uint32_t *codes, int32_t *ids, uint32_t nt, uint32_t depth,
                   const float *bmin, const float *bmax, orc_node *nodes)
{
    uint32_t num_nodes = 1u << depth;
    for (uint32_t i = 0; i < num_nodes; ++i) {
        memset(&nodes[i], 0, sizeof(orc_node));
        nodes[i].prim_count = ORC_INVALID;
        nodes[i].prim_pos = ORC_INVALID;
    }
    build_ctx b = {codes, ids, bmin, bmax, nodes, num_nodes, 0};
    if (nt == 0) return 0;
    build_rec(&b, 0, (int)nt, 1u << 30, 1, 0);
    return b.reached;
}

static int cmp_pair(const void *pa, const void *pb)
{
    const uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
    return (a < b) ? -1 : (a > b) ? 1 : 0;
}

void orc_sort_pairs(uint32_t *codes, int32_t *ids, uint32_t n)
{
    /* key = code << 32 | id + 2^31: unique keys, so the order is the
     * cmpMorton order (code first, then id). */
    uint64_t *k = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i)
        k[i] = ((uint64_t)codes[i] << 32) | (uint32_t)((int64_t)ids[i] + 2147483648LL);
    qsort(k, n, sizeof(uint64_t), cmp_pair);
    for (uint32_t i = 0; i < n; ++i) {
        codes[i] = (uint32_t)(k[i] >> 32);
        ids[i] = (int32_t)((int64_t)(uint32_t)k[i] - 2147483648LL);
    }
    free(k);
}

/* ------------------------------------------------------------------------ */
/* Intersection.hlsl:18-78 and LBVH.hlsl:363-440.                            */
/* ------------------------------------------------------------------------ */
static inline float dot3(const float *a, const float *b)
{
    float s = a[0] * b[0] + a[1] * b[1];
    return s + a[2] * b[2];
}

static inline void cross3(float *r, const float *a, const float *b)
{
    r[0] = (a[1] * b[2]) - (a[2] * b[1]);
    r[1] = (a[2] * b[0]) - (a[0] * b[2]);
    r[2] = (a[0] * b[1]) - (a[1] * b[0]);
}

/* IntersectBox (Intersection.hlsl:18-42); min/max are IEEE minNum/maxNum. */
static inline void intersect_box(const orc_node *n, const float *o,
                                 const float *inv, float *t0, float *t1)
{
    float dmax[3], dmin[3];
    for (int k = 0; k < 3; ++k) {
        dmax[k] = (n->mx[k] - o[k]) * inv[k];
        dmin[k] = (n->mn[k] - o[k]) * inv[k];
    }
    float a = fminf(dmin[0], dmax[0]);
    float b = fmaxf(dmin[0], dmax[0]);
    a = fmaxf(a, fminf(dmin[1], dmax[1]));
    b = fminf(b, fmaxf(dmin[1], dmax[1]));
    a = fmaxf(a, fminf(dmin[2], dmax[2]));
    b = fminf(b, fmaxf(dmin[2], dmax[2]));
    if (a > b) a = b = -1.0f;
    *t0 = a;
    *t1 = b;
}

/* getIntersection (Intersection.hlsl:47-64). */
static inline void moller_trumbore(const float *o, const float *d,
                                   const float *A, const float *B,
                                   const float *C, float *u, float *v, float *t)
{
    float e1[3], e2[3], p[3], tv[3], q[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = B[k] - A[k];
        e2[k] = C[k] - A[k];
    }
    cross3(p, d, e2);
    float det = 1.0f / dot3(e1, p);
    for (int k = 0; k < 3; ++k) tv[k] = o[k] - A[k];
    *u = dot3(tv, p) * det;
    cross3(q, tv, e1);
    *v = dot3(d, q) * det;
    *t = dot3(e2, q) * det;
}

static inline int first_bit_high(uint32_t x) { return 31 - __builtin_clz(x); }
static inline int first_bit_low(uint32_t x) { return __builtin_ctz(x); }

static void trace_one(const orc_node *nodes, const int32_t *ids,
                      const float *xyz, const uint32_t *idx, const orc_ray *ray,
                      int32_t start, int anyhit, orc_hit *out)
{
    orc_hit best;
    best.tri = -1;
    best.u = -1.0f;
    best.v = -1.0f;
    best.t = ray->tmax;
    best.root = 1;
    float inv[3];
    int neg[3];
    for (int k = 0; k < 3; ++k) {
        inv[k] = 1.0f / ray->d[k];
        neg[k] = inv[k] < 0.0f ? 1 : 0;
    }
    int result = 0;
    uint32_t node = (uint32_t)start;
    uint32_t trail = 1u << first_bit_high(node);
    do {
        float t0, t1;
        intersect_box(&nodes[node], ray->o, inv, &t0, &t1);
        int p = first_bit_low(trail + 1);
        result++;
        if ((t0 > best.t) || (t1 < 0.0f)) {
            trail = (trail >> p) + 1;
            node = (node >> p) ^ 1u;
        } else if (nodes[node].prim_count & 4) {
            uint32_t cnt = (uint32_t)nodes[node].prim_count >> 3;
            uint32_t pos = (uint32_t)nodes[node].prim_pos;
            int stop = 0;
            for (uint32_t i = 0; i < cnt; ++i) {
                result += 65536;
                int32_t tri = ids[pos + i];
                const float *A = xyz + 3 * (size_t)idx[3 * (size_t)tri + 0];
                const float *B = xyz + 3 * (size_t)idx[3 * (size_t)tri + 1];
                const float *C = xyz + 3 * (size_t)idx[3 * (size_t)tri + 2];
                float u, v, t;
                moller_trumbore(ray->o, ray->d, A, B, C, &u, &v, &t);
                if ((ray->tri != tri) && (u >= 0.0f) && (v >= 0.0f) &&
                    ((u + v) <= 1.0f) && (t >= 0.0f) && (t < best.t)) {
                    best.tri = tri;
                    best.u = u;
                    best.v = v;
                    best.t = t;
                    best.root = (int32_t)node;
                    if (anyhit) {
                        stop = 1;
                        break;
                    }
                }
            }
            if (stop) break;
            trail = (trail >> p) + 1;
            node = (node >> p) ^ 1u;
        } else {
            trail = trail << 1;
            node = (node << 1) | (uint32_t)neg[nodes[node].prim_count & 3];
        }
    } while (trail > 1);
    best.visited = result;
    *out = best;
}

typedef struct {
    const orc_node *nodes;
    const int32_t *ids;
    const float *xyz;
    const uint32_t *idx;
    const orc_ray *rays;
    const int32_t *start;
    uint32_t begin, end;
    int anyhit;
    orc_hit *hits;
} trace_job;

static void *trace_worker(void *arg)
{
    trace_job *j = (trace_job *)arg;
    for (uint32_t i = j->begin; i < j->end; ++i)
        trace_one(j->nodes, j->ids, j->xyz, j->idx, &j->rays[i],
                  j->start ? j->start[i] : 1, j->anyhit, &j->hits[i]);
    return NULL;
}

void orc_trace(const orc_node *nodes, const int32_t *ids, const float *xyz,
               const uint32_t *idx, const orc_ray *rays, const int32_t *start,
               uint32_t n, int anyhit, orc_hit *hits, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if ((uint32_t)nthreads > n) nthreads = n ? (int)n : 1;
    pthread_t th[256];
    trace_job jobs[256];
    if (nthreads > 256) nthreads = 256;
    for (int t = 0; t < nthreads; ++t) {
        trace_job j = {nodes, ids, xyz, idx, rays, start,
                       (uint32_t)((uint64_t)n * t / nthreads),
                       (uint32_t)((uint64_t)n * (t + 1) / nthreads), anyhit, hits};
        jobs[t] = j;
    }
    if (nthreads == 1) {
        trace_worker(&jobs[0]);
        return;
    }
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, trace_worker, &jobs[t]);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------ */
/* BVH_IntersectP (BVH.hlsl:124-217): stack traversal of the flattened SAH   */
/* BVH, box test on every visited node, near child first by dirIsNeg[axis].  */
/* ------------------------------------------------------------------------ */
static void trace_bvh_one(const orc_node *nodes, const int32_t *ordered, const float *xyz,
                          const uint32_t *idx, const orc_ray *ray, int anyhit, int ref_leaf_loop,
                          orc_hit *out)
{
    orc_hit best;
    best.tri = -1;
    best.u = -1.0f;
    best.v = -1.0f;
    best.t = ray->tmax;
    best.root = 0;
    float inv[3];
    int neg[3];
    for (int k = 0; k < 3; ++k) {
        inv[k] = 1.0f / ray->d[k];
        neg[k] = inv[k] < 0.0f ? 1 : 0;
    }
    int32_t stack[64];
    int sp = 0;
    int32_t node = 0;
    int result = 0;
    int stop = 0;
    while (!stop) {
        float t0, t1;
        intersect_box(&nodes[node], ray->o, inv, &t0, &t1);
        result++;
        if ((t0 > best.t) || (t1 < 0.0f)) {
            if (sp == 0) break;
            node = stack[--sp];
        } else if (nodes[node].prim_count > 0) {
            const int32_t cnt = nodes[node].prim_count;
            const int32_t tests = ref_leaf_loop ? cnt + 1 : cnt;
            for (int32_t i = 0; i < tests; ++i) {
                result += 65536;
                const int32_t tri = ordered[nodes[node].prim_pos + (ref_leaf_loop ? 0 : i)];
                const float *A = xyz + 3 * (size_t)idx[3 * (size_t)tri + 0];
                const float *B = xyz + 3 * (size_t)idx[3 * (size_t)tri + 1];
                const float *C = xyz + 3 * (size_t)idx[3 * (size_t)tri + 2];
                float u, v, t;
                moller_trumbore(ray->o, ray->d, A, B, C, &u, &v, &t);
                if ((ray->tri != tri) && (u >= 0.0f) && (v >= 0.0f) && ((u + v) <= 1.0f) &&
                    (t >= 0.0f) && (t < best.t)) {
                    best.tri = tri;
                    best.u = u;
                    best.v = v;
                    best.t = t;
                    best.root = node;
                    if (anyhit) {
                        stop = 1;
                        break;
                    }
                }
            }
            if (stop || sp == 0) break;
            node = stack[--sp];
        } else {
            const int aux = neg[-nodes[node].prim_count];
            const int32_t first = node + 1, second = nodes[node].prim_pos;
            stack[sp++] = aux ? first : second;
            node = aux ? second : first;
        }
    }
    best.visited = result;
    *out = best;
}

typedef struct {
    const orc_node *nodes;
    const int32_t *ordered;
    const float *xyz;
    const uint32_t *idx;
    const orc_ray *rays;
    orc_hit *hits;
    uint32_t begin, end;
    int anyhit, ref_leaf_loop;
} bvh_job;

static void *bvh_worker(void *arg)
{
    bvh_job *j = (bvh_job *)arg;
    for (uint32_t i = j->begin; i < j->end; ++i)
        trace_bvh_one(j->nodes, j->ordered, j->xyz, j->idx, &j->rays[i], j->anyhit, j->ref_leaf_loop,
                      &j->hits[i]);
    return NULL;
}

void orc_trace_bvh(const orc_node *nodes, const int32_t *ordered, const float *xyz, const uint32_t *idx,
                   const orc_ray *rays, uint32_t n, int anyhit, int ref_leaf_loop, orc_hit *hits,
                   int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((uint32_t)nthreads > n) nthreads = n ? (int)n : 1;
    pthread_t th[256];
    bvh_job jobs[256];
    for (int t = 0; t < nthreads; ++t) {
        bvh_job j = {nodes, ordered, xyz, idx, rays, hits, (uint32_t)((uint64_t)n * t / nthreads),
                     (uint32_t)((uint64_t)n * (t + 1) / nthreads), anyhit, ref_leaf_loop};
        jobs[t] = j;
    }
    if (nthreads == 1) {
        bvh_worker(&jobs[0]);
        return;
    }
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, bvh_worker, &jobs[t]);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* Closest hit over every triangle in index order (no structure): the
 * brute-force check of any traversal.  hit.root is 0. */
void orc_trace_brute(const float *xyz, const uint32_t *idx, uint32_t nt, const orc_ray *rays, uint32_t n,
                     orc_hit *hits)
{
    for (uint32_t r = 0; r < n; ++r) {
        const orc_ray *ray = &rays[r];
        orc_hit best = {-1, -1.0f, -1.0f, ray->tmax, 0, 0};
        for (uint32_t tri = 0; tri < nt; ++tri) {
            const float *A = xyz + 3 * (size_t)idx[3 * (size_t)tri + 0];
            const float *B = xyz + 3 * (size_t)idx[3 * (size_t)tri + 1];
            const float *C = xyz + 3 * (size_t)idx[3 * (size_t)tri + 2];
            float u, v, t;
            moller_trumbore(ray->o, ray->d, A, B, C, &u, &v, &t);
            if ((ray->tri != (int32_t)tri) && (u >= 0.0f) && (v >= 0.0f) && ((u + v) <= 1.0f) && (t >= 0.0f) &&
                (t < best.t)) {
                best.tri = (int32_t)tri;
                best.u = u;
                best.v = v;
                best.t = t;
            }
        }
        hits[r] = best;
    }
}

/* ------------------------------------------------------------------------ */
/* Pinhole ray generation: the pixel-centre rule of CSPrimaryRays.hlsl:37-40 */
/* with an explicit camera basis; maxT and the self id as at :49-53.         */
/* ------------------------------------------------------------------------ */
static void ray_defaults(orc_ray *r)
{
    for (int k = 0; k < 3; ++k) r->refl[k] = 1.0f;
    r->tmin = 0.0f;
    r->tmax = 10000000000000000.0f;
    r->tri = -1;
}

void orc_raygen_pinhole_jitter(const float cam[14], uint32_t w, uint32_t h, const float *jxy,
                               orc_ray *rays)
{
    const float *pos = cam, *fwd = cam + 3, *right = cam + 6, *up = cam + 9;
    const float tx = cam[12], ty = cam[13];
    const float fw = (float)w, fh = (float)h;
    for (uint32_t py = 0; py < h; ++py) {
        for (uint32_t px = 0; px < w; ++px) {
            const size_t pix = (size_t)py * w + px;
            const float jx = jxy ? jxy[2 * pix + 0] : 0.0f;
            const float jy = jxy ? jxy[2 * pix + 1] : 0.0f;
            float x = (((float)(2u * px + 1u) + jx) - fw) / fw;
            float y = -((((float)(2u * py + 1u) + jy) - fh) / fh);
            float sx = x * tx, sy = y * ty;
            float d[3];
            for (int k = 0; k < 3; ++k) {
                float a = fwd[k] + sx * right[k];
                d[k] = a + sy * up[k];
            }
            float len = sqrtf(dot3(d, d));
            orc_ray *r = &rays[pix];
            for (int k = 0; k < 3; ++k) {
                r->o[k] = pos[k];
                r->d[k] = d[k] / len;
            }
            ray_defaults(r);
        }
    }
}

void orc_raygen_pinhole(const float cam[14], uint32_t w, uint32_t h, orc_ray *rays)
{
    orc_raygen_pinhole_jitter(cam, w, h, NULL, rays);
}

void orc_pcg32_draws(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t *out)
{
    orc_pcg32 rng = orc_pcg_seed(initstate, initseq);
    for (uint32_t i = 0; i < n; ++i) out[i] = orc_pcg_next(&rng);
}

void orc_pcg_jitter(uint64_t seed, uint32_t frame, uint32_t sample, uint32_t w, uint32_t h, float *jxy)
{
    for (uint32_t pix = 0; pix < w * h; ++pix) {
        orc_pcg32 rng = orc_pcg_seed(orc_sample_seed(seed, frame, sample), pix);
        const float u0 = (float)(orc_pcg_next(&rng) >> 8) * (1.0f / 16777216.0f);
        jxy[2 * (size_t)pix + 0] = 2.0f * u0 - 1.0f;
        const float u1 = (float)(orc_pcg_next(&rng) >> 8) * (1.0f / 16777216.0f);
        jxy[2 * (size_t)pix + 1] = 2.0f * u1 - 1.0f;
    }
}

/* Row r of the reference's Matrix4 (row-major, translation in column 3)
 * applied to (v, 1): the HLSL mul(float4(v, 1), g_mfWorld) of a column-major
 * cbuffer matrix, summed left to right. */
static inline float mat_row(const float *m, int r, const float v[3])
{
    return ((m[4 * r + 0] * v[0] + m[4 * r + 1] * v[1]) + m[4 * r + 2] * v[2]) + m[4 * r + 3] * 1.0f;
}

void orc_raygen_world(const float world[16], uint32_t w, uint32_t h, const float *jxy, orc_ray *rays)
{
    const float fw = (float)w, fh = (float)h;
    const float inverse = 1.0f / fh;
    const float z = 2.0f;
    const float scale = 0.008f / fabsf(0.035f * z);
    const float zero[3] = {0.0f, 0.0f, 0.0f};
    const float ow = mat_row(world, 3, zero);
    float org[3];
    for (int k = 0; k < 3; ++k) org[k] = mat_row(world, k, zero) / ow;
    for (uint32_t py = 0; py < h; ++py) {
        for (uint32_t px = 0; px < w; ++px) {
            const size_t pix = (size_t)py * w + px;
            const float jx = jxy ? jxy[2 * pix + 0] : 0.0f;
            const float jy = jxy ? jxy[2 * pix + 1] : 0.0f;
            const float y = -((((2.0f * (float)py + jy) + 1.0f) - fh) * inverse);
            const float x = (((2.0f * (float)px + jx) + 1.0f) - fw) * inverse;
            const float p[3] = {scale * x, scale * y, scale * z};
            const float pw = mat_row(world, 3, p);
            float d[3];
            for (int k = 0; k < 3; ++k) d[k] = mat_row(world, k, p) / pw - org[k];
            const float len = sqrtf(dot3(d, d));
            orc_ray *r = &rays[pix];
            for (int k = 0; k < 3; ++k) {
                r->o[k] = org[k];
                r->d[k] = d[k] / len;
            }
            ray_defaults(r);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Direct lighting with shadow rays (CSComputeColorRT.hlsl:54-177, 195-198). */
/* ------------------------------------------------------------------------ */
typedef struct {
    const orc_node *nodes;
    const int32_t *ids;
    const float *xyz;
    const uint32_t *idx;
    const orc_ray *rays;
    const orc_hit *hits;
    const float *lights;
    uint32_t nlights;
    const float *shade;
    float *rgba;
    uint8_t *occluded;
    uint32_t begin, end;
} shade_job;

static void *shade_worker(void *arg)
{
    shade_job *j = (shade_job *)arg;
    const float ambient = j->shade[0], albedo = j->shade[1], env = j->shade[2];
    for (uint32_t i = j->begin; i < j->end; ++i) {
        const orc_ray *ray = &j->rays[i];
        const orc_hit *hit = &j->hits[i];
        float *out = j->rgba + 4 * (size_t)i;
        if (hit->tri < 0) {
            out[0] = out[1] = out[2] = env;
            out[3] = 1.0f;
            if (j->occluded)
                for (uint32_t l = 0; l < j->nlights; ++l)
                    j->occluded[(size_t)i * j->nlights + l] = 2;
            continue;
        }
        const int32_t tri = hit->tri;
        const float *A = j->xyz + 3 * (size_t)j->idx[3 * (size_t)tri + 0];
        const float *B = j->xyz + 3 * (size_t)j->idx[3 * (size_t)tri + 1];
        const float *C = j->xyz + 3 * (size_t)j->idx[3 * (size_t)tri + 2];
        float e1[3], e2[3], n[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = B[k] - A[k];
            e2[k] = C[k] - A[k];
        }
        cross3(n, e1, e2);
        float nl = sqrtf(dot3(n, n));
        for (int k = 0; k < 3; ++k) n[k] = n[k] / nl;
        float nn[3] = {n[0], n[1], n[2]};
        if (dot3(nn, ray->d) > 0.0f)
            for (int k = 0; k < 3; ++k) nn[k] = -nn[k];
        float hp[3];
        for (int k = 0; k < 3; ++k) hp[k] = ray->o[k] + hit->t * ray->d[k];
        float diffuse = 0.0f;
        for (uint32_t l = 0; l < j->nlights; ++l) {
            const float *L = j->lights + 3 * (size_t)l;
            float ld[3];
            for (int k = 0; k < 3; ++k) ld[k] = hp[k] - L[k];
            float dist = sqrtf(dot3(ld, ld));
            for (int k = 0; k < 3; ++k) ld[k] = ld[k] / dist;
            float ls = dot3(n, ld);
            float rs = dot3(n, ray->d);
            uint8_t occ = 2;
            if ((ls * rs) > 0.0f) {
                orc_ray sr;
                for (int k = 0; k < 3; ++k) {
                    sr.o[k] = hp[k];
                    sr.d[k] = -ld[k];
                    sr.refl[k] = ray->refl[k];
                }
                sr.tmin = ray->tmin;
                sr.tmax = dist;
                sr.tri = tri;
                orc_hit sh;
                trace_one(j->nodes, j->ids, j->xyz, j->idx, &sr, hit->root, 0, &sh);
                occ = sh.tri < 0 ? 0 : 1;
                if (sh.tri < 0) {
                    float dd = -dot3(nn, ld);
                    diffuse = diffuse + (dd > 0.0f ? dd : 0.0f);
                }
            }
            if (j->occluded) j->occluded[(size_t)i * j->nlights + l] = occ;
        }
        float c = (ambient + diffuse) * albedo;
        out[0] = out[1] = out[2] = c;
        out[3] = 1.0f;
    }
    return NULL;
}

void orc_render_direct(const orc_node *nodes, const int32_t *ids,
                       const float *xyz, const uint32_t *idx,
                       const orc_ray *rays, const orc_hit *hits, uint32_t n,
                       const float *lights, uint32_t nlights,
                       const float shade[3], float *rgba, uint8_t *occluded,
                       int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((uint32_t)nthreads > n) nthreads = n ? (int)n : 1;
    pthread_t th[256];
    shade_job jobs[256];
    for (int t = 0; t < nthreads; ++t) {
        shade_job j = {nodes, ids, xyz, idx, rays, hits, lights, nlights, shade,
                       rgba, occluded,
                       (uint32_t)((uint64_t)n * t / nthreads),
                       (uint32_t)((uint64_t)n * (t + 1) / nthreads)};
        jobs[t] = j;
    }
    if (nthreads == 1) {
        shade_worker(&jobs[0]);
        return;
    }
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, shade_worker, &jobs[t]);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}
