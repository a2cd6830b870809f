/*
 * Whitted frame oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates, scalar and in source order, the reference's ray-tracing frame:
 * RayTracerCS::Render (RayTracerCS.cpp:286-295) runs the primary-ray stage
 * once, then bounces + 1 times the intersection stage
 * (CSIntersections.hlsl:18-39: LBVH_IntersectP for rays whose id is above -2)
 * and the colour stage (CSComputeColorRT.hlsl:18-262).  This file follows
 * the colour stage line by line:
 *   :28-35   id < g_iEnvMappingFlag -> the result texture is 1 when id > -2;
 *   :56-77   flat normal normalize(cross(B - A, C - A)) or, with Phong on,
 *            normalize(Vec3BaryCentric(nA, nB, nC, (u, v))) (Intersection.hlsl:98-108);
 *   :124-131 the normal is flipped to face the ray ("inverseNormal");
 *   :133-142 hit point o + t d, light direction (hit - light) / |hit - light|,
 *            new direction normalize(reflect(d, n'));
 *   :145-176 when dot(n, L) * dot(n, d) > 0: a shadow ray {o = hit, d = -L,
 *            maxT = |hit - light|, self = hit} traced from the hit's leaf;
 *            if it hits nothing (or shadows are off) specular =
 *            max(0, -dot(L, R))^49 and diffuse = max(0, -dot(n', L)), both
 *            assigned to all four components;
 *   :181-198 colour = ((ambient + diffuse * Kd.w) * Kd
 *            + specular * (1 - Kd.w) * Ks) * (factor, 1), Ks.w times the gloss
 *            switch, added to the accumulator, which becomes the result;
 *   :200-235 continuation: reflect when alpha >= a uniform draw (factor *=
 *            Ks.rgb), else refract(d, n', eta) with eta = Ka.w, inverted when
 *            the normal was not flipped; length < 1e-6 (total internal
 *            reflection) reflects, otherwise factor *= Kd.rgb;
 *   :238-248 a miss adds factor * environment (w = 1) and sets the id to -2.
 * The texture, normal-map and gloss-map samples are left out (the reference
 * with texturing and normal mapping off).  Deliberate choices, shared with
 * the device kernel: pow(x, 49) is x^32 * x^16 * x by squaring; refract is
 * the HLSL definition t = (eta i + (eta cosi - sqrt|cost2|) n) * (cost2 > 0),
 * cosi = dot(-i, n), cost2 = 1 - eta^2 (1 - cosi^2); reflect(i, n) =
 * i - (2 dot(n, i)) n; the reflect/refract draw is PCG32 keyed by (seed,
 * frame, pixel) instead of the reference's random texture; the environment
 * is a constant colour (a uniform cube map).
 *
 * ref_order = 1 starts every closest-hit and shadow ray at the previous hit's
 * leaf, as the reference passes g_uIntersections[index].iRoot; ref_order = 0
 * starts them at the root, which is what the device does.  Both visit every
 * leaf, so they differ only where two triangles tie at exactly the same t.
 *
 * Every float expression is IEEE single precision in the order written,
 * with no FMA contraction.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lbvh_oracle.h"
#include "pcg32.h"

/* The float draw of the device's pcg_unit: the top 24 bits, exact. */
static float pcg_unit_f(orc_pcg32 *r) { return (float)(orc_pcg_next(r) >> 8) * (1.0f / 16777216.0f); }

static inline float d3(const float *a, const float *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

static inline float pow49(float x)
{
    float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8, x32 = x16 * x16;
    return (x32 * x16) * x;
}

static const orc_material k_default_material = {
    {0.5f, 0.5f, 0.5f, 0.8f}, {0.0f, 0.0f, 0.0f, 1.0f}, {0.8f, 0.8f, 0.8f, 1.0f}, 0.0f, 0.0f, 0u, 0u};

typedef struct {
    const orc_node *nodes;
    const int32_t *ids;
    const float *xyz;
    const uint32_t *idx;
    const float *normals;
    const uint32_t *mat_ids;
    const orc_material *mats;
    const orc_ray *prim;
    uint32_t w;
    const orc_whitted_params *p;
    float *rgba;
    orc_hit *hits0;
    uint8_t *occ0;
    uint64_t rays;
    uint32_t begin, end;
} wjob;

static void *whitted_worker(void *arg)
{
    wjob *j = (wjob *)arg;
    const orc_whitted_params *p = j->p;
    const int shadows = (p->switches & 1u) != 0, phong = (p->switches & 2u) != 0 && j->normals != NULL;
    const float gloss = (p->switches & 4u) ? 1.0f : 0.0f;
    for (uint32_t pix = j->begin; pix < j->end; ++pix) {
        orc_ray ray = j->prim[pix];
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f}, res[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        orc_pcg32 rng = orc_pcg_seed(orc_sample_seed(p->seed, p->frame, 0u), pix);
        int32_t start = 1;
        if (j->occ0) j->occ0[pix] = 2;
        for (uint32_t pass = 0; pass <= p->bounces; ++pass) {
            orc_hit h;
            orc_trace(j->nodes, j->ids, j->xyz, j->idx, &ray, &start, 1, 0, &h, 1);
            j->rays += 1;
            if (j->hits0 && pass == 0) j->hits0[pix] = h;
            if (h.tri < p->env_flag) {
                res[0] = res[1] = res[2] = res[3] = 1.0f;
                break;
            }
            if (h.tri < 0) {
                for (int c = 0; c < 3; ++c) acc[c] = acc[c] + ray.refl[c] * p->env[c];
                acc[3] = acc[3] + 1.0f;
                memcpy(res, acc, sizeof(res));
                break;
            }
            const int32_t tri = h.tri;
            const orc_material *M = j->mats ? &j->mats[j->mat_ids ? j->mat_ids[tri] : 0u] : &k_default_material;
            const uint32_t ia = j->idx[3 * (size_t)tri], ib = j->idx[3 * (size_t)tri + 1],
                           ic = j->idx[3 * (size_t)tri + 2];
            float n[3];
            if (phong) {
                const float *na = j->normals + 3 * (size_t)ia, *nb = j->normals + 3 * (size_t)ib,
                            *nc = j->normals + 3 * (size_t)ic;
                const float w = 1.0f - (h.u + h.v);
                for (int c = 0; c < 3; ++c) n[c] = (na[c] * w + nb[c] * h.u) + nc[c] * h.v;
            } else {
                const float *A = j->xyz + 3 * (size_t)ia, *B = j->xyz + 3 * (size_t)ib, *C = j->xyz + 3 * (size_t)ic;
                const float e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
                const float e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
                n[0] = (e1[1] * e2[2]) - (e1[2] * e2[1]);
                n[1] = (e1[2] * e2[0]) - (e1[0] * e2[2]);
                n[2] = (e1[0] * e2[1]) - (e1[1] * e2[0]);
            }
            const float nl = sqrtf(d3(n, n));
            for (int c = 0; c < 3; ++c) n[c] = n[c] / nl;
            float m[3] = {n[0], n[1], n[2]};
            const int inverse = d3(m, ray.d) > 0.0f;
            if (inverse)
                for (int c = 0; c < 3; ++c) m[c] = -m[c];
            float hp[3], L[3], R[3];
            for (int c = 0; c < 3; ++c) hp[c] = ray.o[c] + h.t * ray.d[c];
            for (int c = 0; c < 3; ++c) L[c] = hp[c] - p->light[c];
            const float dist = sqrtf(d3(L, L));
            for (int c = 0; c < 3; ++c) L[c] = L[c] / dist;
            const float k2 = 2.0f * d3(m, ray.d);
            for (int c = 0; c < 3; ++c) R[c] = ray.d[c] - k2 * m[c];
            const float rl = sqrtf(d3(R, R));
            for (int c = 0; c < 3; ++c) R[c] = R[c] / rl;
            int lit = 0;
            float D = 0.0f, S = 0.0f;
            if ((d3(n, L) * d3(n, ray.d)) > 0.0f) {
                int unocc = 1;
                if (shadows) {
                    orc_ray sr = ray;
                    for (int c = 0; c < 3; ++c) {
                        sr.o[c] = hp[c];
                        sr.d[c] = -L[c];
                    }
                    sr.tri = tri;
                    sr.tmax = dist;
                    int32_t sstart = p->ref_order ? h.root : 1;
                    orc_hit sh;
                    orc_trace(j->nodes, j->ids, j->xyz, j->idx, &sr, &sstart, 1, 0, &sh, 1);
                    j->rays += 1;
                    unocc = sh.tri < 0;
                    if (j->occ0 && pass == 0) j->occ0[pix] = unocc ? 0 : 1;
                }
                if (unocc) {
                    lit = 1;
                    const float sp = -d3(L, R);
                    S = pow49(sp > 0.0f ? sp : 0.0f);
                    const float dd = -d3(m, L);
                    D = dd > 0.0f ? dd : 0.0f;
                }
            }
            const float tw = M->diffuse[3];
            for (int c = 0; c < 4; ++c) {
                const float Dc = lit ? D : (c == 3 ? 1.0f : 0.0f);
                const float Sc = lit ? S : (c == 3 ? 1.0f : 0.0f);
                const float spc = c == 3 ? M->specular[3] * gloss : M->specular[c];
                const float fc = c == 3 ? 1.0f : ray.refl[c];
                const float dcol = (M->ambient[c] + Dc * tw) * M->diffuse[c];
                const float scol = (Sc * (1.0f - tw)) * spc;
                acc[c] = acc[c] + (dcol + scol) * fc;
                res[c] = acc[c];
            }
            const float u = pcg_unit_f(&rng);
            if (M->alpha >= u) {
                for (int c = 0; c < 3; ++c) {
                    ray.refl[c] = ray.refl[c] * M->specular[c];
                    ray.d[c] = R[c];
                }
            } else {
                float eta = M->ambient[3];
                if (!inverse) eta = 1.0f / eta;
                const float cosi = -d3(ray.d, m);
                const float cost2 = 1.0f - (eta * eta) * (1.0f - cosi * cosi);
                const float g = eta * cosi - sqrtf(fabsf(cost2));
                const float keep = cost2 > 0.0f ? 1.0f : 0.0f;
                float t[3];
                for (int c = 0; c < 3; ++c) t[c] = (eta * ray.d[c] + g * m[c]) * keep;
                const float tl = sqrtf(d3(t, t));
                if (tl < 0.000001f) {
                    for (int c = 0; c < 3; ++c) ray.d[c] = R[c];
                } else {
                    for (int c = 0; c < 3; ++c) {
                        ray.d[c] = t[c] / tl;
                        ray.refl[c] = ray.refl[c] * M->diffuse[c];
                    }
                }
            }
            for (int c = 0; c < 3; ++c) ray.o[c] = hp[c];
            ray.tri = tri;
            start = p->ref_order ? h.root : 1;
        }
        memcpy(j->rgba + 4 * (size_t)pix, res, sizeof(res));
    }
    return NULL;
}

uint64_t orc_render_whitted(const orc_node *nodes, const int32_t *ids, const float *xyz,
                            const uint32_t *idx, const float *normals, const uint32_t *mat_ids,
                            const orc_material *mats, const float cam[14], uint32_t w, uint32_t h,
                            const orc_whitted_params *p, float *rgba, orc_hit *hits0, uint8_t *occ0,
                            int nthreads)
{
    const uint32_t n = w * h;
    orc_ray *prim = (orc_ray *)malloc(sizeof(orc_ray) * (n ? n : 1));
    if (!prim) return 0;
    orc_raygen_pinhole(cam, w, h, prim);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((uint32_t)nthreads > n) nthreads = n ? (int)n : 1;
    pthread_t th[256];
    wjob jobs[256];
    for (int t = 0; t < nthreads; ++t) {
        wjob j = {nodes, ids, xyz, idx, normals, mat_ids, mats, prim, w, p, rgba, hits0, occ0, 0,
                  (uint32_t)((uint64_t)n * t / nthreads), (uint32_t)((uint64_t)n * (t + 1) / nthreads)};
        jobs[t] = j;
    }
    if (nthreads == 1) {
        whitted_worker(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, whitted_worker, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    free(prim);
    uint64_t rays = 0;
    for (int t = 0; t < nthreads; ++t) rays += jobs[t].rays;
    return rays;
}
