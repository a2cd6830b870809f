/*
 * Sanitizer driver for the CPU oracle (TEST INFRASTRUCTURE ONLY): runs every
 * oracle entry point on small seeded scenes, in an AddressSanitizer +
 * UndefinedBehaviorSanitizer build (tests/sanitize/Makefile), so that an out
 * of bounds access, a leak or undefined arithmetic aborts the run.  Exit 0
 * when every call returned and the basic invariants hold.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lbvh_oracle.h"

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static float frand(void)
{
    rng = rng * 6364136223846793005ull + 1442695040888963407ull;
    return (float)(rng >> 40) / 16777216.0f;
}

#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            exit(1);                                                      \
        }                                                                 \
    } while (0)

static void run_scene(uint32_t nt, uint32_t depth, int dup)
{
    const uint32_t nv = 3 * nt;
    float *xyz = malloc(sizeof(float) * 3 * nv);
    uint32_t *idx = malloc(sizeof(uint32_t) * 3 * nt);
    for (uint32_t t = 0; t < nt; ++t) {
        const float cx = dup ? 0.5f : frand(), cy = dup ? 0.5f : frand(), cz = frand();
        for (int v = 0; v < 3; ++v) {
            xyz[9 * t + 3 * v + 0] = cx + 0.02f * frand();
            xyz[9 * t + 3 * v + 1] = cy + 0.02f * frand();
            xyz[9 * t + 3 * v + 2] = cz + 0.02f * frand();
            idx[3 * t + v] = 3 * t + v;
        }
    }
    float *bmin = malloc(12 * nt), *bmax = malloc(12 * nt), *cent = malloc(12 * nt), scene[6];
    orc_prim_bounds(xyz, idx, nt, bmin, bmax, cent, scene);
    uint32_t *codes = malloc(4 * nt);
    int32_t *ids = malloc(4 * nt);
    orc_morton_codes(cent, nt, scene, codes);
    for (uint32_t t = 0; t < nt; ++t) {
        CHECK(codes[t] < (1u << 30));
        ids[t] = (int32_t)t;
    }
    uint32_t *c2 = malloc(4 * nt);
    int32_t *i2 = malloc(4 * nt);
    memcpy(c2, codes, 4 * nt);
    memcpy(i2, ids, 4 * nt);
    orc_sort_pairs(c2, i2, nt);
    for (uint32_t t = 1; t < nt; ++t) CHECK(c2[t - 1] <= c2[t]);
    orc_node *nodes = malloc(sizeof(orc_node) * ((size_t)1 << depth));
    const int reached = orc_lbvh_build(codes, ids, nt, depth, bmin, bmax, nodes);  /* in-place partition */
    CHECK(reached >= 1);
    const int reached2 = orc_lbvh_build(c2, i2, nt, depth, bmin, bmax, nodes);     /* presorted */
    CHECK(reached2 == reached);

    /* rays: a pinhole view, jittered, and the camera-matrix rule */
    const uint32_t W = 24, H = 16, n = W * H;
    const float cam[14] = {0.5f, 0.5f, -1.5f, 0, 0, 1, 1, 0, 0, 0, 1, 0, 0.8f, 0.55f};
    orc_ray *rays = malloc(sizeof(orc_ray) * n);
    orc_hit *hits = malloc(sizeof(orc_hit) * n), *h2 = malloc(sizeof(orc_hit) * n);
    float *jit = malloc(8 * n);
    orc_pcg_jitter(7, 1, 0, W, H, jit);
    orc_raygen_pinhole_jitter(cam, W, H, jit, rays);
    float world[16] = {1, 0, 0, 0.5f, 0, 1, 0, 0.5f, 0, 0, 1, -1.5f, 0, 0, 0, 1};
    orc_raygen_world(world, W, H, NULL, rays);
    orc_raygen_pinhole(cam, W, H, rays);
    orc_trace(nodes, i2, xyz, idx, rays, NULL, n, 0, hits, 2);
    int32_t *start = malloc(4 * n);
    for (uint32_t i = 0; i < n; ++i) start[i] = hits[i].root;
    orc_trace(nodes, i2, xyz, idx, rays, start, n, 1, h2, 1);
    orc_trace_brute(xyz, idx, nt, rays, n, h2);
    for (uint32_t i = 0; i < n; ++i)
        if (hits[i].tri >= 0) CHECK(h2[i].tri >= 0);
    const float lights[6] = {0.5f, 2.5f, -1.0f, 2.0f, 0.5f, 0.5f};
    const float shade[3] = {0.1f, 0.8f, 0.25f};
    float *rgba = malloc(16 * n);
    uint8_t *occ = malloc(2 * n);
    orc_render_direct(nodes, i2, xyz, idx, rays, hits, n, lights, 2, shade, rgba, occ, 3);
    double *gi = malloc(8 * n);
    orc_render_gi(nodes, i2, xyz, idx, cam, W, H, 2, 3, 0.7, 1.0, 5, 0, 0, gi, 2);
    orc_material mats[2];
    memset(mats, 0, sizeof(mats));
    for (int k = 0; k < 3; ++k) {
        mats[0].diffuse[k] = 0.7f;
        mats[1].specular[k] = 0.9f;
    }
    mats[0].diffuse[3] = 1.0f;
    mats[1].alpha = 0.5f;
    mats[1].ambient[3] = 1.5f;
    uint32_t *mid = malloc(4 * nt);
    for (uint32_t t = 0; t < nt; ++t) mid[t] = t & 1;
    float *nrm = malloc(12 * nv);
    for (uint32_t v = 0; v < nv; ++v) {
        nrm[3 * v] = 0.0f;
        nrm[3 * v + 1] = 0.0f;
        nrm[3 * v + 2] = 1.0f;
    }
    orc_whitted_params p;
    memset(&p, 0, sizeof(p));
    p.light[0] = 0.5f;
    p.light[1] = 2.5f;
    p.light[2] = -1.0f;
    p.bounces = 3;
    p.switches = 3;
    p.env_flag = -1;
    p.seed = 3;
    orc_render_whitted(nodes, i2, xyz, idx, nrm, mid, mats, cam, W, H, &p, rgba, h2, occ, 2);
    p.ref_order = 1;
    orc_render_whitted(nodes, i2, xyz, idx, NULL, NULL, NULL, cam, W, H, &p, rgba, NULL, NULL, 1);

    /* the SAH-layout traversal over a hand-made two-leaf tree of the first primitives */
    if (nt >= 2) {
        orc_node b[3];
        memset(b, 0, sizeof(b));
        for (int k = 0; k < 3; ++k) {
            b[0].mn[k] = scene[k] - 1.0f;
            b[0].mx[k] = scene[3 + k] + 1.0f;
            b[1].mn[k] = b[2].mn[k] = b[0].mn[k];
            b[1].mx[k] = b[2].mx[k] = b[0].mx[k];
        }
        b[0].prim_count = 0;
        b[0].prim_pos = 2;
        b[1].prim_count = 1;
        b[1].prim_pos = 0;
        b[2].prim_count = 1;
        b[2].prim_pos = 1;
        int32_t ord[2] = {0, 1};
        orc_trace_bvh(b, ord, xyz, idx, rays, n, 0, 1, h2, 2);
        orc_trace_bvh(b, ord, xyz, idx, rays, n, 1, 0, h2, 1);
    }
    uint32_t draws[6];
    orc_pcg32_draws(42, 54, 6, draws);
    CHECK(draws[0] == 0xA15C02B7u);
    CHECK(orc_morton_closed_form(1023, 1023, 1023) == (1u << 30) - 1);
    free(xyz); free(idx); free(bmin); free(bmax); free(cent); free(codes); free(ids); free(c2); free(i2);
    free(nodes); free(rays); free(hits); free(h2); free(jit); free(start); free(rgba); free(occ); free(gi);
    free(mid); free(nrm);
}

int main(void)
{
    run_scene(1, 4, 0);
    run_scene(2, 6, 1);     /* identical centroids */
    run_scene(37, 5, 0);    /* depth limit: multi-primitive leaves */
    run_scene(500, 12, 0);
    run_scene(300, 10, 1);  /* many equal codes */
    printf("oracle sanitizer run ok\n");
    return 0;
}
