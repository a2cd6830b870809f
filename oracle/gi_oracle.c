/*
 * FP64 CPU path tracer -- TEST INFRASTRUCTURE ONLY (the GI oracle).
 *
 * Renders the diffuse path-tracing job of rt_render_jobs in double precision:
 * the loop structure of CSComputeColorGI.hlsl:153-230 (bounce, hemisphere
 * flip to the incoming side, throughput, sky on a miss) with the sampling
 * the batched workload defines: cosine-weighted bounce directions (basis of
 * Duff et al., JCGT 2017), Lambert albedo, pixel jitter, PCG32 streams keyed
 * by (seed, frame, sample, pixel).  Traversal is the trail algorithm of
 * LBVH.hlsl:363-440 over the oracle's heap with box and Moller-Trumbore tests
 * evaluated in double on the float scene data.
 *
 * With first_sample = 0 it draws the same random numbers as the device for
 * the same samples (same-stream comparison); with a large first_sample and
 * many samples it gives an independent converged reference.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <string.h>

#include "lbvh_oracle.h"
#include "pcg32.h"

static double pcg_unit(orc_pcg32 *r) { return (double)(orc_pcg_next(r) >> 8) * (1.0 / 16777216.0); }

static int fbh(uint32_t x) { return 31 - __builtin_clz(x); }
static int fbl(uint32_t x) { return __builtin_ctz(x); }

/* Closest hit in double; returns the triangle id or -1, sets *t_out and the
 * triangle's sorted slot in *slot_out. */
static int32_t trace_d(const orc_node *nodes, const int32_t *ids, const float *xyz,
                       const uint32_t *idx, const double o[3], const double d[3], int32_t self,
                       double *t_out)
{
    double inv[3];
    int neg[3];
    for (int k = 0; k < 3; ++k) {
        inv[k] = 1.0 / d[k];
        neg[k] = inv[k] < 0.0;
    }
    double best = 1e16;
    int32_t best_id = -1;
    uint32_t node = 1, trail = 1u << fbh(node);
    do {
        const orc_node *n = &nodes[node];
        double a = 0, b = 0;
        {
            double dmin[3], dmax[3];
            for (int k = 0; k < 3; ++k) {
                dmax[k] = ((double)n->mx[k] - o[k]) * inv[k];
                dmin[k] = ((double)n->mn[k] - o[k]) * inv[k];
            }
            a = fmin(dmin[0], dmax[0]);
            b = fmax(dmin[0], dmax[0]);
            a = fmax(a, fmin(dmin[1], dmax[1]));
            b = fmin(b, fmax(dmin[1], dmax[1]));
            a = fmax(a, fmin(dmin[2], dmax[2]));
            b = fmin(b, fmax(dmin[2], dmax[2]));
            if (a > b) a = b = -1.0;
        }
        int p = fbl(trail + 1);
        if (a > best || b < 0.0) {
            trail = (trail >> p) + 1;
            node = (node >> p) ^ 1u;
        } else if (n->prim_count & 4) {
            uint32_t cnt = (uint32_t)n->prim_count >> 3, pos = (uint32_t)n->prim_pos;
            for (uint32_t i = 0; i < cnt; ++i) {
                int32_t tri = ids[pos + i];
                const float *A = xyz + 3 * (size_t)idx[3 * (size_t)tri + 0];
                const float *B = xyz + 3 * (size_t)idx[3 * (size_t)tri + 1];
                const float *C = xyz + 3 * (size_t)idx[3 * (size_t)tri + 2];
                double e1[3], e2[3], pv[3], tv[3], q[3];
                for (int k = 0; k < 3; ++k) {
                    e1[k] = (double)B[k] - (double)A[k];
                    e2[k] = (double)C[k] - (double)A[k];
                    tv[k] = o[k] - (double)A[k];
                }
                pv[0] = d[1] * e2[2] - d[2] * e2[1];
                pv[1] = d[2] * e2[0] - d[0] * e2[2];
                pv[2] = d[0] * e2[1] - d[1] * e2[0];
                double det = 1.0 / (e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2]);
                double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * det;
                q[0] = tv[1] * e1[2] - tv[2] * e1[1];
                q[1] = tv[2] * e1[0] - tv[0] * e1[2];
                q[2] = tv[0] * e1[1] - tv[1] * e1[0];
                double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * det;
                double t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * det;
                if (tri != self && u >= 0.0 && v >= 0.0 && u + v <= 1.0 && t >= 0.0 && t < best) {
                    best = t;
                    best_id = tri;
                }
            }
            trail = (trail >> p) + 1;
            node = (node >> p) ^ 1u;
        } else {
            trail <<= 1;
            node = (node << 1) | (uint32_t)neg[n->prim_count & 3];
        }
    } while (trail > 1);
    *t_out = best;
    return best_id;
}

typedef struct {
    const orc_node *nodes;
    const int32_t *ids;
    const float *xyz;
    const uint32_t *idx;
    const float *cam;
    uint32_t w, h, spp, max_bounces, frame, first_sample;
    double albedo, sky;
    uint64_t seed;
    double *out;
    uint64_t rays;
    uint32_t row0, row1;
} gi_job;

static void *gi_worker(void *arg)
{
    gi_job *J = (gi_job *)arg;
    const float *cam = J->cam;
    const double fw = J->w, fh = J->h;
    uint64_t rays = 0;
    for (uint32_t py = J->row0; py < J->row1; ++py) {
        for (uint32_t px = 0; px < J->w; ++px) {
            const uint32_t pix = py * J->w + px;
            double sum = 0.0;
            for (uint32_t s = J->first_sample; s < J->first_sample + J->spp; ++s) {
                orc_pcg32 rng = orc_pcg_seed(orc_sample_seed(J->seed, J->frame, s), pix);
                double jx = 2.0 * pcg_unit(&rng) - 1.0, jy = 2.0 * pcg_unit(&rng) - 1.0;
                double x = ((2.0 * px + 1.0 + jx) - fw) / fw;
                double y = -(((2.0 * py + 1.0 + jy) - fh) / fh);
                double sx = x * cam[12], sy = y * cam[13];
                double d[3], o[3];
                for (int k = 0; k < 3; ++k) {
                    d[k] = (double)cam[3 + k] + sx * cam[6 + k] + sy * cam[9 + k];
                    o[k] = cam[k];
                }
                double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                for (int k = 0; k < 3; ++k) d[k] /= len;
                double thr = 1.0, L = 0.0;
                int32_t self = -1;
                for (uint32_t b = 0; b <= J->max_bounces; ++b) {
                    double t;
                    int32_t tri = trace_d(J->nodes, J->ids, J->xyz, J->idx, o, d, self, &t);
                    rays++;
                    if (tri < 0) {
                        L += thr * J->sky;
                        break;
                    }
                    if (b == J->max_bounces) break;
                    const float *A = J->xyz + 3 * (size_t)J->idx[3 * (size_t)tri + 0];
                    const float *B = J->xyz + 3 * (size_t)J->idx[3 * (size_t)tri + 1];
                    const float *C = J->xyz + 3 * (size_t)J->idx[3 * (size_t)tri + 2];
                    double e1[3], e2[3], n[3];
                    for (int k = 0; k < 3; ++k) {
                        e1[k] = (double)B[k] - (double)A[k];
                        e2[k] = (double)C[k] - (double)A[k];
                    }
                    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
                    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
                    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
                    double nl = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                    for (int k = 0; k < 3; ++k) n[k] /= nl;
                    if (n[0] * d[0] + n[1] * d[1] + n[2] * d[2] > 0.0)
                        for (int k = 0; k < 3; ++k) n[k] = -n[k];
                    for (int k = 0; k < 3; ++k) o[k] = o[k] + t * d[k];
                    double r1 = pcg_unit(&rng), r2 = pcg_unit(&rng);
                    double phi = 6.283185307179586476925 * r1, rr = sqrt(r2);
                    double lx = rr * cos(phi), ly = rr * sin(phi), lz = sqrt(fmax(0.0, 1.0 - r2));
                    double sign = n[2] >= 0.0 ? 1.0 : -1.0;
                    double a = -1.0 / (sign + n[2]);
                    double bb = n[0] * n[1] * a;
                    double T[3] = {1.0 + sign * n[0] * n[0] * a, sign * bb, -sign * n[0]};
                    double Bv[3] = {bb, sign + n[1] * n[1] * a, -n[1]};
                    for (int k = 0; k < 3; ++k) d[k] = lx * T[k] + ly * Bv[k] + lz * n[k];
                    thr *= J->albedo;
                    self = tri;
                }
                sum += L;
            }
            J->out[pix] = sum;
        }
    }
    J->rays = rays;
    return NULL;
}

uint64_t orc_render_gi(const orc_node *nodes, const int32_t *ids, const float *xyz,
                       const uint32_t *idx, const float cam[14], uint32_t w, uint32_t h,
                       uint32_t spp, uint32_t max_bounces, double albedo, double sky,
                       uint64_t seed, uint32_t frame, uint32_t first_sample, double *out,
                       int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((uint32_t)nthreads > h) nthreads = h ? (int)h : 1;
    pthread_t th[256];
    gi_job jobs[256];
    for (int t = 0; t < nthreads; ++t) {
        gi_job j = {nodes, ids, xyz, idx, cam, w, h, spp, max_bounces, frame, first_sample,
                    albedo, sky, seed, out, 0,
                    (uint32_t)((uint64_t)h * t / nthreads), (uint32_t)((uint64_t)h * (t + 1) / nthreads)};
        jobs[t] = j;
    }
    if (nthreads == 1) {
        gi_worker(&jobs[0]);
        return jobs[0].rays;
    }
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, gi_worker, &jobs[t]);
    uint64_t rays = 0;
    for (int t = 0; t < nthreads; ++t) {
        pthread_join(th[t], NULL);
        rays += jobs[t].rays;
    }
    return rays;
}
