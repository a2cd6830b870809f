// Seeded synthetic scenes (host only).  The reference loads OBJ models
// (Common/Scenes/Model.cpp:210-381) that are not shipped; these generators
// produce the scenes the workloads are defined on.  All randomness comes from
// PCG32 (O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically
// Good Algorithms for Random Number Generation", 2014), so every scene is a
// pure function of its parameters and seed.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "rtcore.h"

namespace {

struct Pcg32 {
    uint64_t state, inc;
    Pcg32(uint64_t seed, uint64_t stream)
    {
        state = 0u;
        inc = (stream << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next()
    {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31));
    }
    // uniform in [0, 1) with 53 random bits
    double uniform()
    {
        const uint64_t hi = next() >> 5, lo = next() >> 6;
        return ((double)hi * 67108864.0 + (double)lo) * (1.0 / 9007199254740992.0);
    }
    double normal()  // Box-Muller, one value per call
    {
        double u1 = uniform();
        const double u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925 * u2);
    }
};

constexpr double kPi = 3.14159265358979323846;

// Value noise on the integer lattice: a hash of the cell corner gives a value
// in [-1, 1]; trilinear interpolation with the smoothstep fade.
double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed)
{
    uint64_t h = seed ^ 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)x * 0xBF58476D1CE4E5B9ull;
    h = (h ^ (h >> 31)) * 0x94D049BB133111EBull;
    h ^= (uint64_t)y * 0xD6E8FEB86659FD93ull;
    h = (h ^ (h >> 29)) * 0xBF58476D1CE4E5B9ull;
    h ^= (uint64_t)z * 0x9E3779B97F4A7C15ull;
    h = (h ^ (h >> 32)) * 0x94D049BB133111EBull;
    h ^= h >> 31;
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

double value_noise(double x, double y, double z, uint64_t seed)
{
    const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    auto fade = [](double t) { return t * t * (3.0 - 2.0 * t); };
    const double tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
    const double x00 = lerp(c[0][0][0], c[1][0][0], tx), x10 = lerp(c[0][1][0], c[1][1][0], tx);
    const double x01 = lerp(c[0][0][1], c[1][0][1], tx), x11 = lerp(c[0][1][1], c[1][1][1], tx);
    return lerp(lerp(x00, x10, ty), lerp(x01, x11, ty), tz);
}

struct Mesh {
    std::vector<float> xyz;
    std::vector<uint32_t> idx;
    void vert(double x, double y, double z)
    {
        xyz.push_back((float)x);
        xyz.push_back((float)y);
        xyz.push_back((float)z);
    }
    void tri(uint32_t a, uint32_t b, uint32_t c)
    {
        idx.push_back(a);
        idx.push_back(b);
        idx.push_back(c);
    }
    uint32_t nv() const { return (uint32_t)(xyz.size() / 3); }
};

int emit(const Mesh &m, float *xyz, uint32_t *nv, uint32_t *idx, uint32_t *nt)
{
    if (!nv || !nt) return RT_ERR_ARG;
    const uint32_t v = m.nv(), t = (uint32_t)(m.idx.size() / 3);
    if (xyz && idx) {
        if (*nv < v || *nt < t) {
            *nv = v;
            *nt = t;
            return RT_ERR_CAPACITY;
        }
        memcpy(xyz, m.xyz.data(), sizeof(float) * m.xyz.size());
        memcpy(idx, m.idx.data(), sizeof(uint32_t) * m.idx.size());
    }
    *nv = v;
    *nt = t;
    return RT_OK;
}

// UV sphere with nlat rings and nlon segments; poles are single vertices, so
// the polar rows are triangle fans and every triangle has area > 0.
void add_sphere(Mesh &m, double cx, double cy, double cz, double r, uint32_t nlat, uint32_t nlon)
{
    const uint32_t base = m.nv();
    m.vert(cx, cy + r, cz);
    for (uint32_t i = 1; i < nlat; ++i) {
        const double th = kPi * i / nlat;
        for (uint32_t j = 0; j < nlon; ++j) {
            const double ph = 2.0 * kPi * j / nlon;
            m.vert(cx + r * std::sin(th) * std::cos(ph), cy + r * std::cos(th),
                   cz + r * std::sin(th) * std::sin(ph));
        }
    }
    m.vert(cx, cy - r, cz);
    const uint32_t south = m.nv() - 1;
    auto ring = [&](uint32_t i, uint32_t j) { return base + 1 + (i - 1) * nlon + (j % nlon); };
    for (uint32_t j = 0; j < nlon; ++j) m.tri(base, ring(1, j + 1), ring(1, j));
    for (uint32_t i = 1; i + 1 < nlat; ++i)
        for (uint32_t j = 0; j < nlon; ++j) {
            m.tri(ring(i, j), ring(i, j + 1), ring(i + 1, j + 1));
            m.tri(ring(i, j), ring(i + 1, j + 1), ring(i + 1, j));
        }
    for (uint32_t j = 0; j < nlon; ++j) m.tri(south, ring(nlat - 1, j), ring(nlat - 1, j + 1));
}

// Axis-aligned unit cube (12 triangles) with its minimum corner at (x, y, z).
void add_cube(Mesh &m, double x, double y, double z)
{
    const uint32_t b = m.nv();
    for (int k = 0; k < 8; ++k) m.vert(x + (k & 1), y + ((k >> 1) & 1), z + ((k >> 2) & 1));
    static const uint32_t f[12][3] = {{0, 2, 1}, {1, 2, 3}, {4, 5, 6}, {5, 7, 6},
                                      {0, 1, 4}, {1, 5, 4}, {2, 6, 3}, {3, 6, 7},
                                      {0, 4, 2}, {2, 4, 6}, {1, 3, 5}, {3, 7, 5}};
    for (auto &t : f) m.tri(b + t[0], b + t[1], b + t[2]);
}

}  // namespace

extern "C" {

// Torus in the XZ plane around the origin: nu segments around the major
// circle, nv_ring around the tube, 2 triangles per quad; every vertex
// coordinate jittered by U(-jitter, jitter).
int rt_gen_torus(float R, float r, uint32_t nu, uint32_t nv_ring, float jitter, uint64_t seed,
                 float *xyz, uint32_t *nv, uint32_t *idx, uint32_t *nt)
{
    if (nu < 3 || nv_ring < 3) return RT_ERR_ARG;
    Mesh m;
    Pcg32 rng(seed, 1);
    for (uint32_t i = 0; i < nu; ++i) {
        const double th = 2.0 * kPi * i / nu;
        for (uint32_t j = 0; j < nv_ring; ++j) {
            const double ph = 2.0 * kPi * j / nv_ring;
            const double rr = R + r * std::cos(ph);
            double p[3] = {rr * std::cos(th), r * std::sin(ph), rr * std::sin(th)};
            for (double &c : p) c += (2.0 * rng.uniform() - 1.0) * jitter;
            m.vert(p[0], p[1], p[2]);
        }
    }
    auto at = [&](uint32_t i, uint32_t j) { return (i % nu) * nv_ring + (j % nv_ring); };
    for (uint32_t i = 0; i < nu; ++i)
        for (uint32_t j = 0; j < nv_ring; ++j) {
            m.tri(at(i, j), at(i + 1, j), at(i + 1, j + 1));
            m.tri(at(i, j), at(i + 1, j + 1), at(i, j + 1));
        }
    return emit(m, xyz, nv, idx, nt);
}

// ntri independent triangles: centroid c ~ U[0,1]^3, edges e1, e2 ~ N(0, sigma^2 I),
// v0 = c - (e1 + e2)/3, v1 = v0 + e1, v2 = v0 + e2 (computed in double).
int rt_gen_triangle_soup(uint32_t ntri, float sigma, uint64_t seed, float *xyz, uint32_t *nv,
                         uint32_t *idx, uint32_t *nt)
{
    Mesh m;
    m.xyz.reserve(9ull * ntri);
    m.idx.reserve(3ull * ntri);
    Pcg32 rng(seed, 2);
    for (uint32_t t = 0; t < ntri; ++t) {
        double c[3], e1[3], e2[3];
        for (double &x : c) x = rng.uniform();
        for (double &x : e1) x = rng.normal() * sigma;
        for (double &x : e2) x = rng.normal() * sigma;
        double v0[3];
        for (int k = 0; k < 3; ++k) v0[k] = c[k] - (e1[k] + e2[k]) / 3.0;
        m.vert(v0[0], v0[1], v0[2]);
        m.vert(v0[0] + e1[0], v0[1] + e1[1], v0[2] + e1[2]);
        m.vert(v0[0] + e2[0], v0[1] + e2[1], v0[2] + e2[2]);
        m.tri(3 * t, 3 * t + 1, 3 * t + 2);
    }
    return emit(m, xyz, nv, idx, nt);
}

// Unit sphere, (nlat - 2) * nlon * 2 + 2 * nlon triangles, radius displaced by
// amplitude * (3-octave value noise at frequencies 4, 8, 16 with weights 4/7,
// 2/7, 1/7, so the displacement stays within +-amplitude).
int rt_gen_noisy_sphere(uint32_t nlat, uint32_t nlon, float amplitude, uint64_t seed, float *xyz,
                        uint32_t *nv, uint32_t *idx, uint32_t *nt)
{
    if (nlat < 3 || nlon < 3) return RT_ERR_ARG;
    Mesh m;
    add_sphere(m, 0.0, 0.0, 0.0, 1.0, nlat, nlon);
    for (size_t v = 0; v < m.xyz.size() / 3; ++v) {
        const double x = m.xyz[3 * v], y = m.xyz[3 * v + 1], z = m.xyz[3 * v + 2];
        const double n = (4.0 * value_noise(4 * x, 4 * y, 4 * z, seed) +
                          2.0 * value_noise(8 * x, 8 * y, 8 * z, seed + 1) +
                          1.0 * value_noise(16 * x, 16 * y, 16 * z, seed + 2)) /
                         7.0;
        const double len = std::sqrt(x * x + y * y + z * z);
        const double s = (1.0 + amplitude * n) / len;
        m.xyz[3 * v] = (float)(x * s);
        m.xyz[3 * v + 1] = (float)(y * s);
        m.xyz[3 * v + 2] = (float)(z * s);
    }
    return emit(m, xyz, nv, idx, nt);
}

// A random mix of unit cubes (12 triangles) and UV spheres (radius 0.5,
// 16 x 32 -> 960 triangles) inside a box of side ~ (target/200)^(1/3) * 4,
// added until the triangle count reaches target_tris; the last object is
// clipped to hit the count exactly.
int rt_gen_object_mix(uint32_t target_tris, uint64_t seed, float *xyz, uint32_t *nv, uint32_t *idx,
                      uint32_t *nt)
{
    Mesh m;
    Pcg32 rng(seed, 3);
    const double side = 4.0 * std::cbrt(std::max(1.0, target_tris / 200.0));
    while (m.idx.size() / 3 < target_tris) {
        const double x = rng.uniform() * side, y = rng.uniform() * side, z = rng.uniform() * side;
        if (rng.uniform() < 0.5) add_cube(m, x, y, z);
        else add_sphere(m, x + 0.5, y + 0.5, z + 0.5, 0.5, 16, 32);
    }
    m.idx.resize(3ull * target_tris);
    return emit(m, xyz, nv, idx, nt);
}

}  // extern "C"
