// Scene files: the reference's two model formats, loaded to the vertex and
// index arrays rt_scene_upload takes, with the reference's normalisation.
//
// Model::LoadFile (Common/Scenes/Model.cpp:210-381) reads
//   - ".txt": vertex count, triangle count, per vertex "x y z nx ny nz u v",
//     per triangle "a b c" (:214-265);
//   - ".obj": through ModelOBJ (Common/Scenes/Loaders/model_obj.cpp), which
//     fan-triangulates faces of the forms v, v/vt, v//vn, v/vt/vn in file
//     order (:922-1116), gives each distinct (position index, texture
//     coordinate, normal) its own vertex in order of first use (addVertex,
//     :451-493), then normalize() centres the box and divides by its largest
//     extent (:228-256, bounds :77-128);
// and Model::Scale (Model.cpp:647-709) then fits every model into a sphere
// of radius 0.49: centre the box and scale its largest extent to 1.8, then
// scale by 0.49 / max |p|.  Float arithmetic follows those functions'
// operations and order.
//
// Two reference behaviours are not reproduced: a negative OBJ index -k here
// is the k-th most recent vertex (the OBJ convention; the reference maps it
// one further back), and the bounds start from the lowest float (the
// reference starts its maxima at the smallest positive float, which breaks
// models lying entirely below zero on an axis).  Materials, normals and
// texture coordinates only steer vertex numbering; the core uses positions.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "rtcore.h"

namespace {

struct Loaded {
    std::vector<float> xyz;
    std::vector<uint32_t> idx;
};

bool ends_with_ci(const std::string &s, const char *suffix)
{
    const size_t n = strlen(suffix);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; ++i) {
        char a = s[s.size() - n + i], b = suffix[i];
        if (a >= 'A' && a <= 'Z') a = (char)(a - 'A' + 'a');
        if (a != b) return false;
    }
    return true;
}

// ModelOBJ::normalize(1, true): p = (p + (-centre)) * (1 / largest extent).
void obj_normalize(std::vector<float> &xyz)
{
    const size_t nv = xyz.size() / 3;
    if (nv == 0) return;
    float lo[3], hi[3];
    for (int c = 0; c < 3; ++c) {
        lo[c] = std::numeric_limits<float>::max();
        hi[c] = std::numeric_limits<float>::lowest();
    }
    for (size_t i = 0; i < nv; ++i)
        for (int c = 0; c < 3; ++c) {
            const float x = xyz[3 * i + c];
            if (x < lo[c]) lo[c] = x;
            if (x > hi[c]) hi[c] = x;
        }
    float centre[3], ext[3];
    for (int c = 0; c < 3; ++c) {
        centre[c] = (lo[c] + hi[c]) / 2.0f;
        ext[c] = hi[c] - lo[c];
    }
    const float radius = std::max(std::max(ext[0], ext[1]), ext[2]);
    const float factor = 1.0f / radius;
    for (size_t i = 0; i < nv; ++i)
        for (int c = 0; c < 3; ++c) {
            float p = xyz[3 * i + c];
            p += -centre[c];
            p *= factor;
            xyz[3 * i + c] = p;
        }
}

// Model::Scale: fit the largest extent to 1.8 about the box centre, then
// scale so that the farthest vertex is at 0.49.
void model_scale(std::vector<float> &xyz)
{
    const size_t nv = xyz.size() / 3;
    if (nv == 0) return;
    float mn[3] = {xyz[0], xyz[1], xyz[2]}, mx[3] = {xyz[0], xyz[1], xyz[2]};
    for (size_t i = 1; i < nv; ++i)
        for (int c = 0; c < 3; ++c) {
            const float x = xyz[3 * i + c];
            if (x < mn[c]) mn[c] = x;
            if (x > mx[c]) mx[c] = x;
        }
    for (int c = 0; c < 3; ++c) mx[c] = mx[c] - mn[c];
    float scale = 1.8f / std::max(mx[0], std::max(mx[1], mx[2]));
    for (int c = 0; c < 3; ++c) mn[c] = mn[c] + 0.5f * mx[c];
    for (size_t i = 0; i < nv; ++i)
        for (int c = 0; c < 3; ++c) {
            float p = xyz[3 * i + c] - mn[c];
            xyz[3 * i + c] = scale * p;
        }
    scale = 0.0f;
    for (size_t i = 0; i < nv; ++i) {
        const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        const float lng = (x * x + y * y) + z * z;
        if (lng > scale) scale = lng;
    }
    scale = std::sqrt(scale);
    scale = 0.49f / scale;
    for (size_t i = 0; i < 3 * nv; ++i) xyz[i] *= scale;
}

bool parse_index(const char *s, int &out)
{
    char *end = nullptr;
    errno = 0;
    const long v = strtol(s, &end, 10);
    if (end == s || errno != 0 || v == 0 || v > INT32_MAX || v < INT32_MIN) return false;
    out = (int)v;
    return true;
}

// One face corner "v", "v/vt", "v//vn" or "v/vt/vn" -> 0-based indices
// (-1 when absent).  Relative (negative) indices count back from the
// elements read so far.
bool parse_corner(const char *tok, int nv, int nvt, int nvn, int &v, int &vt, int &vn)
{
    v = vt = vn = -1;
    const char *p[3] = {tok, nullptr, nullptr};
    const char *s1 = strchr(tok, '/');
    if (s1) {
        p[1] = s1 + 1;
        const char *s2 = strchr(s1 + 1, '/');
        if (s2) p[2] = s2 + 1;
    }
    const int counts[3] = {nv, nvt, nvn};
    int *outs[3] = {&v, &vt, &vn};
    for (int k = 0; k < 3; ++k) {
        if (!p[k] || *p[k] == '/' || *p[k] == '\0') {
            if (k == 0) return false;
            continue;
        }
        int x;
        if (!parse_index(p[k], x)) return false;
        x = x < 0 ? counts[k] + x : x - 1;
        if (x < 0 || x >= counts[k]) return false;
        *outs[k] = x;
    }
    return true;
}

int load_obj(FILE *f, Loaded &out, std::string &err)
{
    std::vector<float> pos, tex, nrm;
    std::map<std::tuple<int, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>, uint32_t> cache;
    char *line = nullptr;
    size_t cap = 0;
    long lineno = 0;
    struct Free {
        char **p;
        ~Free() { free(*p); }
    } guard{&line};
    while (getline(&line, &cap, f) >= 0) {
        ++lineno;
        char *save = nullptr;
        char *tok = strtok_r(line, " \t\r\n", &save);
        if (!tok || tok[0] == '#') continue;
        if (strcmp(tok, "v") == 0 || strcmp(tok, "vn") == 0 || strcmp(tok, "vt") == 0) {
            const int n = tok[1] == 't' ? 2 : 3;
            std::vector<float> &dst = tok[1] == '\0' ? pos : (tok[1] == 'n' ? nrm : tex);
            for (int k = 0; k < n; ++k) {
                char *t = strtok_r(nullptr, " \t\r\n", &save);
                char *end = nullptr;
                const float x = t ? strtof(t, &end) : 0.0f;
                if (!t || end == t) {
                    err = "line " + std::to_string(lineno) + ": bad " + std::string(tok) + " record";
                    return RT_ERR_ARG;
                }
                dst.push_back(x);
            }
        } else if (strcmp(tok, "f") == 0) {
            std::vector<uint32_t> corners;
            for (char *t = strtok_r(nullptr, " \t\r\n", &save); t; t = strtok_r(nullptr, " \t\r\n", &save)) {
                int v, vt, vn;
                if (!parse_corner(t, (int)(pos.size() / 3), (int)(tex.size() / 2), (int)(nrm.size() / 3), v, vt,
                                  vn)) {
                    err = "line " + std::to_string(lineno) + ": bad face corner '" + t + "'";
                    return RT_ERR_ARG;
                }
                // the vertex identity of ModelOBJ::addVertex: position index
                // plus texture coordinate and normal values (zero if absent)
                float tc[2] = {0.0f, 0.0f}, nn[3] = {0.0f, 0.0f, 0.0f};
                if (vt >= 0) memcpy(tc, &tex[2 * vt], sizeof(tc));
                if (vn >= 0) memcpy(nn, &nrm[3 * vn], sizeof(nn));
                uint32_t b[5];
                memcpy(b, tc, 8);
                memcpy(b + 2, nn, 12);
                const auto key = std::make_tuple(v, b[0], b[1], b[2], b[3], b[4]);
                auto it = cache.find(key);
                uint32_t id;
                if (it == cache.end()) {
                    id = (uint32_t)(out.xyz.size() / 3);
                    out.xyz.insert(out.xyz.end(), &pos[3 * v], &pos[3 * v] + 3);
                    cache.emplace(key, id);
                } else {
                    id = it->second;
                }
                corners.push_back(id);
            }
            if (corners.size() < 3) {
                err = "line " + std::to_string(lineno) + ": face with fewer than 3 corners";
                return RT_ERR_ARG;
            }
            for (size_t k = 2; k < corners.size(); ++k) {  // fan: (0, k-1, k)
                out.idx.push_back(corners[0]);
                out.idx.push_back(corners[k - 1]);
                out.idx.push_back(corners[k]);
            }
        }
        // mtllib, usemtl, o, g, s and the rest carry no geometry
    }
    return RT_OK;
}

int load_txt(FILE *f, Loaded &out, std::string &err)
{
    long nv = 0, nt = 0;
    if (fscanf(f, "%ld %ld", &nv, &nt) != 2 || nv <= 0 || nt <= 0 || nv > (1L << 27) || nt > (1L << 26)) {
        err = "bad header (vertex and triangle counts)";
        return RT_ERR_ARG;
    }
    out.xyz.resize(3 * (size_t)nv);
    for (long i = 0; i < nv; ++i) {
        float r[8];
        if (fscanf(f, "%f %f %f %f %f %f %f %f", &r[0], &r[1], &r[2], &r[3], &r[4], &r[5], &r[6], &r[7]) != 8) {
            err = "vertex " + std::to_string(i) + ": expected 8 numbers";
            return RT_ERR_ARG;
        }
        memcpy(&out.xyz[3 * (size_t)i], r, 12);
    }
    out.idx.resize(3 * (size_t)nt);
    for (long i = 0; i < nt; ++i) {
        long a, b, c;
        if (fscanf(f, "%ld %ld %ld", &a, &b, &c) != 3 || a < 0 || b < 0 || c < 0 || a >= nv || b >= nv || c >= nv) {
            err = "triangle " + std::to_string(i) + ": expected 3 vertex indices in range";
            return RT_ERR_ARG;
        }
        out.idx[3 * i] = (uint32_t)a;
        out.idx[3 * i + 1] = (uint32_t)b;
        out.idx[3 * i + 2] = (uint32_t)c;
    }
    return RT_OK;
}

thread_local std::string g_load_err;

}  // namespace

extern "C" {

int rt_load_model(const char *path, uint32_t flags, float *xyz, uint32_t *nv, uint32_t *idx, uint32_t *nt)
{
    if (!path || !nv || !nt) return RT_ERR_ARG;
    g_load_err.clear();
    const std::string p(path);
    const bool obj = ends_with_ci(p, ".obj"), txt = ends_with_ci(p, ".txt");
    if (!obj && !txt) {
        g_load_err = "rt_load_model: the file name must end in .obj or .txt";
        return RT_ERR_ARG;
    }
    FILE *f = fopen(path, "r");
    if (!f) {
        g_load_err = "rt_load_model: cannot open " + p;
        return RT_ERR_ARG;
    }
    Loaded m;
    std::string err;
    const int rc = obj ? load_obj(f, m, err) : load_txt(f, m, err);
    fclose(f);
    if (rc != RT_OK) {
        g_load_err = "rt_load_model: " + p + ": " + err;
        return rc;
    }
    if (m.idx.empty() || m.xyz.size() / 3 > 0xFFFFFFFFull || m.idx.size() / 3 > 0xFFFFFFFFull) {
        g_load_err = "rt_load_model: " + p + ": no triangles";
        return RT_ERR_ARG;
    }
    if (!(flags & RT_LOAD_RAW)) {
        if (obj) obj_normalize(m.xyz);
        model_scale(m.xyz);
    }
    const uint32_t cv = (uint32_t)(m.xyz.size() / 3), ct = (uint32_t)(m.idx.size() / 3);
    if (xyz || idx) {
        if (!xyz || !idx || *nv < cv || *nt < ct) {
            g_load_err = "rt_load_model: output arrays too small";
            return RT_ERR_CAPACITY;
        }
        memcpy(xyz, m.xyz.data(), m.xyz.size() * sizeof(float));
        memcpy(idx, m.idx.data(), m.idx.size() * sizeof(uint32_t));
    }
    *nv = cv;
    *nt = ct;
    return RT_OK;
}

const char *rt_load_last_error(void) { return g_load_err.c_str(); }

}  // extern "C"
