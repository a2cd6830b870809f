// Scene files: the reference's two model formats, loaded to the vertex and
// index arrays rt_scene_upload takes, with the reference's normalisation, and
// to the normals, material ids and materials rt_scene_shading takes.
//
// Model::LoadFile (Common/Scenes/Model.cpp:210-381) reads
//   - ".txt": vertex count, triangle count, per vertex "x y z nx ny nz u v",
//     per triangle "a b c" (:214-265); normals are normalised (Normalize,
//     Common/Geometry/Geometry.h:268-275) and every triangle gets the default
//     MaterialData (MaterialData.cpp:17-24);
//   - ".obj": through ModelOBJ (Common/Scenes/Loaders/model_obj.cpp), which
//     fan-triangulates faces of the forms v, v/vt, v//vn, v/vt/vn in file
//     order (:922-1116), gives each distinct (position index, texture
//     coordinate, normal) its own vertex in order of first use (addVertex,
//     :451-493), builds normals from the raw positions when the file has none
//     (generateNormals, :541-616), reads the MTL file named by mtllib
//     (importMaterials, :1119-1297), then normalize() centres the box and
//     divides by its largest extent (:228-256, bounds :77-128).  Model then
//     takes one default MaterialData when the OBJ brought no materials, gives
//     every triangle material 0 when it brought one, and otherwise the
//     material active (usemtl) at the triangle's face (:284-378).
// Model::Scale (Model.cpp:647-709) then fits every model into a sphere of
// radius 0.49: centre the box and scale its largest extent to 1.8, then scale
// by 0.49 / max |p|.  Float arithmetic follows those functions' operations
// and order.
//
// Reference behaviours not reproduced: a negative OBJ index -k here is the
// k-th most recent vertex (the OBJ convention; the reference maps it one
// further back); the bounds start from the lowest float (the reference
// starts its maxima at the smallest positive float, which breaks models lying
// entirely below zero on an axis); of several mtllib lines the last one
// names the materials (the reference merges their name tables); a material
// property before any newmtl is an error (the reference writes through a
// null pointer).  Texture file names are not kept: no map is sampled.
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
    std::vector<float> nrm;  // 3 per vertex
    std::vector<uint32_t> idx;
    std::vector<uint32_t> mat;  // 1 per triangle
    std::vector<rt_material> mats;
};

// MaterialData() (Common/Scenes/MaterialData.cpp:17-24).
rt_material default_material_data()
{
    rt_material m;
    memset(&m, 0, sizeof(m));
    m.diffuse[0] = m.diffuse[1] = m.diffuse[2] = 0.5f;
    m.diffuse[3] = 0.8f;
    m.ambient[3] = 1.0f;
    m.specular[0] = m.specular[1] = m.specular[2] = 0.8f;
    m.specular[3] = 1.0f;
    return m;
}

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

// The name argument of an mtllib / newmtl / usemtl line as the reference's
// sscanf(buffer, "%s %s", buffer, buffer) leaves it: the second of the first
// two words when there are two, else the first.
std::string name_arg(char *rest)
{
    char *save = nullptr;
    char *w0 = strtok_r(rest, " \t\r\n", &save);
    if (!w0) return std::string();
    char *w1 = strtok_r(nullptr, " \t\r\n", &save);
    return std::string(w1 ? w1 : w0);
}

struct MtlTable {
    std::vector<rt_material> mats;
    std::map<std::string, int> cache;  // name -> index (a repeated name maps to its last definition)
};

// ModelOBJ::importMaterials (model_obj.cpp:1119-1297): "newmtl" starts a
// material with ModelOBJ's defaults (ambient 0.2, diffuse 0.8, specular 0,
// each with w = 1, shininess 0, alpha 1); Ka / Kd / Ks read up to four
// numbers and set w = 1 when fewer than four convert; Ns is divided by 1000;
// Tr sets alpha = 1 - value, d sets alpha; illum 1 zeroes specular.rgb and
// sets its w to 1.  A file that cannot be opened brings no materials.
int load_mtl(const std::string &path, MtlTable &t, std::string &err)
{
    t.mats.clear();
    t.cache.clear();
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return RT_OK;
    char *line = nullptr;
    size_t cap = 0;
    long lineno = 0;
    int illum = 0;  // keeps its last value when an illum line does not convert, as in the reference
    rt_material *m = nullptr;
    int rc = RT_OK;
    while (getline(&line, &cap, f) >= 0) {
        ++lineno;
        char *save = nullptr;
        char *tok = strtok_r(line, " \t\r\n", &save);
        if (!tok) continue;
        char *rest = save ? save : tok + strlen(tok);
        const char c0 = tok[0], c1 = tok[1];
        if (c0 == 'n') {  // newmtl
            rt_material d;
            memset(&d, 0, sizeof(d));
            d.ambient[0] = d.ambient[1] = d.ambient[2] = 0.2f;
            d.ambient[3] = 1.0f;
            d.diffuse[0] = d.diffuse[1] = d.diffuse[2] = 0.8f;
            d.diffuse[3] = 1.0f;
            d.specular[3] = 1.0f;
            d.shininess = 0.0f;
            d.alpha = 1.0f;
            d.id = (uint32_t)t.mats.size();
            t.cache[name_arg(rest)] = (int)t.mats.size();
            t.mats.push_back(d);
            m = &t.mats.back();
            continue;
        }
        const bool prop = (c0 == 'N' && c1 == 's') || (c0 == 'K' && (c1 == 'a' || c1 == 'd' || c1 == 's')) ||
                          (c0 == 'T' && c1 == 'r') || c0 == 'd' || c0 == 'i';
        if (!prop) continue;  // comments, maps, Ni, Ke, Tf and the rest
        if (!m) {
            err = path + ": line " + std::to_string(lineno) + ": '" + tok + "' before any newmtl";
            rc = RT_ERR_ARG;
            break;
        }
        if (c0 == 'N') {
            sscanf(rest, "%f", &m->shininess);
            m->shininess /= 1000.0f;
        } else if (c0 == 'K') {
            float *dst = c1 == 'a' ? m->ambient : (c1 == 'd' ? m->diffuse : m->specular);
            if (sscanf(rest, "%f %f %f %f", &dst[0], &dst[1], &dst[2], &dst[3]) < 4) dst[3] = 1.0f;
        } else if (c0 == 'T') {
            sscanf(rest, "%f", &m->alpha);
            m->alpha = 1.0f - m->alpha;
        } else if (c0 == 'd') {
            sscanf(rest, "%f", &m->alpha);
        } else {  // illum
            sscanf(rest, "%d", &illum);
            if (illum == 1) {
                m->specular[0] = m->specular[1] = m->specular[2] = 0.0f;
                m->specular[3] = 1.0f;
            }
        }
    }
    free(line);
    fclose(f);
    return rc;
}

// ModelOBJ::generateNormals: per vertex the sum of its faces' unnormalised
// cross(B - A, C - A), times 1 / length.
void generate_normals(const std::vector<float> &xyz, const std::vector<uint32_t> &idx, std::vector<float> &nrm)
{
    nrm.assign(xyz.size(), 0.0f);
    for (size_t t = 0; t + 2 < idx.size(); t += 3) {
        const float *A = &xyz[3 * (size_t)idx[t]], *B = &xyz[3 * (size_t)idx[t + 1]], *C = &xyz[3 * (size_t)idx[t + 2]];
        const float e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        const float e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
        const float n[3] = {(e1[1] * e2[2]) - (e1[2] * e2[1]), (e1[2] * e2[0]) - (e1[0] * e2[2]),
                            (e1[0] * e2[1]) - (e1[1] * e2[0])};
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) nrm[3 * (size_t)idx[t + k] + c] += n[c];
    }
    for (size_t v = 0; v + 2 < nrm.size(); v += 3) {
        float *n = &nrm[v];
        const float len = 1.0f / sqrtf((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
        n[0] *= len;
        n[1] *= len;
        n[2] *= len;
    }
}

std::string dir_of(const std::string &path)
{
    size_t k = path.find_last_of('\\');
    if (k == std::string::npos) k = path.find_last_of('/');
    return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

int load_obj(const std::string &path, FILE *f, bool want_materials, Loaded &out, std::string &err)
{
    std::vector<float> pos, tex, nrm;
    std::map<std::tuple<int, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>, uint32_t> cache;
    MtlTable mtl;
    char *line = nullptr;
    size_t cap = 0;
    long lineno = 0;
    struct Free {
        char **p;
        ~Free() { free(*p); }
    } guard{&line};
    // ModelOBJ reads the materials in its first pass, so usemtl resolves
    // against them wherever the mtllib line stands.
    if (want_materials) {
        std::string lib;
        bool have = false;
        while (getline(&line, &cap, f) >= 0) {
            char *save = nullptr;
            char *tok = strtok_r(line, " \t\r\n", &save);
            if (tok && strcmp(tok, "mtllib") == 0) {
                lib = name_arg(save ? save : tok + strlen(tok));
                have = true;
            }
        }
        rewind(f);
        if (have) {
            const int rc = load_mtl(dir_of(path) + lib, mtl, err);
            if (rc != RT_OK) return rc;
        }
    }
    int active = 0;
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
        } else if (strcmp(tok, "usemtl") == 0) {
            if (want_materials) {
                const auto it = mtl.cache.find(name_arg(save ? save : tok + strlen(tok)));
                active = it == mtl.cache.end() ? 0 : it->second;
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
                    out.nrm.insert(out.nrm.end(), nn, nn + 3);
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
                out.mat.push_back((uint32_t)active);
            }
        }
        // o, g, s and the rest carry nothing the renderer uses
    }
    if (nrm.empty()) generate_normals(out.xyz, out.idx, out.nrm);
    if (mtl.mats.empty()) {
        out.mats.assign(1, default_material_data());
        std::fill(out.mat.begin(), out.mat.end(), 0u);
    } else {
        out.mats = mtl.mats;
        if (out.mats.size() == 1) std::fill(out.mat.begin(), out.mat.end(), 0u);
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
    out.nrm.resize(3 * (size_t)nv);
    for (long i = 0; i < nv; ++i) {
        float r[8];
        if (fscanf(f, "%f %f %f %f %f %f %f %f", &r[0], &r[1], &r[2], &r[3], &r[4], &r[5], &r[6], &r[7]) != 8) {
            err = "vertex " + std::to_string(i) + ": expected 8 numbers";
            return RT_ERR_ARG;
        }
        memcpy(&out.xyz[3 * (size_t)i], r, 12);
        const float len = sqrtf((r[3] * r[3] + r[4] * r[4]) + r[5] * r[5]);
        out.nrm[3 * (size_t)i] = r[3] / len;
        out.nrm[3 * (size_t)i + 1] = r[4] / len;
        out.nrm[3 * (size_t)i + 2] = r[5] / len;
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
    out.mat.assign((size_t)nt, 0u);
    out.mats.assign(1, default_material_data());
    return RT_OK;
}

thread_local std::string g_load_err;

int load_any(const char *path, uint32_t flags, bool want_materials, Loaded &m)
{
    g_load_err.clear();
    if (!path) {
        g_load_err = "rt_load_model: null path";
        return RT_ERR_ARG;
    }
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
    std::string err;
    const int rc = obj ? load_obj(p, f, want_materials, m, err) : load_txt(f, m, err);
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
    return RT_OK;
}

}  // namespace

extern "C" {

int rt_load_model(const char *path, uint32_t flags, float *xyz, uint32_t *nv, uint32_t *idx, uint32_t *nt)
{
    if (!path || !nv || !nt) return RT_ERR_ARG;
    Loaded m;
    const int rc = load_any(path, flags, false, m);
    if (rc != RT_OK) return rc;
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

int rt_load_model_ex(const char *path, uint32_t flags, rt_model *model)
{
    if (!path || !model) return RT_ERR_ARG;
    Loaded m;
    const int rc = load_any(path, flags, true, m);
    if (rc != RT_OK) return rc;
    const uint32_t cv = (uint32_t)(m.xyz.size() / 3), ct = (uint32_t)(m.idx.size() / 3);
    const uint32_t cm = (uint32_t)m.mats.size();
    if (model->xyz || model->idx) {
        if (!model->xyz || !model->idx || model->nv < cv || model->nt < ct ||
            (model->materials && model->nmat < cm)) {
            g_load_err = "rt_load_model_ex: output arrays too small";
            return RT_ERR_CAPACITY;
        }
        memcpy(model->xyz, m.xyz.data(), m.xyz.size() * sizeof(float));
        memcpy(model->idx, m.idx.data(), m.idx.size() * sizeof(uint32_t));
        if (model->normals) memcpy(model->normals, m.nrm.data(), m.nrm.size() * sizeof(float));
        if (model->mat_ids) memcpy(model->mat_ids, m.mat.data(), m.mat.size() * sizeof(uint32_t));
        if (model->materials) memcpy(model->materials, m.mats.data(), m.mats.size() * sizeof(rt_material));
    }
    model->nv = cv;
    model->nt = ct;
    model->nmat = cm;
    return RT_OK;
}

const char *rt_load_last_error(void) { return g_load_err.c_str(); }

}  // extern "C"
