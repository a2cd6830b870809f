// Sanitizer driver for the library's host-only code (TEST INFRASTRUCTURE
// ONLY): scene generators, the OBJ / MTL / .txt loaders, the SAH BVH host
// build, PNG / PPM writing and the render server in dry-run mode, linked
// against the library built host-only with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/sanitize/Makefile).  No GPU is touched.
// Exit 0 when every call behaved as documented.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rtcore.h"

#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            std::exit(1);                                                        \
        }                                                                        \
    } while (0)

struct Mesh {
    std::vector<float> xyz;
    std::vector<uint32_t> idx;
};

template <class F>
static Mesh gen(F f)
{
    uint32_t nv = 0, nt = 0;
    CHECK(f(nullptr, &nv, nullptr, &nt) == RT_OK);
    Mesh m;
    m.xyz.resize(3ull * nv);
    m.idx.resize(3ull * nt);
    CHECK(f(m.xyz.data(), &nv, m.idx.data(), &nt) == RT_OK);
    for (uint32_t i : m.idx) CHECK(i < nv);
    return m;
}

static void write_file(const std::string &path, const char *text)
{
    FILE *f = std::fopen(path.c_str(), "w");
    CHECK(f);
    std::fputs(text, f);
    std::fclose(f);
}

static void loaders(const std::string &dir)
{
    write_file(dir + "/m.mtl",
               "newmtl red\nKd 0.8 0.1 0.1\nKa 0.1 0.1 0.1\nKs 0.5 0.5 0.5\nNs 32\nd 0.5\nNi 1.4\n"
               "newmtl glass\nKd 0.9 0.9 0.9\nd 0.2\nillum 7\n");
    write_file(dir + "/m.obj",
               "mtllib m.mtl\n# a quad, a pentagon fan, negative indices, normals and texcoords\n"
               "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 1.5 0.2\nvt 0 0\nvt 1 0\nvn 0 0 1\n"
               "usemtl red\nf 1/1/1 2/2/1 3//1 4//1\nusemtl glass\nf -5 -4 -3 -2 -1\nusemtl nothing\nf 1 3 5\n"
               "s off\ng grp\no obj\nbogus line\n");
    write_file(dir + "/bad.obj", "v 0 0 0\nf 1 2 3\n");
    write_file(dir + "/m.txt", "3 1\n0 0 0 0 0 1 0 0\n1 0 0 0 0 1 1 0\n0 1 0 0 0 2 0 1\n0 1 2\n");
    write_file(dir + "/short.txt", "3 1\n0 0 0 0 0 1\n");
    for (const char *name : {"m.obj", "m.txt"}) {
        for (uint32_t flags : {0u, (uint32_t)RT_LOAD_RAW}) {
            uint32_t nv = 0, nt = 0;
            const std::string p = dir + "/" + name;
            CHECK(rt_load_model(p.c_str(), flags, nullptr, &nv, nullptr, &nt) == RT_OK);
            std::vector<float> xyz(3ull * nv);
            std::vector<uint32_t> idx(3ull * nt);
            CHECK(rt_load_model(p.c_str(), flags, xyz.data(), &nv, idx.data(), &nt) == RT_OK);
            rt_model m{};
            CHECK(rt_load_model_ex(p.c_str(), flags, &m) == RT_OK);
            std::vector<float> x2(3ull * m.nv), n2(3ull * m.nv);
            std::vector<uint32_t> i2(3ull * m.nt), mid(m.nt);
            std::vector<rt_material> mats(m.nmat);
            m.xyz = x2.data();
            m.normals = n2.data();
            m.idx = i2.data();
            m.mat_ids = mid.data();
            m.materials = mats.data();
            CHECK(rt_load_model_ex(p.c_str(), flags, &m) == RT_OK);
            CHECK(x2 == xyz && i2 == idx);
            for (uint32_t k : mid) CHECK(k < m.nmat);
        }
    }
    uint32_t nv = 0, nt = 0;
    CHECK(rt_load_model((dir + "/bad.obj").c_str(), 0, nullptr, &nv, nullptr, &nt) != RT_OK);
    CHECK(std::strlen(rt_load_last_error()) > 0);
    CHECK(rt_load_model((dir + "/missing.obj").c_str(), 0, nullptr, &nv, nullptr, &nt) != RT_OK);
    CHECK(rt_load_model((dir + "/short.txt").c_str(), 0, nullptr, &nv, nullptr, &nt) != RT_OK);
}

static void bvh(const Mesh &m)
{
    const uint32_t nt = (uint32_t)(m.idx.size() / 3), nv = (uint32_t)(m.xyz.size() / 3);
    for (uint32_t split : {0u, 1u, 2u}) {
        for (uint32_t mp : {1u, 2u, 4u, 255u}) {
            std::vector<rt_node32> nodes(2ull * nt - 1);
            std::vector<int32_t> ord(nt);
            rt_bvh_info info{};
            CHECK(rt_bvh_build_host(m.xyz.data(), nv, m.idx.data(), nt, mp, split, nodes.data(), ord.data(),
                                    &info) == RT_OK);
            std::vector<int> seen(nt, 0);
            for (int32_t o : ord) seen.at((size_t)o)++;
            for (int s : seen) CHECK(s == 1);
        }
    }
    std::vector<rt_node32> nodes(1);
    CHECK(rt_bvh_build_host(m.xyz.data(), nv, m.idx.data(), nt, 0, 0, nodes.data(), nullptr, nullptr) != RT_OK);
}

static void images(const std::string &dir)
{
    std::vector<uint8_t> px(4 * 37 * 21);
    for (size_t i = 0; i < px.size(); ++i) px[i] = (uint8_t)(i * 7);
    CHECK(rt_image_write((dir + "/a.png").c_str(), px.data(), 37, 21) == RT_OK);
    CHECK(rt_image_write((dir + "/a.ppm").c_str(), px.data(), 37, 21) == RT_OK);
    CHECK(rt_image_write((dir + "/a.xyz").c_str(), px.data(), 37, 21) != RT_OK);
}

// Bytes the dry-run server accounts for scene `m` (depth 14) plus one 64 x 48 stream.
static uint64_t probe_bytes(const Mesh &m)
{
    rt_server_config cfg{};
    cfg.hbm_budget = 1ull << 40;
    cfg.max_inflight = 1;
    cfg.tiles_per_step = 1 << 20;
    cfg.dry_run = 1;
    cfg.dry_rate = 1e9;
    rt_server *s = nullptr;
    CHECK(rt_server_create(nullptr, &cfg, &s) == RT_OK);
    CHECK(rt_server_add_scene(s, 9, m.xyz.data(), (uint32_t)m.xyz.size() / 3, m.idx.data(),
                              (uint32_t)m.idx.size() / 3, 14) == RT_OK);
    rt_request q{};
    q.scene_key = 9;
    q.width = 64;
    q.height = 48;
    q.mode = RT_MODE_GI;
    q.spp = 1;
    CHECK(rt_server_submit(s, &q, nullptr) == RT_OK);
    double dt = 0.0;
    CHECK(rt_server_step(s, 0.0, &dt) == RT_OK);
    rt_server_stats st{};
    CHECK(rt_server_status(s, &st) == RT_OK);
    rt_server_destroy(s);
    return st.bytes_used;
}

static void server(const Mesh &a, const Mesh &b)
{
    for (double window : {0.0, 0.002}) {
        rt_server_config cfg{};
        // room for the larger scene and a few streams: scenes and accumulators are evicted
        cfg.hbm_budget = probe_bytes(b) + 3 * 16 * 64 * 48;
        cfg.max_inflight = 4;
        cfg.tiles_per_step = 64;
        cfg.dry_run = 1;
        cfg.dry_rate = 1e9;
        cfg.tile_order = RT_TILES_CENTRE;
        cfg.batch_window = window;
        rt_server *s = nullptr;
        CHECK(rt_server_create(nullptr, &cfg, &s) == RT_OK);
        CHECK(rt_server_add_scene(s, 1, a.xyz.data(), (uint32_t)a.xyz.size() / 3, a.idx.data(),
                                  (uint32_t)a.idx.size() / 3, 12) == RT_OK);
        CHECK(rt_server_add_scene(s, 2, b.xyz.data(), (uint32_t)b.xyz.size() / 3, b.idx.data(),
                                  (uint32_t)b.idx.size() / 3, 14) == RT_OK);
        CHECK(rt_server_add_scene(s, 2, b.xyz.data(), 3, b.idx.data(), 1, 14) != RT_OK);  // key taken
        const float light[3] = {0.0f, 3.0f, 0.0f};
        rt_camera cam{};
        const float pos[3] = {0, 0.5f, -3}, tgt[3] = {0, 0, 0}, up[3] = {0, 1, 0};
        rt_camera_look_at(&cam, pos, tgt, up, 60.0f, 1.0f);
        const uint32_t N = 40;
        for (uint32_t r = 0; r < N; ++r) {
            rt_request q{};
            q.scene_key = 1 + r % 2;
            q.view = r % 3;
            q.cam = cam;
            q.width = 32 + 8 * (r % 4);
            q.height = 24 + 8 * (r % 3);
            q.mode = r % 2 ? RT_MODE_GI : RT_MODE_DIRECT;
            q.spp = 1 + r % 2;
            q.max_bounces = 2;
            q.albedo = 0.7f;
            q.sky = 1.0f;
            q.lights = q.mode == RT_MODE_DIRECT ? light : nullptr;
            q.nlights = q.mode == RT_MODE_DIRECT ? 1 : 0;
            q.arrival = 0.0005 * r;
            uint64_t id = 0;
            CHECK(rt_server_submit(s, &q, &id) == RT_OK);
        }
        rt_request big{};
        big.scene_key = 2;
        big.cam = cam;
        big.width = big.height = 4096;  // 256 MiB accumulator: never fits
        big.mode = RT_MODE_GI;
        big.spp = 1;
        big.arrival = 0.0;
        CHECK(rt_server_submit(s, &big, nullptr) == RT_OK);
        double now = 0.0;
        uint32_t got = 0;
        for (int step = 0; step < 100000 && got < N + 1; ++step) {
            double dt = 0.0;
            CHECK(rt_server_step(s, now, &dt) == RT_OK);
            rt_request_result res[8];
            uint32_t n = 0;
            CHECK(rt_server_poll(s, res, 8, &n) == RT_OK);
            got += n;
            rt_server_stats st{};
            CHECK(rt_server_status(s, &st) == RT_OK);
            now += dt > 0.0 ? dt : (st.wake > now ? st.wake - now : 0.0005);
        }
        CHECK(got == N + 1);
        rt_server_stats st{};
        CHECK(rt_server_status(s, &st) == RT_OK);
        if (!(st.completed == N && st.rejected == 1 && st.evictions > 0))
            std::fprintf(stderr, "completed %llu rejected %llu evictions %llu builds %llu\n",
                         (unsigned long long)st.completed, (unsigned long long)st.rejected,
                         (unsigned long long)st.evictions, (unsigned long long)st.builds);
        CHECK(st.completed == N && st.rejected == 1 && st.evictions > 0);
        std::vector<float> img(4 * 48 * 40);
        uint32_t samples = 0;
        rt_server_read_stream(s, 1, 0, 32, 24, RT_MODE_DIRECT, img.data(), &samples);
        rt_server_destroy(s);
    }
    rt_server_config bad{};
    rt_server *s = nullptr;
    CHECK(rt_server_create(nullptr, &bad, &s) != RT_OK);
}

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    const Mesh torus = gen([](float *x, uint32_t *nv, uint32_t *i, uint32_t *nt) {
        return rt_gen_torus(1.0f, 0.3f, 16, 8, 1e-3f, 1, x, nv, i, nt);
    });
    const Mesh soup = gen([](float *x, uint32_t *nv, uint32_t *i, uint32_t *nt) {
        return rt_gen_triangle_soup(2000, 0.01f, 2, x, nv, i, nt);
    });
    const Mesh sphere = gen([](float *x, uint32_t *nv, uint32_t *i, uint32_t *nt) {
        return rt_gen_noisy_sphere(17, 16, 0.05f, 3, x, nv, i, nt);
    });
    const Mesh mix = gen([](float *x, uint32_t *nv, uint32_t *i, uint32_t *nt) {
        return rt_gen_object_mix(3000, 4, x, nv, i, nt);
    });
    loaders(dir);
    bvh(torus);
    bvh(mix);
    bvh(sphere);
    images(dir);
    server(torus, soup);
    rt_camera cam{};
    const float world[16] = {1, 0, 0, 0.5f, 0, 1, 0, 0, 0, 0, 1, -2, 0, 0, 0, 1};
    rt_camera_from_world(&cam, world);
    CHECK(cam.kind == RT_CAMERA_WORLD && cam.pos[0] == 0.5f && cam.pos[2] == -2.0f);
    std::printf("host sanitizer run ok\n");
    return 0;
}
