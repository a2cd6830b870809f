// C++ facades over the rtcore C ABI (see facade.h).
#include "facade.h"

#include <cstring>
#include <string>

namespace rtgpu {

Device::Device(int device)
{
    const int rc = rt_create(&ctx_, device);
    if (rc != RT_OK) throw Error(rc, "rt_create failed on device " + std::to_string(device));
}

Device::~Device() { rt_destroy(ctx_); }

void Device::check(int rc, const char *what) const
{
    if (rc != RT_OK) throw Error(rc, std::string(what) + ": " + rt_last_error(ctx_));
}

static unsigned int log2_exact(int n)
{
    if (n < 4 || (n & (n - 1)) != 0) throw Error(RT_ERR_ARG, "node count must be a power of two >= 4");
    unsigned int d = 0;
    while ((1 << d) < n) ++d;
    return d;
}

LbvhGpu::LbvhGpu(Device &dev, const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt, int n)
    : dev_(dev), depth_(log2_exact(n)), nt_(nt)
{
    dev_.check(rt_scene_create(dev_.ctx(), nt, nv, depth_, &scene_), "rt_scene_create");
    dev_.check(rt_scene_upload(dev_.ctx(), scene_, xyz, nv, idx, nt), "rt_scene_upload");
}

LbvhGpu::~LbvhGpu()
{
    if (scene_ >= 0) rt_scene_destroy(dev_.ctx(), scene_);
}

void LbvhGpu::Build()
{
    dev_.check(rt_lbvh_build(dev_.ctx(), scene_, &times_), "rt_lbvh_build");
    fetched_ = false;
}

void LbvhGpu::fetch()
{
    if (fetched_) return;
    nodes_.resize(size_t(1) << depth_);
    ids_.resize(nt_);
    dev_.check(rt_lbvh_read(dev_.ctx(), scene_, nodes_.data(), nullptr, ids_.data(), nullptr),
               "rt_lbvh_read");
    fetched_ = true;
}

void *LbvhGpu::GetPtrToElements()
{
    fetch();
    return nodes_.data();
}

unsigned int *LbvhGpu::GetOrderedElementsIds()
{
    fetch();
    return reinterpret_cast<unsigned int *>(ids_.data());
}

BBox LbvhGpu::WorldBound()
{
    fetch();
    BBox b{};
    if (nt_ == 0) return b;
    memcpy(b.pMin, nodes_[1].mn, sizeof(b.pMin));
    memcpy(b.pMax, nodes_[1].mx, sizeof(b.pMax));
    return b;
}

void LbvhGpu::IntersectP(const Ray *rays, uint32_t count, TIntersection *out)
{
    std::vector<rt_hit24> hits(count);
    dev_.check(rt_trace(dev_.ctx(), scene_, reinterpret_cast<const rt_ray48 *>(rays), count,
                        hits.data(), 0, nullptr),
               "rt_trace");
    for (uint32_t i = 0; i < count; ++i) {
        out[i].IDTr = hits[i].tri;
        out[i].u = hits[i].u;
        out[i].v = hits[i].v;
        out[i].t = hits[i].t;
        out[i].tmp[0] = hits[i].visited;
        out[i].tmp[1] = hits[i].root;
    }
}

TIntersection LbvhGpu::IntersectP(Ray &ray)
{
    TIntersection r{};
    IntersectP(&ray, 1, &r);
    return r;
}

void StructureConstructionHip::GPUBuildStructure(unsigned int uiDepth, LbvhGpu &lbvh)
{
    if (uiDepth != lbvh.depth())
        throw Error(RT_ERR_ARG, "GPUBuildStructure: depth differs from the scene's heap depth");
    dev_.check(rt_lbvh_build(dev_.ctx(), lbvh.scene(), &times_), "rt_lbvh_build");
}

RayTracerHip::RayTracerHip(Device &dev, LbvhGpu &lbvh, unsigned int width, unsigned int height,
                           bool dynamic)
    : dev_(dev), lbvh_(lbvh), build_(dev), dynamic_(dynamic)
{
    desc_.width = width;
    desc_.height = height;
    desc_.ambient = 0.1f;
    desc_.albedo = 0.8f;
    desc_.env = 0.0f;
    // Input's defaults (Common/Input.cpp:23-40): shadows on, Phong on (it
    // needs normals, so it starts off until SetMaterials brings them),
    // gloss mapping on, environment-mapping flag -1, no bounces
    wdesc_.width = width;
    wdesc_.height = height;
    wdesc_.switches = RT_WHITTED_SHADOWS | RT_WHITTED_GLOSS;
    wdesc_.env_flag = -1;
    wdesc_.bounces = 0;
    wdesc_.seed = 1;
    const float pos[3] = {0.0f, 0.0f, 3.0f}, tgt[3] = {0.0f, 0.0f, 0.0f}, up[3] = {0.0f, 1.0f, 0.0f};
    SetCamera(pos, tgt, up, 60.0f);
}

void RayTracerHip::SetCamera(const float pos[3], const float target[3], const float up[3],
                             float fov_y_deg)
{
    rt_camera_look_at(&desc_.cam, pos, target, up, fov_y_deg,
                      (float)desc_.width / (float)desc_.height);
    wdesc_.cam = desc_.cam;
}

void RayTracerHip::SetCamera(const float world[16])
{
    rt_camera_from_world(&desc_.cam, world);
    wdesc_.cam = desc_.cam;
}

void RayTracerHip::SetMaterials(const float *normals, const uint32_t *mat_ids, const rt_material *materials,
                                uint32_t nmat)
{
    uint32_t nt = 0, nv = 0, depth = 0;
    dev_.check(rt_scene_info(dev_.ctx(), lbvh_.scene(), &nt, &nv, &depth), "rt_scene_info");
    dev_.check(rt_scene_shading(dev_.ctx(), lbvh_.scene(), normals, nv, mat_ids, nt, materials, nmat),
               "rt_scene_shading");
    if (normals) set_switch(RT_WHITTED_PHONG, true);
}

void RayTracerHip::SetEnvironment(float r, float g, float b)
{
    wdesc_.env[0] = r;
    wdesc_.env[1] = g;
    wdesc_.env[2] = b;
}

void RayTracerHip::SelectAccelerationStructure(int id, uint32_t max_prims, uint32_t split)
{
    if (id == RT_STRUCTURE_BVH && !bvh_built_) {
        rt_bvh_info info;
        dev_.check(rt_bvh_build(dev_.ctx(), lbvh_.scene(), max_prims, split, &info), "rt_bvh_build");
        bvh_built_ = true;
    }
    dev_.check(rt_scene_select(dev_.ctx(), lbvh_.scene(), id), "rt_scene_select");
    structure_ = id;
}

void RayTracerHip::TakeScreenshot(const char *path)
{
    std::vector<uint8_t> img(4ull * desc_.width * desc_.height);
    dev_.check(rt_present(dev_.ctx(), nullptr, desc_.width, desc_.height, 0, img.data()), "rt_present");
    const int rc = rt_image_write(path, img.data(), desc_.width, desc_.height);
    if (rc != RT_OK) throw Error(rc, rt_image_last_error());
}

void RayTracerHip::SetLights(const std::vector<float> &xyz)
{
    if (xyz.size() % 3 != 0 || xyz.size() > 24) throw Error(RT_ERR_ARG, "SetLights: 0..8 lights");
    lights_ = xyz;
}

void RayTracerHip::SetShading(float ambient, float albedo, float env)
{
    desc_.ambient = ambient;
    desc_.albedo = albedo;
    desc_.env = env;
}

void RayTracerHip::Render()
{
    // RayTracerCS::Render (RayTracerCS.cpp:258-295): rebuild when dynamic,
    // then primary rays, intersections and shading with shadow rays.
    // The SAH BVH is static (the reference builds it once on the host), so a
    // dynamic rebuild applies to the LBVH only.
    if (dynamic_ && structure_ == RT_STRUCTURE_LBVH) build_.GPUBuildStructure(lbvh_.depth(), lbvh_);
    if (whitted_) {
        if (lights_.size() < 3) throw Error(RT_ERR_ARG, "Render: the ray-tracing stage needs a light");
        for (int k = 0; k < 3; ++k) wdesc_.light[k] = lights_[k];
        wdesc_.frame = samples_;
        dev_.check(rt_render_whitted(dev_.ctx(), lbvh_.scene(), &wdesc_, 0, &stats_), "rt_render_whitted");
    } else {
        desc_.lights = lights_.data();
        desc_.nlights = (uint32_t)(lights_.size() / 3);
        dev_.check(rt_render_direct(dev_.ctx(), lbvh_.scene(), &desc_, 0, &stats_), "rt_render_direct");
    }
    ++samples_;
}

void RayTracerHip::ReadFrame(std::vector<float> &rgba)
{
    rgba.resize(4ull * desc_.width * desc_.height);
    dev_.check(rt_frame_read(dev_.ctx(), rgba.data(), nullptr, nullptr), "rt_frame_read");
}

}  // namespace rtgpu
