// C++ facades over the rtcore C ABI that keep the reference's roles:
//   LbvhGpu                  -- AccelerationStructure for the LBVH
//                               (AccelerationStructures/AccelerationStructure.h:40-67,
//                                LBVH/LBVH.h:70-118), built on the device
//   StructureConstructionHip -- StructureConstruction::GPUBuildStructure
//                               (RayTracerCS/StructureConstruction.h:67-91)
//   RayTracerHip             -- RayTracerCS::Render (RayTracerCS/RayTracerCS.h:64-144)
// The reference's methods return void and report nothing; these throw
// std::runtime_error carrying rt_last_error() when the device path fails.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "rtcore.h"

namespace rtgpu {

// TIntersection (Common/Geometry/Geometry.h:116-122): IDTr, u, v, t; tmp[0]
// carries the visit counter and tmp[1] the hit leaf's heap index.
struct TIntersection {
    int IDTr;
    float u, v, t;
    int tmp[2];
};

// Ray (Common/Geometry/Ray.h:22-49) in the device byte layout (48 bytes).
struct Ray {
    float o[3];
    float d[3];
    float factor[3];
    float mint, maxt;
    int TriangleID;
};
static_assert(sizeof(Ray) == sizeof(rt_ray48), "Ray layout");

struct BBox {
    float pMin[3];
    float pMax[3];
};

class Error : public std::runtime_error {
   public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

   private:
    int code_;
};

// Owns one rtcore context on a device.
class Device {
   public:
    explicit Device(int device = 0);
    ~Device();
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    rt_ctx *ctx() const { return ctx_; }
    void check(int rc, const char *what) const;

   private:
    rt_ctx *ctx_ = nullptr;
};

// AccelerationStructure role.  n = number of heap nodes (a power of two),
// as the LBVH constructor's third argument (LBVH.h:104).
class LbvhGpu {
   public:
    LbvhGpu(Device &dev, const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt,
            int n = 1 << 21);
    ~LbvhGpu();
    LbvhGpu(const LbvhGpu &) = delete;
    LbvhGpu &operator=(const LbvhGpu &) = delete;

    void Build();
    TIntersection IntersectP(Ray &ray);
    void IntersectP(const Ray *rays, uint32_t count, TIntersection *out);

    unsigned int GetNumberOfElements() const { return 1u << depth_; }
    size_t GetSizeOfElement() const { return sizeof(rt_node32); }
    void *GetPtrToElements();
    unsigned int *GetOrderedElementsIds();
    unsigned int GetKeyboardId() const { return 9; }
    BBox WorldBound();
    const char *GetName() const { return "LBVH"; }
    unsigned int GetNumPrimitives() const { return nt_; }

    int scene() const { return scene_; }
    unsigned int depth() const { return depth_; }
    const rt_build_times &LastBuildTimes() const { return times_; }

   private:
    void fetch();
    Device &dev_;
    int scene_ = -1;
    unsigned int depth_ = 0, nt_ = 0;
    bool fetched_ = false;
    rt_build_times times_{};
    std::vector<rt_node32> nodes_;
    std::vector<int32_t> ids_;
};

// StructureConstruction role: the per-frame GPU rebuild of a scene's LBVH.
class StructureConstructionHip {
   public:
    explicit StructureConstructionHip(Device &dev) : dev_(dev) {}
    // The reference passes the depth per call; here the depth is fixed when
    // the scene is created, and a mismatch is an error.
    void GPUBuildStructure(unsigned int uiDepth, LbvhGpu &lbvh);
    const rt_build_times &LastTimes() const { return times_; }

   private:
    Device &dev_;
    rt_build_times times_{};
};

// RayTracerCS role: one frame = optional rebuild + primary rays + the colour
// stage.  By default the direct-lighting frame (white Lambert, any number of
// lights); SetRayTracing(true) switches to the reference's ray-tracing
// colour stage with materials, one light and bounces (CSComputeColorRT.hlsl),
// driven by the toggles of its Input class (Common/Input.h:40-72).
class RayTracerHip {
   public:
    RayTracerHip(Device &dev, LbvhGpu &lbvh, unsigned int width, unsigned int height,
                 bool dynamic);
    void SetCamera(const float pos[3], const float target[3], const float up[3], float fov_y_deg);
    // the reference's own camera: the Matrix4 that cbCamera::Update uploads
    // (Camera::GetInverseMatrix, RayTracerCS/ConstantBuffers.h:28-31)
    void SetCamera(const float world[16]);
    void SetLights(const std::vector<float> &xyz);
    void SetShading(float ambient, float albedo, float env);
    // ray-tracing colour stage and its toggles
    void SetRayTracing(bool on) { whitted_ = on; }
    void SetMaterials(const float *normals, const uint32_t *mat_ids, const rt_material *materials,
                      uint32_t nmat);
    void SetNumBounces(unsigned int n) { wdesc_.bounces = n; }
    unsigned int GetNumBounces() const { return wdesc_.bounces; }
    void SetShadows(bool on) { set_switch(RT_WHITTED_SHADOWS, on); }
    void SetPhongShading(bool on) { set_switch(RT_WHITTED_PHONG, on); }
    void SetGlossMapping(bool on) { set_switch(RT_WHITTED_GLOSS, on); }
    void ChangeEnvMappingFlag() { wdesc_.env_flag = wdesc_.env_flag == 0 ? -1 : 0; }
    int GetEnvMappingFlag() const { return wdesc_.env_flag; }
    void SetEnvironment(float r, float g, float b);
    // acceleration structure switch (g_iAccelerationStructure: 0 BVH, 9 LBVH);
    // 0 builds the SAH BVH on first use (bvh_max_prims_node, bvh_split_algorithm)
    void SelectAccelerationStructure(int id, uint32_t max_prims = 1, uint32_t split = RT_BVH_SAH);
    int GetAccelerationStructure() const { return structure_; }
    void Render();
    void ReadFrame(std::vector<float> &rgba);
    // RendererDX11::TakeScreenshot: the UNORM8 frame as PNG or PPM
    void TakeScreenshot(const char *path);
    unsigned int Samples() const { return samples_; }
    const rt_frame_stats &LastStats() const { return stats_; }
    const rt_build_times &LastBuildTimes() const { return build_.LastTimes(); }

   private:
    void set_switch(uint32_t bit, bool on) { wdesc_.switches = on ? (wdesc_.switches | bit) : (wdesc_.switches & ~bit); }
    Device &dev_;
    LbvhGpu &lbvh_;
    StructureConstructionHip build_;
    rt_render_desc desc_{};
    rt_whitted_desc wdesc_{};
    std::vector<float> lights_;
    bool dynamic_;
    bool whitted_ = false;
    bool bvh_built_ = false;
    int structure_ = RT_STRUCTURE_LBVH;
    unsigned int samples_ = 0;
    rt_frame_stats stats_{};
};

}  // namespace rtgpu
