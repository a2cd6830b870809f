// Drives the C++ facades end to end: torus scene (16,384 triangles), LBVH built
// on the device through LbvhGpu / StructureConstructionHip, a few frames of
// RayTracerHip::Render, and one batch of LbvhGpu::IntersectP.  Prints one JSON
// line with what it measured; exits non-zero on any error.
#include <cstdio>
#include <vector>

#include "facade.h"

int main(int argc, char **argv)
{
    const unsigned int frames = argc > 1 ? (unsigned)atoi(argv[1]) : 3u;
    try {
        uint32_t nv = 0, nt = 0;
        rt_gen_torus(1.0f, 0.3f, 128, 64, 1e-3f, 7, nullptr, &nv, nullptr, &nt);
        std::vector<float> xyz(3ull * nv);
        std::vector<uint32_t> idx(3ull * nt);
        if (rt_gen_torus(1.0f, 0.3f, 128, 64, 1e-3f, 7, xyz.data(), &nv, idx.data(), &nt) != RT_OK)
            return 2;
        rtgpu::Device dev(0);
        rtgpu::LbvhGpu lbvh(dev, xyz.data(), nv, idx.data(), nt, 1 << 20);
        rtgpu::RayTracerHip rt(dev, lbvh, 512, 512, true);
        const float pos[3] = {0.0f, 0.5f, 3.0f}, tgt[3] = {0, 0, 0}, up[3] = {0, 1, 0};
        rt.SetCamera(pos, tgt, up, 60.0f);
        rt.SetLights({0.0f, 0.0f, 0.0f});
        double trace_ms = 0.0;
        uint64_t rays = 0;
        for (unsigned int f = 0; f < frames; ++f) {
            rt.Render();
            trace_ms += rt.LastStats().trace_ms;
            rays += rt.LastStats().primary_rays + rt.LastStats().shadow_rays;
        }
        std::vector<float> rgba;
        rt.ReadFrame(rgba);
        double sum = 0.0;
        for (size_t i = 0; i < rgba.size(); i += 4) sum += rgba[i];
        // the ray-tracing colour stage with 2 bounces on the SAH BVH, then a
        // screenshot (argv[2], optional)
        rt.SetRayTracing(true);
        rt.SetNumBounces(2);
        rt.SetEnvironment(0.3f, 0.4f, 0.5f);
        rt.SelectAccelerationStructure(RT_STRUCTURE_BVH);
        rt.Render();
        const uint64_t whitted_rays =
            rt.LastStats().primary_rays + rt.LastStats().shadow_rays + rt.LastStats().bounce_rays;
        if (argc > 2) rt.TakeScreenshot(argv[2]);
        rt.SelectAccelerationStructure(RT_STRUCTURE_LBVH);
        // one ray straight down the +z axis at the torus tube (x = 1)
        rtgpu::Ray r{};
        r.o[0] = 1.0f;
        r.o[1] = 0.0f;
        r.o[2] = 3.0f;
        r.d[2] = -1.0f;
        r.maxt = 1e16f;
        r.TriangleID = -1;
        rtgpu::TIntersection h = lbvh.IntersectP(r);
        const rtgpu::BBox wb = lbvh.WorldBound();
        printf("{\"frames\": %u, \"samples\": %u, \"rays\": %llu, \"trace_ms\": %.4f, "
               "\"build_ms\": %.4f, \"image_mean\": %.6f, \"probe_tri\": %d, \"probe_t\": %.6f, "
               "\"root_box\": [%.4f, %.4f, %.4f, %.4f, %.4f, %.4f], \"keyboard_id\": %u, "
               "\"name\": \"%s\", \"whitted_rays\": %llu}\n",
               frames, rt.Samples(), (unsigned long long)rays, trace_ms,
               rt.LastBuildTimes().total_ms, sum / (rgba.size() / 4), h.IDTr, h.t, wb.pMin[0],
               wb.pMin[1], wb.pMin[2], wb.pMax[0], wb.pMax[1], wb.pMax[2], lbvh.GetKeyboardId(),
               lbvh.GetName(), (unsigned long long)whitted_rays);
        return 0;
    } catch (const rtgpu::Error &e) {
        fprintf(stderr, "rt_frame_demo: %s (code %d)\n", e.what(), e.code());
        return 1;
    }
}
