// Drives the C++ facades end to end: torus scene (16,384 triangles), LBVH built
// on the device through LbvhGpu / StructureConstructionHip, a few frames of
// RayTracerHip::Render, one batch of LbvhGpu::IntersectP, and a frame from the
// reference's camera matrix (RayTracerHip::SetCamera(world)).  Prints one JSON
// line with what it measured; exits non-zero on any error.
#include <cmath>
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
        // the reference's own camera: a Matrix4 (rotation columns right, up,
        // forward; translation in column 3) at (0.3, 0.8, -2.6) looking at the
        // origin; one direct frame, written raw (argv[3], optional)
        const float eye[3] = {0.3f, 0.8f, -2.6f};
        float fw[3] = {-eye[0], -eye[1], -eye[2]}, rt3[3], up3[3];
        float fl = std::sqrt(fw[0] * fw[0] + fw[1] * fw[1] + fw[2] * fw[2]);
        for (float &c : fw) c /= fl;
        rt3[0] = fw[2];  // (0, 1, 0) x fw
        rt3[1] = 0.0f;
        rt3[2] = -fw[0];
        float rl = std::sqrt(rt3[0] * rt3[0] + rt3[2] * rt3[2]);
        rt3[0] /= rl;
        rt3[2] /= rl;
        up3[0] = fw[1] * rt3[2] - fw[2] * rt3[1];
        up3[1] = fw[2] * rt3[0] - fw[0] * rt3[2];
        up3[2] = fw[0] * rt3[1] - fw[1] * rt3[0];
        float world[16] = {};
        for (int k = 0; k < 3; ++k) {
            world[4 * k + 0] = rt3[k];
            world[4 * k + 1] = up3[k];
            world[4 * k + 2] = fw[k];
            world[4 * k + 3] = eye[k];
        }
        world[15] = 1.0f;
        rt.SetRayTracing(false);
        rt.SetCamera(world);
        rt.Render();
        std::vector<float> wframe;
        rt.ReadFrame(wframe);
        if (argc > 3) {
            FILE *f = fopen(argv[3], "wb");
            if (!f || fwrite(wframe.data(), sizeof(float), wframe.size(), f) != wframe.size()) return 3;
            fclose(f);
        }
        char wtxt[512];
        int at = 0;
        for (int k = 0; k < 16; ++k) at += snprintf(wtxt + at, sizeof(wtxt) - at, "%s%.9g", k ? ", " : "", world[k]);
        printf("{\"frames\": %u, \"samples\": %u, \"rays\": %llu, \"trace_ms\": %.4f, "
               "\"build_ms\": %.4f, \"image_mean\": %.6f, \"probe_tri\": %d, \"probe_t\": %.6f, "
               "\"root_box\": [%.4f, %.4f, %.4f, %.4f, %.4f, %.4f], \"keyboard_id\": %u, "
               "\"name\": \"%s\", \"whitted_rays\": %llu, \"world\": [%s]}\n",
               frames, rt.Samples(), (unsigned long long)rays, trace_ms,
               rt.LastBuildTimes().total_ms, sum / (rgba.size() / 4), h.IDTr, h.t, wb.pMin[0],
               wb.pMin[1], wb.pMin[2], wb.pMax[0], wb.pMax[1], wb.pMax[2], lbvh.GetKeyboardId(),
               lbvh.GetName(), (unsigned long long)whitted_rays, wtxt);
        return 0;
    } catch (const rtgpu::Error &e) {
        fprintf(stderr, "rt_frame_demo: %s (code %d)\n", e.what(), e.code());
        return 1;
    }
}
