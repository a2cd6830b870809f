/*
 * rtcore -- MI355X-native LBVH build + trace core (C ABI).
 *
 * One context = one HIP device + one HIP stream (default priority).  All
 * calls on a context are ordered on that stream; a context is not
 * thread-safe (the caller serialises), as the reference's single D3D
 * immediate context is not.  Every entry point returns RT_OK (0) or a
 * negative rt_status; rt_last_error() then holds a message.  No entry point
 * aborts the process.
 *
 * Byte layouts are the reference's own (RayTracerCS/Shaders/Structures.hlsl),
 * so facades can hand buffers across without repacking.
 */
#ifndef RTCORE_H
#define RTCORE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rt_status {
    RT_OK = 0,
    RT_ERR_ARG = -1,      /* bad argument (null pointer, size, unknown scene) */
    RT_ERR_HIP = -2,      /* a HIP runtime call failed */
    RT_ERR_NOMEM = -3,    /* device allocation failed */
    RT_ERR_STATE = -4,    /* call out of order (e.g. trace before build) */
    RT_ERR_CAPACITY = -5, /* input exceeds what the scene was created for */
    RT_ERR_BUDGET = -6,   /* the device-memory budget cannot hold the request */
    RT_ERR_NODEV = -7     /* no usable HIP device */
} rt_status;

/* Node / LBVHNode: Structures.hlsl:18-26, LBVH.h:25-28 (32 bytes).
 * prim_count = (count << 3) | leaf(4) | axis(0..2); prim_pos = first index
 * into the sorted primitive arrays.  Heap layout: node 0 unused, root 1,
 * children of n at 2n and 2n+1. */
typedef struct rt_node32 {
    float mn[3];
    float mx[3];
    int32_t prim_count;
    int32_t prim_pos;
} rt_node32;

/* Ray: Structures.hlsl:28-35 (48 bytes).  tri = triangle id to skip (self),
 * -1 for none. */
typedef struct rt_ray48 {
    float o[3];
    float d[3];
    float refl[3];
    float tmin, tmax;
    int32_t tri;
} rt_ray48;

/* Intersection: Structures.hlsl:44-50 (24 bytes).  visited counts +1 per box
 * test and +65536 per triangle test (LBVH.hlsl:389,404); root is the heap
 * index of the leaf holding the hit (1 when there is none). */
typedef struct rt_hit24 {
    int32_t tri;
    float u, v, t;
    int32_t visited;
    int32_t root;
} rt_hit24;

/* Device time of the build stages of the last rt_lbvh_build, milliseconds. */
typedef struct rt_build_times {
    float bounds_ms; /* per-primitive boxes + scene-box partials */
    float morton_ms; /* scene box, 30-bit codes, digit histograms */
    float sort_ms;   /* onesweep radix sort of (code, id) */
    float emit_ms;   /* split tables + emission of the nodes above 512 primitives */
    float refit_ms;  /* in-LDS emission + refit of the small subtrees, refit of
                        the big nodes, triangle relayout */
    float total_ms;  /* first launch to last, on the device */
} rt_build_times;

/* Primary-ray camera (CSPrimaryRays.hlsl:22-69), one of two kinds:
 *  - RT_CAMERA_BASIS (0): a pinhole with a unit basis and half-angle tangents
 *    (rt_camera_look_at).  Pixel (px, py) of a W x H image:
 *    x = ((2px+1) - W)/W, y = -((2py+1) - H)/H,
 *    d = normalize((fwd + x*tan_x*right) + y*tan_y*up), origin pos.
 *  - RT_CAMERA_WORLD (1): the reference's own camera, the cbCamera matrix
 *    g_mfWorld (Shaders/ConstantBuffers.hlsl:18-21) as the host's Matrix4
 *    holds it (Camera::GetInverseMatrix, Common/Camera.h:64; row-major, the
 *    translation in column 3; RayTracerCS/ConstantBuffers.h:23-31), with the
 *    rays of CSPrimaryRays.hlsl:39-52: x = ((2px+1) - W)/H,
 *    y = -((2py+1) - H)/H, z = 2, scale = 0.008/|0.035 z|,
 *    origin = (M (0,0,0,1)).xyz / w,
 *    d = normalize((M (scale x, scale y, scale z, 1)).xyz / w - origin).
 *    For the reference's square images this is its rule exactly; wider
 *    images keep its pixel pitch.  Set it with rt_camera_from_world; the
 *    library derives pos from the matrix.
 * Each matrix row is summed left to right; normalize is v / sqrt(v.v). */
enum { RT_CAMERA_BASIS = 0, RT_CAMERA_WORLD = 1 };
typedef struct rt_camera {
    float pos[3];
    float fwd[3];
    float right[3];
    float up[3];
    float tan_x, tan_y;
    uint32_t kind;   /* RT_CAMERA_BASIS or RT_CAMERA_WORLD */
    float world[16]; /* RT_CAMERA_WORLD: the Matrix4's 16 floats, xx xy xz xt yx ... tt */
} rt_camera;

/* Direct-lighting frame: primary rays, one shadow ray per light per hit
 * (CSComputeColorRT.hlsl:126-177), flat normals, a white Lambert material
 * of alpha 1: color = (ambient + sum_lights max(0, -n.L)) * albedo; misses
 * return env. */
typedef struct rt_render_desc {
    rt_camera cam;
    uint32_t width, height;
    const float *lights; /* host array, 3 floats per point light */
    uint32_t nlights;    /* 0..8 */
    float ambient, albedo, env;
} rt_render_desc;

/* Counters of the last frame / trace call. */
typedef struct rt_frame_stats {
    float trace_ms;         /* device time of the trace kernel(s) */
    uint64_t primary_rays;
    uint64_t shadow_rays;
    uint64_t bounce_rays;
    uint64_t box_tests;     /* filled only with RT_FLAG_COUNT */
    uint64_t tri_tests;     /* filled only with RT_FLAG_COUNT */
} rt_frame_stats;

enum {
    RT_FLAG_ANYHIT = 1u << 0,    /* stop at the first accepted triangle */
    RT_FLAG_COUNT = 1u << 1,     /* count box / triangle tests into stats */
    RT_FLAG_KEEP_HITS = 1u << 2, /* keep per-pixel primary hits + shadow results */
    RT_FLAG_TIMING = 1u << 3,    /* time the call with HIP events (synchronises) */
    RT_FLAG_PERSISTENT = 1u << 4, /* rt_render_jobs: one persistent launch over all tiles */
    RT_FLAG_START_AT_IROOT = 1u << 5 /* rt_trace: ray i starts at heap node hits[i].root, the
                                        previous intersection's iRoot, as CSComputeIntersections
                                        passes it to LBVH_IntersectP (CSIntersections.hlsl:32);
                                        hits are read, then overwritten; 0 and 1 are the root, a
                                        node below a leaf starts at that leaf; LBVH only */
};

/* Launch shape of the traversal kernels (rt_autotune_jobs picks one). */
typedef struct rt_tuning {
    int32_t block;         /* threads per block: 64, 128 or 256 (tile 16 x block/16) */
    int32_t lds_stack;     /* stack entries per lane kept in LDS (8..26); deeper ones spill */
    int32_t blocks_per_cu; /* resident blocks per CU for persistent job launches */
    int32_t traversal;     /* 0: loop form, 1: step form, 2: loop form fetching nodes and
                              leaves with one shared load (same visit order and results;
                              the fastest one depends on the scene) */
    int32_t quad_tiles;    /* rt_render_direct of an LBVH scene without RT_FLAG_KEEP_HITS or
                              RT_FLAG_COUNT: the quad_tiles heaviest tiles of the last frame's
                              tile order get four lanes per pixel, each searching every fourth
                              subtree of heap depth 8 along the ray (same image; shortens the
                              longest pixels' chains).  0: off */
} rt_tuning;

/* One image of one scene for rt_render_jobs: spp samples per pixel are added
 * into accum (device memory, width * height * 4 floats: r, g, b, samples).
 * RT_MODE_DIRECT: primary + shadow rays as rt_render_desc (ambient, albedo,
 * lights; misses add sky).  RT_MODE_GI: diffuse path tracing, up to
 * max_bounces cosine-weighted bounces, Lambert albedo, constant sky emitter
 * (CSComputeColorGI.hlsl:153-230 structure).  Primary rays are jittered in the
 * pixel; random numbers are PCG32 streams keyed by (seed, frame, sample,
 * pixel), so a pixel's result does not depend on the other jobs of a launch. */
enum { RT_MODE_DIRECT = 0, RT_MODE_GI = 1 };
typedef struct rt_job_desc {
    int scene;
    rt_camera cam;
    uint32_t width, height;
    uint32_t mode, spp, max_bounces;
    float albedo, sky, ambient;
    const float *lights; /* host, 3 floats per light (direct mode), up to 8 */
    uint32_t nlights;
    uint64_t seed;
    uint32_t frame;
    float *accum;
    /* optional tile range (tiles of 16 x block/16 pixels, positions in the
     * job's tile order): tile_count == 0 renders the whole image */
    uint32_t tile_begin, tile_count;
    /* order in which a persistent launch's queue hands out the job's tiles:
     * RT_TILES_ROWS (row-major) or RT_TILES_CENTRE (nearest the image centre
     * first, where a camera aimed at its scene puts the longest paths, so
     * they start early and the launch's tail is made of short tiles).  The
     * image does not depend on it. */
    uint32_t tile_order;
} rt_job_desc;
enum { RT_TILES_ROWS = 0, RT_TILES_CENTRE = 1 };

typedef struct rt_ctx rt_ctx;

/* ---- context (DeviceCore role: Device::Initialize, DeviceCore/Device.hpp:68-87) */
int rt_create(rt_ctx **out, int device);
void rt_destroy(rt_ctx *ctx);
const char *rt_last_error(const rt_ctx *ctx);
int rt_sync(rt_ctx *ctx);
int rt_get_stream(rt_ctx *ctx, void **hip_stream);
/* Device bytes held by the context (all scenes and frame buffers). */
int rt_memory_in_use(rt_ctx *ctx, uint64_t *bytes);

/* ---- scenes (Model + its primitive/vertex buffers, Common/Scenes/Model.cpp:163-205;
 *      buffer creation RayTracerCS.cpp:343-462) */
/* Reserve device storage for up to max_tris triangles over max_verts vertices
 * and a heap of 1 << depth nodes (conf.ini lbvh_depth; 2 <= depth <= 26). */
int rt_scene_create(rt_ctx *ctx, uint32_t max_tris, uint32_t max_verts,
                    uint32_t depth, int *scene);
/* A batch: nscenes scenes in one set of device arrays, built together by one
 * sequence of launches (a plain scene is a batch of one).  Scene handles are
 * first_scene .. first_scene + nscenes - 1; rt_lbvh_build on any of them
 * rebuilds every scene of the batch.  At most 256 scenes per batch. */
int rt_batch_create(rt_ctx *ctx, uint32_t nscenes, const uint32_t *max_tris,
                    const uint32_t *max_verts, const uint32_t *depths, int *first_scene);
/* Releases the handle; a batch's memory is freed with its last handle. */
int rt_scene_destroy(rt_ctx *ctx, int scene);
/* Copy vertices (3 floats each) and triangle indices (3 uint32 each; the
 * reference's DWORD, which is 8 bytes on LP64 Linux, is not used). */
int rt_scene_upload(rt_ctx *ctx, int scene, const float *xyz, uint32_t nv,
                    const uint32_t *idx, uint32_t nt);
/* Same, from device pointers (stream-ordered device-to-device copy). */
int rt_scene_upload_device(rt_ctx *ctx, int scene, const float *d_xyz,
                           uint32_t nv, const uint32_t *d_idx, uint32_t nt);
/* Rigid motion of the uploaded (rest) vertices: v = R (v0 - c) + c, with
 * R row-major in m[0..8] and c in m[9..11]; float arithmetic in the order
 * ((R0 dx + R1 dy) + R2 dz) + cx, no FMA. */
int rt_scene_transform(rt_ctx *ctx, int scene, const float m[12]);
int rt_scene_info(rt_ctx *ctx, int scene, uint32_t *nt, uint32_t *nv,
                  uint32_t *depth);

/* ---- build (StructureConstruction::GPUBuildStructure, StructureConstruction.cpp:86-131) */
/* Rebuilds the LBVH of the scene's current vertices (and of every other scene
 * of its batch) on the context stream.
 * times != NULL synchronises and fills the per-stage device times; the event
 * at each stage boundary stalls the stream (about 5 us each on gfx950), so
 * total_ms is then a little above an unobserved build's time. */
int rt_lbvh_build(rt_ctx *ctx, int scene, rt_build_times *times);
/* Copies build results to the host; any pointer may be NULL.
 * nodes: 1 << depth entries (AccelerationStructure::GetPtrToElements);
 * sorted_codes / sorted_ids: nt entries (GetOrderedElementsIds);
 * scene_box: {min xyz, max xyz} used for the Morton quantisation;
 * Unreachable heap slots hold unspecified values, as in the CPU builder
 * (LBVH.cpp:577); children of leaves hold -2147483647 in both ints. */
int rt_lbvh_read(rt_ctx *ctx, int scene, rt_node32 *nodes,
                 uint32_t *sorted_codes, int32_t *sorted_ids, float scene_box[6]);

/* ---- trace (CSComputeIntersections -> LBVH_IntersectP, CSIntersections.hlsl:18-39,
 *      LBVH.hlsl:363-440).  Closest hit from the root, strict '<' in the
 *      reference visit order; rays with tri <= -2 are dead (hit.tri = -2). */
int rt_trace(rt_ctx *ctx, int scene, const rt_ray48 *rays, uint32_t n,
             rt_hit24 *hits, uint32_t flags, rt_frame_stats *stats);
/* Same with device pointers; asynchronous unless RT_FLAG_TIMING/COUNT. */
int rt_trace_device(rt_ctx *ctx, int scene, const rt_ray48 *d_rays, uint32_t n,
                    rt_hit24 *d_hits, uint32_t flags, rt_frame_stats *stats);

/* ---- frame (RayTracerCS::Render, RayTracerCS.cpp:241-311, without rebuild) */
/* Primary + shadow rays + direct shading into the context's frame buffer. */
int rt_render_direct(rt_ctx *ctx, int scene, const rt_render_desc *desc,
                     uint32_t flags, rt_frame_stats *stats);
/* Copy the last frame to the host; any pointer may be NULL.  rgba: 4 floats
 * per pixel; hits: primary hits (needs RT_FLAG_KEEP_HITS); occluded: one byte
 * per pixel per light, 0 lit, 1 occluded, 2 no shadow ray (KEEP_HITS). */
int rt_frame_read(rt_ctx *ctx, float *rgba, rt_hit24 *hits, uint8_t *occluded);

/* ---- Whitted frames (the reference's ray-tracing colour stage) */
/* Material: Structures.hlsl:69-78 (64 bytes), the layout of MaterialData
 * (Common/Scenes/MaterialData.h:24-28) that fills the material buffer.  As the
 * colour stage uses it (CSComputeColorRT.hlsl:179-236): diffuse.rgb is the
 * surface colour and diffuse.w the weight of the lit diffuse term (1 - w
 * weights the specular term); ambient.rgb is added to the diffuse term and
 * ambient.w is the index of refraction; specular.rgb colours the specular
 * term and the reflected continuation, specular.w is the gloss (times the
 * RT_WHITTED_GLOSS switch); alpha is the probability that the continuation
 * ray reflects rather than refracts.  shininess is carried but unused (the
 * specular exponent is the constant 49). */
typedef struct rt_material {
    float diffuse[4];
    float ambient[4];
    float specular[4];
    float shininess;
    float alpha;
    uint32_t id;
    uint32_t pad;
} rt_material;

/* Shading attributes of an uploaded scene: vertex normals (Model vertices,
 * Model.cpp:320-333), one material id per triangle (PrimitiveCS,
 * RayTracerCS.cpp:322-327) and the material table (RayTracerCS.cpp:348-363).
 * normals: 3 floats per vertex in rt_scene_upload's vertex order, nv equal
 * to the uploaded count; NULL keeps flat shading only.  mat_ids: nt entries
 * (nt equal to the uploaded count), each < nmat; NULL means material 0
 * everywhere.  materials: nmat entries (1..65536); NULL with nmat 0 installs
 * the reference's default MaterialData (MaterialData.cpp:17-24).  Upload the
 * geometry first; a later rt_scene_upload with other counts voids these. */
int rt_scene_shading(rt_ctx *ctx, int scene, const float *normals, uint32_t nv,
                     const uint32_t *mat_ids, uint32_t nt,
                     const rt_material *materials, uint32_t nmat);

/* Whitted frame: RayTracerCS::Render (RayTracerCS.cpp:286-295) in the
 * ray-tracing mode -- primary rays (CSPrimaryRays.hlsl, no jitter), then
 * bounces + 1 passes of closest hit (CSIntersections.hlsl:18-39) and colour
 * (CSComputeColorRT.hlsl:18-262), one lane carrying its pixel through all
 * passes.  A hit adds its material colour lit by one point light (diffuse
 * max(0, -n.L), specular max(0, -L.R)^49, a shadow ray when
 * RT_WHITTED_SHADOWS) times the ray's reflective factor, then continues:
 * reflected with probability alpha (factor *= specular.rgb), else refracted
 * (factor *= diffuse.rgb; total internal reflection reflects).  A miss adds
 * factor * env and ends the path; with env_flag 0 a miss instead paints the
 * pixel (1, 1, 1, 1), as the reference's environment-mapping switch does.
 * Normals are flat, or interpolated vertex normals with RT_WHITTED_PHONG
 * (needs rt_scene_shading normals).  The reflect/refract draw is a PCG32
 * stream keyed by (seed, frame, pixel).  Texture, normal and gloss maps are
 * not sampled: this is the reference with texturing and normal mapping off.
 * Output: rt_frame_read's rgba is the reference's result texture; with
 * RT_FLAG_KEEP_HITS its hits / occluded are the first pass's (one light). */
enum {
    RT_WHITTED_SHADOWS = 1u << 0, /* g_bIsShadowOn */
    RT_WHITTED_PHONG = 1u << 1,   /* g_bIsPhongShadingOn */
    RT_WHITTED_GLOSS = 1u << 2,   /* g_bIsGlossMappingOn (specular.w) */
    RT_WHITTED_ONE_PASS = 1u << 3, /* one launch, each lane runs its pixel's passes; by default
                                      pass 0 runs per pixel and each bounce pass over a compacted
                                      queue of the live rays (same results) */
    RT_WHITTED_SPLIT = 1u << 4,    /* pass 0 per pixel, then one launch in which each live path
                                      of the compacted queue runs all its remaining passes (same
                                      results; not with RT_WHITTED_ONE_PASS) */
    RT_WHITTED_QUADS = 1u << 5     /* with RT_WHITTED_SPLIT: four lanes per live path share each
                                      remaining pass's closest-hit and shadow traversals (same
                                      results; an LBVH scene without RT_FLAG_COUNT, else one lane
                                      per path) */
};
typedef struct rt_whitted_desc {
    rt_camera cam;
    uint32_t width, height;
    float light[3];    /* point light position (g_vfLightPosition) */
    uint32_t bounces;  /* 0..16; passes = bounces + 1 (g_iNumBounces) */
    uint32_t switches; /* RT_WHITTED_* */
    int32_t env_flag;  /* -1 (reference default) or 0 (g_iEnvMappingFlag) */
    float env[3];      /* colour of a miss (a uniform environment map) */
    uint32_t frame;
    uint64_t seed;
} rt_whitted_desc;
int rt_render_whitted(rt_ctx *ctx, int scene, const rt_whitted_desc *desc,
                      uint32_t flags, rt_frame_stats *stats);
/* Times every launch shape, traversal form and the four Whitted pass forms
 * (per-pass wavefront, RT_WHITTED_ONE_PASS, RT_WHITTED_SPLIT, RT_WHITTED_SPLIT
 * | RT_WHITTED_QUADS) on this frame (renders into the context's frame
 * buffer), keeps the fastest shape in the context and sets RT_WHITTED_ONE_PASS
 * / RT_WHITTED_SPLIT / RT_WHITTED_QUADS in desc->switches to the fastest
 * form.  cache_path as for rt_autotune_jobs. */
int rt_autotune_whitted(rt_ctx *ctx, int scene, rt_whitted_desc *desc,
                        const char *cache_path, rt_tuning *best);

/* ---- second structure: the SAH BVH of the reference (BVH::Build,
 *      AccelerationStructures/BVH/bvh.cpp:64-349; traversal BVH_IntersectP,
 *      Shaders/AccelerationStructures/BVH.hlsl:124-217; keyboard id 0) */
enum { RT_BVH_SAH = 0, RT_BVH_MIDDLE = 1, RT_BVH_EQUAL = 2 }; /* conf.ini bvh_split_algorithm */
enum { RT_STRUCTURE_BVH = 0, RT_STRUCTURE_LBVH = 9 };         /* g_iAccelerationStructure */
typedef struct rt_bvh_info {
    uint32_t nodes;          /* LinearBVHNode count */
    uint32_t leaves;
    uint32_t max_depth;      /* levels of the deepest leaf + 1 (at most 27) */
    uint32_t forced_medians; /* splits replaced by the median to keep the depth */
    float sah_cost;          /* sum over nodes of area / root area x (0.125 inner, prims leaf) */
    double build_ms;         /* host build time */
} rt_bvh_info;
/* Builds the BVH of the scene's current vertices on the host (max_prims per
 * leaf, conf.ini bvh_max_prims_node; 1..255), stores its device layout and
 * selects it for every later rt_trace / rt_render_* of the scene (one
 * closest-hit order, the reference's BVH_IntersectP: near child first by the
 * split axis and the ray's sign).  Synchronises.  The BVH is static, as in
 * the reference: a new upload or rt_scene_transform drops it and selects
 * the LBVH. */
int rt_bvh_build(rt_ctx *ctx, int scene, uint32_t max_prims, uint32_t split, rt_bvh_info *info);
/* Which structure the scene's traversals use (the other must exist). */
int rt_scene_select(rt_ctx *ctx, int scene, int structure);
/* The built BVH in the reference's layout: nodes (info.nodes LinearBVHNode =
 * rt_node32: leaf prim_count > 0 primitives from prim_pos; inner prim_count =
 * -axis, first child next, second child at prim_pos) and the ordered
 * primitive ids (nt entries); either may be NULL. */
int rt_bvh_read(rt_ctx *ctx, int scene, rt_node32 *nodes, int32_t *ordered_ids);
/* Host only, no context: the same build from arrays.  nodes needs 2 nt - 1
 * entries, ordered_ids nt; info (optional) as above. */
int rt_bvh_build_host(const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt,
                      uint32_t max_prims, uint32_t split, rt_node32 *nodes,
                      int32_t *ordered_ids, rt_bvh_info *info);

/* ---- presentation (the swap chain's R8G8B8A8_UNORM back buffer,
 *      Renderer/DX11/RendererDX11.inl:60-100; TakeScreenshot, :134-150) */
/* 8-bit RGBA of a float RGBA image, each channel stored as D3D stores a
 * float into a UNORM8 texture: NaN -> 0, clamp to [0, 1], x 255, rounded to
 * nearest.  d_rgba: device image of width * height float4, or NULL for the
 * context's frame buffer (the last rt_render_direct / rt_render_whitted, of
 * this size).  RT_PRESENT_AVERAGE first divides r, g, b by w (the sample
 * count of an rt_render_jobs accumulation buffer) and sets alpha to 1.
 * rgba8: host, 4 bytes per pixel, rows top to bottom.  Synchronises. */
enum { RT_PRESENT_AVERAGE = 1u << 0 };
int rt_present(rt_ctx *ctx, const float *d_rgba, uint32_t width, uint32_t height,
               uint32_t flags, uint8_t *rgba8);
/* Host only: writes an 8-bit RGBA image as ".png" (RGBA; zlib stream of
 * stored blocks) or ".ppm" (binary P6, RGB), chosen by the file name.
 * Errors: rt_image_last_error() (this thread). */
int rt_image_write(const char *path, const uint8_t *rgba8, uint32_t width, uint32_t height);
const char *rt_image_last_error(void);

/* ---- profiling window (Timer role, DeviceCore/Timer.hpp:27-66; Performance,
 *      Common/Performance.cpp:52-101).  Between begin and end every build,
 *      trace and render call records HIP events around its kernels on the
 *      context stream, without synchronising, and the device counters
 *      (shadow / bounce rays always, box / triangle tests with RT_FLAG_COUNT)
 *      accumulate.  end synchronises and sums.  Per-call stats arguments
 *      reset the counters, so do not pass them inside a window. */
typedef struct rt_profile {
    uint32_t builds;     /* rt_lbvh_build calls in the window */
    float build_ms[6];   /* [5]: summed end-to-end build times; [0..4] are 0: a
                            window records no stage events, which would each
                            stall the stream between two kernels (rt_lbvh_build
                            with rt_build_times gives the stages) */
    uint32_t launches;   /* trace / render kernel launches in the window */
    float trace_ms;      /* summed device time of those launches */
    uint64_t primary_rays, shadow_rays, bounce_rays, box_tests, tri_tests;
} rt_profile;
int rt_profile_begin(rt_ctx *ctx);
int rt_profile_end(rt_ctx *ctx, rt_profile *out);

/* ---- batched rendering (no reference counterpart: the reference renders one
 *      scene per frame, RayTracerCS.cpp:241-311) */
int rt_tuning_get(rt_ctx *ctx, rt_tuning *t);
int rt_tuning_set(rt_ctx *ctx, const rt_tuning *t);
/* Renders all jobs: with RT_FLAG_PERSISTENT one launch whose blocks take tiles
 * of every job from one queue; without it one launch per job.  Jobs may name
 * scenes of different batches.  stats (optional, synchronises) reports the
 * kernel time and ray counts; with RT_FLAG_COUNT also the box and triangle
 * tests, and every ray then takes one lane (without it, the live paths and
 * shadow rays of an LBVH scene's waves share their idle lanes: the same
 * accumulations, untallied tests). */
int rt_render_jobs(rt_ctx *ctx, const rt_job_desc *jobs, uint32_t njobs,
                   uint32_t flags, rt_frame_stats *stats);
/* rt_render_jobs without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, code that must not call hipMalloc).  The job table, the
 * tile counter and the ray counters live in one static device array per GPU
 * (RT_NOALLOC_MAX_JOBS job records + 512 bytes, about 512 KiB, held until
 * the process ends); the caller passes the scenes and accumulators as for
 * rt_render_jobs, and the results are the same.  More than
 * RT_NOALLOC_MAX_JOBS jobs return RT_ERR_CAPACITY.  Every context on one GPU
 * shares that array: calls that overlap on one GPU (two contexts, or two
 * host threads) race and must be serialised by the caller.  Not allowed
 * inside a profiling window (RT_ERR_STATE).  A RT_TILES_CENTRE job uses the
 * order an earlier rt_render_jobs of that image size made, or else renders
 * in row order (the same image); a tile range over an order not made yet is
 * RT_ERR_STATE. */
#define RT_NOALLOC_MAX_JOBS 1024
int rt_render_jobs_noalloc(rt_ctx *ctx, const rt_job_desc *jobs, uint32_t njobs,
                           uint32_t flags, rt_frame_stats *stats);
/* Times every launch shape on these jobs (into scratch accumulators) and
 * keeps the fastest in the context.  cache_path (optional) is a text file of
 * earlier results keyed by the jobs' shape; a hit skips the search. */
int rt_autotune_jobs(rt_ctx *ctx, const rt_job_desc *jobs, uint32_t njobs,
                     const char *cache_path, rt_tuning *best);
/* Times every launch shape and traversal form on rt_render_direct of this
 * scene and frame (renders into the context's frame buffer), then, for an
 * LBVH scene, quad_tiles 64, 256 and 1024 with the fastest of them, and keeps
 * the fastest (quads only when at least 3 % faster); the blocks_per_cu field
 * is left as it is.  cache_path as for rt_autotune_jobs, keyed by triangle
 * count, depth, image size and light count. */
int rt_autotune_direct(rt_ctx *ctx, int scene, const rt_render_desc *desc,
                       const char *cache_path, rt_tuning *best);
/* Captures rebuild (optional) + one persistent rt_render_jobs launch of these
 * jobs as a HIP graph; each rt_graph_launch replays it with frame + n for the
 * n-th replay.  Jobs must name scenes of the batch of `scene`.  The graph
 * holds that batch (its memory stays allocated, and counted by
 * rt_memory_in_use, until rt_graph_destroy); the accumulators it names are
 * the caller's and must outlive it.  rt_graph_launch returns RT_ERR_STATE,
 * and runs nothing, once a scene handle of the batch has been destroyed, a
 * BVH of it built or dropped (rt_bvh_build, an upload), or its structure
 * switched (rt_scene_select) since the capture: capture again. */
int rt_graph_capture_frame(rt_ctx *ctx, int scene, int rebuild, const rt_job_desc *jobs,
                           uint32_t njobs, int *graph);
int rt_graph_launch(rt_ctx *ctx, int graph);
int rt_graph_destroy(rt_ctx *ctx, int graph);

/* ---- render server (no reference counterpart; it keeps the reference's
 *      progressive accumulation, RayTracerCS.cpp:246-250 with
 *      CSPrimaryRays.hlsl:62-63, per stream of requests).
 * Scenes are registered once by key and kept resident (built) while the
 * device-memory budget allows; least recently used idle scenes and
 * accumulators are evicted.  A request renders spp samples per pixel of one
 * view into the accumulator of its stream (scene_key, view, width, height,
 * mode): the stream's n-th request uses frame = n, so successive requests
 * refine one image.  Each rt_server_step admits arrived requests in arrival
 * order (one in flight per stream) and renders a share of every in-flight
 * request's tiles in one persistent rt_render_jobs launch; a request's pixels
 * are the same alone or in any mix.  dry_run performs no device work (no
 * context needed): a step takes estimated rays / dry_rate seconds. */
typedef struct rt_server rt_server;
typedef struct rt_server_config {
    uint64_t hbm_budget;     /* device bytes for scenes + accumulators */
    uint32_t max_inflight;   /* requests rendered concurrently */
    uint32_t tiles_per_step; /* tiles per launch, shared by in-flight requests */
    int32_t dry_run;
    double dry_rate;         /* rays per second assumed by dry_run */
    int32_t autotune;        /* 1: once RT_SERVER_TUNE_REQUESTS requests have completed,
                                time the launch shapes of the context's block size on
                                their whole images (scratch accumulators, as
                                rt_autotune_jobs) and launch every later step with the
                                fastest; the context's own tuning is left as it is */
    uint32_t tile_order;     /* tile order of the server's jobs (RT_TILES_ROWS / _CENTRE) */
    double batch_window;     /* seconds an arrived request may wait, while nothing is in
                                flight, for later arrivals to share its launch (0: none;
                                a full batch of max_inflight requests starts at once) */
} rt_server_config;
#define RT_SERVER_TUNE_REQUESTS 8
typedef struct rt_request {
    uint32_t scene_key, view;
    rt_camera cam;
    uint32_t width, height;
    uint32_t mode, spp, max_bounces;
    float albedo, sky, ambient;
    const float *lights;     /* host, 3 floats per light, up to 8; copied at submit */
    uint32_t nlights;
    uint64_t seed;
    double arrival;          /* seconds, caller's clock; admitted once now >= arrival */
} rt_request;
typedef struct rt_request_result {
    uint64_t id;
    double arrival, admitted, completed; /* completed = step's now + step duration */
    uint64_t rays;           /* rays of this request (share of its launches) */
    uint32_t frame;          /* pass index of the request in its stream */
    uint32_t samples;        /* samples per pixel in the stream after this request */
    int32_t status;          /* RT_OK, RT_ERR_BUDGET (never fits) or the failing call's code */
} rt_request_result;
typedef struct rt_server_stats {
    uint64_t submitted, completed, rejected, steps, builds, evictions, rays;
    uint64_t bytes_used;
    uint32_t pending, inflight, resident_scenes, accumulators;
    double device_s;         /* summed device time of the step launches (dry run: simulated) */
    double tune_s;           /* wall time of the launch-shape search (autotune) */
    rt_tuning tuning;        /* the shape the steps launch with */
    double wake;             /* when the last step held arrived requests for the batch
                                window: the time it ends (a step then does work); else 0 */
} rt_server_stats;
int rt_server_create(rt_ctx *ctx, const rt_server_config *cfg, rt_server **out);
void rt_server_destroy(rt_server *srv);
const char *rt_server_last_error(const rt_server *srv);
/* Registers a scene (copied to the host side of the server; built on demand). */
int rt_server_add_scene(rt_server *srv, uint32_t key, const float *xyz, uint32_t nv,
                        const uint32_t *idx, uint32_t nt, uint32_t depth);
int rt_server_submit(rt_server *srv, const rt_request *req, uint64_t *id);
/* One scheduling step at time `now` (seconds); *elapsed (optional) = the
 * step's duration: wall time of admission, builds and the synchronised
 * launch (dry run: the simulated launch time).  Requests retired by the step
 * complete at now + *elapsed; a step with nothing to render takes 0. */
int rt_server_step(rt_server *srv, double now, double *elapsed);
/* Moves up to cap finished results into out; *n = how many. */
int rt_server_poll(rt_server *srv, rt_request_result *out, uint32_t cap, uint32_t *n);
int rt_server_status(rt_server *srv, rt_server_stats *out);
/* Copies a resident stream's accumulator (width * height * 4 floats). */
int rt_server_read_stream(rt_server *srv, uint32_t scene_key, uint32_t view, uint32_t width,
                          uint32_t height, uint32_t mode, float *rgba_sum, uint32_t *samples);

/* ---- device buffers for accumulators (Buffer<T> role, DeviceCore/Buffer.hpp:49-77) */
int rt_device_alloc(rt_ctx *ctx, uint64_t bytes, void **ptr);
int rt_device_free(rt_ctx *ctx, void *ptr);
int rt_device_zero(rt_ctx *ctx, void *ptr, uint64_t bytes);
int rt_device_read(rt_ctx *ctx, void *host, const void *ptr, uint64_t bytes);

/* ---- camera helper (Camera::SetCamera role, Common/Camera.cpp:154-211) */
/* The reference's camera: kind RT_CAMERA_WORLD, world copied, pos set to
 * the matrix's origin (M (0,0,0,1)).xyz / w, the basis fields zeroed.
 * Replaces the upload of cbCamera (RayTracerCS.cpp:253, ConstantBuffers.h:28-31). */
void rt_camera_from_world(rt_camera *cam, const float world[16]);
/* Fills a pinhole camera (RT_CAMERA_BASIS) at pos looking at target, vertical
 * field of view fov_y_deg, aspect = width / height (basis computed in double,
 * rounded). */
void rt_camera_look_at(rt_camera *cam, const float pos[3], const float target[3],
                       const float up[3], float fov_y_deg, float aspect);

/* ---- seeded synthetic scenes (host only, no device work) */
/* Each generator writes vertices (3 floats each) and indices (3 uint32 per
 * triangle); call with NULL outputs to get the sizes in *nv / *nt. */
int rt_gen_torus(float R, float r, uint32_t nu, uint32_t nv_ring, float jitter,
                 uint64_t seed, float *xyz, uint32_t *nv, uint32_t *idx,
                 uint32_t *nt);
int rt_gen_triangle_soup(uint32_t ntri, float sigma, uint64_t seed, float *xyz,
                         uint32_t *nv, uint32_t *idx, uint32_t *nt);
int rt_gen_noisy_sphere(uint32_t nlat, uint32_t nlon, float amplitude,
                        uint64_t seed, float *xyz, uint32_t *nv, uint32_t *idx,
                        uint32_t *nt);
int rt_gen_object_mix(uint32_t target_tris, uint64_t seed, float *xyz,
                      uint32_t *nv, uint32_t *idx, uint32_t *nt);

/* ---- scene files (Model::LoadFile + Model::Scale, Common/Scenes/Model.cpp:210-381,647-709;
 *      ModelOBJ, Common/Scenes/Loaders/model_obj.cpp) */
/* Loads a ".obj" (faces fan-triangulated in file order, one vertex per
 * distinct position index + texture coordinate + normal, in order of first
 * use) or the reference's ".txt" model, then applies the reference's
 * normalisation (OBJ: centre and divide by the largest extent; both: fit into
 * a sphere of radius 0.49 about the origin) unless flags has RT_LOAD_RAW.
 * Call with xyz = idx = NULL to get the sizes in *nv / *nt; then with arrays
 * of at least those sizes.  Errors: rt_load_last_error() (this thread). */
enum { RT_LOAD_RAW = 1u << 0 };
int rt_load_model(const char *path, uint32_t flags, float *xyz, uint32_t *nv,
                  uint32_t *idx, uint32_t *nt);
const char *rt_load_last_error(void);

/* The same model with its shading attributes, as Model::LoadFile hands them
 * to the renderer.  Normals: the file's (".txt": normalised; ".obj": as
 * written, zero for corners without one) or, for an OBJ with no "vn" at all,
 * ModelOBJ::generateNormals of the raw positions (model_obj.cpp:541-616).
 * Materials: an OBJ's MTL file (mtllib, read as ModelOBJ::importMaterials,
 * model_obj.cpp:1119-1297) gives the table, and with more than one material
 * each triangle the one active at its face (usemtl; unknown names -> 0);
 * without materials, and for ".txt", one default MaterialData.
 * In: capacities in nv, nt, nmat; out: counts.  xyz = idx = NULL asks for
 * the counts only; normals, mat_ids and materials are optional outputs. */
typedef struct rt_model {
    float *xyz;             /* 3 per vertex */
    float *normals;         /* 3 per vertex */
    uint32_t *idx;          /* 3 per triangle */
    uint32_t *mat_ids;      /* 1 per triangle */
    rt_material *materials; /* nmat */
    uint32_t nv, nt, nmat;
} rt_model;
int rt_load_model_ex(const char *path, uint32_t flags, rt_model *model);

#ifdef __cplusplus
}
#endif

#endif /* RTCORE_H */
