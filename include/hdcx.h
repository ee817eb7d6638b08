/*
 * hdcx — C ABI of the MI355X spectral path-tracing core.
 *
 * This is the drop-in boundary that replaces the reference integrator class
 * hdcodex::VulkanPathTracer (include/hdcodex/gpu/vulkan_path_tracer.h:26-52).
 * Every entry point returns an int status (HDCX_OK = 0) and never throws;
 * the message of the last failure on the calling thread is returned by
 * hdcx_last_error(). The C++ wrapper in hdcx_path_tracer.hpp restores the
 * reference's exact method surface on top of this ABI.
 *
 * Threading: one tracer is driven by one thread at a time, like the
 * reference (one Hydra render thread). Distinct tracers are independent.
 */
#ifndef HDCX_H
#define HDCX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HDCX_OK 0
#define HDCX_ERR_INVALID_ARGUMENT 1
#define HDCX_ERR_INVALID_SCENE 2
#define HDCX_ERR_OUT_OF_MEMORY 3
#define HDCX_ERR_DEVICE 4
#define HDCX_ERR_NO_GEOMETRY 5
#define HDCX_ERR_UNSUPPORTED 6
#define HDCX_ERR_INTERNAL 7

/* Byte sizes of the packed records; identical to the reference's std430
 * structs GpuMaterial (src/gpu/vulkan_path_tracer.cpp:1907-1933) and
 * GpuLight (:1935-1952). */
#define HDCX_MATERIAL_BYTES 368
#define HDCX_LIGHT_BYTES 224
#define HDCX_MAX_LIGHTS 64
#define HDCX_MAX_TEXTURES 256
#define HDCX_MAX_BOUNCES 12
#define HDCX_MISSING_TEXTURE 0xffffffffu

/* Texel formats of hdcx_texture_view (reference CreateTexture :2215-2221). */
#define HDCX_TEXTURE_RGBA8_UNORM 0
#define HDCX_TEXTURE_RGBA8_SRGB 1
#define HDCX_TEXTURE_RGBA32F 2

/* Shading modes (include/hdcodex/core/shading_mode.h). */
#define HDCX_SHADING_FUSED 0
#define HDCX_SHADING_MODULAR 1
#define HDCX_SHADING_RASTER_PREVIEW 2

typedef struct hdcx_tracer hdcx_tracer;
typedef struct hdcx_scene hdcx_scene;

/* Pinhole camera, the reference's PathTracerCamera
 * (include/hdcodex/gpu/vulkan_path_tracer.h:16-23). Pixel row 0 is the
 * bottom row (v = 0 at lower_left). */
typedef struct hdcx_camera {
    float origin[3];
    float lower_left[3];
    float horizontal[3];
    float vertical[3];
} hdcx_camera;

typedef struct hdcx_texture_view {
    uint32_t width;
    uint32_t height;
    uint32_t format;     /* HDCX_TEXTURE_* */
    uint32_t reserved;
    const void* texels;  /* width*height*4 bytes (RGBA8) or floats (RGBA32F) */
} hdcx_texture_view;

/* A flattened, packed scene: exactly the buffers the reference uploads in
 * BuildScene (src/gpu/vulkan_path_tracer.cpp:3004-3030), in its order and
 * byte layout. All pointers are host memory owned by the caller for the
 * duration of hdcx_set_scene. */
/* Instancing (no counterpart in the reference, which flattens every mesh
 * into one BLAS under a one-instance TLAS, vulkan_path_tracer.cpp:3032-3125).
 * A prototype is a range of the view's triangles that is not drawn itself;
 * prototypes follow the drawn triangles and tile [drawn, triangle_count) in
 * order. An instance draws one prototype with a row-major 3x4
 * object-to-world transform and, unless material == 0xffffffff, one
 * material for all its triangles. Primitive ids are those of the scene the
 * reference would flatten: drawn triangles first, then instance i's
 * triangles at base_i + local, base_0 = drawn, base_{i+1} = base_i +
 * triangles of instance i's prototype. */
typedef struct hdcx_prototype {
    uint32_t first_triangle;
    uint32_t triangle_count;
} hdcx_prototype;

typedef struct hdcx_instance {
    float object_to_world[12]; /* rows: x' = m0 x + m1 y + m2 z + m3, ... */
    uint32_t prototype;
    uint32_t material;         /* 0xffffffff: the prototype triangles' own */
} hdcx_instance;

typedef struct hdcx_scene_view {
    uint32_t vertex_count;
    uint32_t triangle_count;
    const float* positions;            /* float4 per vertex, w = 1 */
    const uint32_t* indices;           /* 3 per triangle, world-flattened */
    const uint32_t* triangle_materials;/* 1 per triangle */
    const float* texcoords;            /* 6 per triangle (uv per corner) */
    const float* normals;              /* 12 per triangle (float4 per corner) */
    const float* emissive;             /* 4 per triangle: cdf, pdf/area, area, 0;
                                          NULL: built on the device (below) */
    uint32_t material_count;
    uint32_t light_count;              /* authored lights, <= HDCX_MAX_LIGHTS */
    const void* materials;             /* material_count * HDCX_MATERIAL_BYTES */
    const void* lights;                /* max(light_count,1) * HDCX_LIGHT_BYTES */
    uint32_t texture_count;
    uint32_t opaque;                   /* 1 iff every material is opaque */
    const hdcx_texture_view* textures;
    uint32_t prototype_count;          /* 0: no instancing */
    const hdcx_prototype* prototypes;
    uint32_t instance_count;
    const hdcx_instance* instances;
    /* Used only when emissive is NULL. Emission power per material
     * (max(emissionWeight, 0) x emission luminance, the reference's
     * materialEmissionPowers, vulkan_path_tracer.cpp:2873-2880), or NULL:
     * then derived from the records (luminance 0.2126 r + 0.7152 g +
     * 0.0722 b of max(emission, 0), or the mean texel luminance of the
     * emission texture, of a UDIM set the mean of its tiles' means). */
    const float* material_emission_power;
    /* Used only when emissive is NULL: triangles [0, n) are sampled as mesh
     * lights, the others are not (cdf 1, pdf 0); 0 = every drawn triangle. */
    uint32_t light_triangle_count;
    uint32_t reserved;
} hdcx_scene_view;

/* Light tables built on the device by hdcx_set_scene (no host pass):
 * - emissive == NULL: the emissive-triangle distribution of the reference's
 *   BuildScene (vulkan_path_tracer.cpp:2944-2993): weight = area x power of
 *   the triangle's material, area from the FP64 cross product, FP64 prefix
 *   sum of weight / total, the last sampled triangle's cdf 1; every entry 0
 *   when no weight is positive;
 * - a dome light record (textureInfo.z == 0) whose textureInfo.w equals
 *   textureInfo.x: the lat-long importance tables of its own texture
 *   (BuildLatLongImportanceTexture, :1998-2066: luminance x row solid
 *   angle, FP64 marginal and conditional CDFs, pdf = luminance / total);
 *   a dome with no positive weight is sampled uniformly (as if w were
 *   0xffffffff). Otherwise w names a host-built RGBA32F importance texture
 *   (r conditional CDF, g marginal CDF, b pdf), as the reference packs it.
 * The device sums run in a fixed tree order: deterministic, and within a
 * few FP64 ulps of the reference's sequential sums before the FP32 store. */

/* One query ray: origin, tmin, direction, tmax (32 bytes). */
typedef struct hdcx_ray {
    float origin[3];
    float tmin;
    float direction[3];
    float tmax;
} hdcx_ray;

/* Closest-hit record (16 bytes). primitive = flattened triangle index (for
 * instanced triangles the id in the flattened scene, see hdcx_instance),
 * or 0xffffffff on a miss; (u, v) are the barycentrics of vertices 1 and 2. */
typedef struct hdcx_hit {
    float t;
    uint32_t primitive;
    float u;
    float v;
} hdcx_hit;

/* Thread-local text of the last error on this thread ("" if none). */
const char* hdcx_last_error(void);

/* Tracer lifetime. Replaces VulkanPathTracer(VulkanContext&, ShaderCache&)
 * (vulkan_path_tracer.h:28) and ~VulkanPathTracer. */
int hdcx_create(int device_ordinal, hdcx_tracer** out_tracer);
int hdcx_destroy(hdcx_tracer* tracer);

/* Replaces VulkanPathTracer::SetScene (vulkan_path_tracer.h:34; body
 * src/gpu/vulkan_path_tracer.cpp:2604-3128): uploads the packed scene and
 * builds the BVH on the device. A NULL view, or a view without vertices or
 * triangles, clears the scene (HasGeometry() becomes false). */
int hdcx_set_scene(hdcx_tracer* tracer, const hdcx_scene_view* scene);

/* The checks hdcx_set_scene runs before touching the device: indices and
 * material ids in range, finite positions, texture references (plain and
 * UDIM) in range, <= 64 lights, host dome importance tables RGBA32F. Returns
 * HDCX_ERR_INVALID_SCENE with the first violation. Needs no device; the
 * reference has no such check (the driver consumes the buffers as given). */
int hdcx_validate_scene(const hdcx_scene_view* scene);

/* Replaces SetShadingMode / GetShadingMode (vulkan_path_tracer.h:35-36).
 * HDCX_SHADING_RASTER_PREVIEW is rejected with HDCX_ERR_INVALID_ARGUMENT as
 * the reference rejects it with std::invalid_argument (:2495-2498). */
int hdcx_set_shading_mode(hdcx_tracer* tracer, int mode);
int hdcx_get_shading_mode(const hdcx_tracer* tracer, int* out_mode);

/* Replaces VulkanPathTracer::Render (vulkan_path_tracer.h:39-45; body
 * :3230-3306). Traces sample_count progressive samples per pixel starting at
 * sample number sample_index and writes the running mean of samples
 * [0, sample_index + sample_count) as linear RGBA32F, alpha 1, row 0 at the
 * bottom, into out_rgba (width*height*4 floats, caller-owned). out_rgba may
 * be NULL: the film then stays on the device (hdcx_read_film reads it).
 * Returns HDCX_ERR_NO_GEOMETRY (writing nothing) when there is no geometry
 * or the size is zero, where the reference returns an empty vector. */
int hdcx_render(hdcx_tracer* tracer, const hdcx_camera* camera,
                uint32_t width, uint32_t height, uint32_t sample_index,
                uint32_t max_bounces, uint32_t sample_count, float* out_rgba);

/* Independent sample streams (no counterpart in the reference, whose random
 * seeds are fixed by the pixel and the sample number, :1198-1207): sample
 * number n of hdcx_render is seeded as n + offset, while the progressive
 * mean still counts samples from 0. Default 0, the reference's streams. */
int hdcx_set_sample_offset(hdcx_tracer* tracer, uint32_t offset);

/* Copies the current device film (RGBA32F, width*height*4) to out_rgba. */
int hdcx_read_film(hdcx_tracer* tracer, float* out_rgba);

/* The current film upscaled to width x height by nearest neighbour on the
 * device: replaces the render pass's host UpscaleNearest
 * (src/hydra/render_pass.cpp:45-68) of the interactive half-resolution
 * frame (:191-208), pixel for pixel (destination (x, y) takes source
 * (min(w0-1, x*w0/width), min(h0-1, y*h0/height)), unsigned 32-bit
 * arithmetic). out_rgba holds width*height*4 floats: host memory, or, with
 * out_on_device != 0, device memory of the tracer's device (written by the
 * kernel, no host copy). Sizes whose index products reach 2^32 are
 * rejected (HDCX_ERR_INVALID_ARGUMENT). */
int hdcx_read_film_upscaled(hdcx_tracer* tracer, uint32_t width, uint32_t height, float* out_rgba,
                            int out_on_device);

/* Replaces VulkanPathTracer::HasGeometry (vulkan_path_tracer.h:47).
 * Returns 1 or 0; 0 for a NULL tracer. */
int hdcx_has_geometry(const hdcx_tracer* tracer);

/* Closest-hit queries against the current scene: the reference's committed
 * ray query (rayQueryInitializeEXT ... rayQueryProceedEXT,
 * src/gpu/vulkan_path_tracer.cpp:1228-1268) exposed for batches of rays in
 * host memory. Opaque semantics (no any-hit confirmation). */
int hdcx_trace_closest(hdcx_tracer* tracer, const hdcx_ray* rays,
                       uint32_t count, hdcx_hit* out_hits);

/* Any-hit (occlusion) queries with terminate-on-first-hit semantics
 * (shadowVisibility / meshLightVisible, :855-912) in an opaque scene:
 * out_occluded[i] = 1 when any triangle lies in [tmin, tmax]. */
int hdcx_trace_occluded(hdcx_tracer* tracer, const hdcx_ray* rays,
                        uint32_t count, uint8_t* out_occluded);

/* ---- Procedural scenes (no counterpart in the reference: its scenes come
 * from Hydra). They produce the packed view of a SceneSnapshot through the
 * same packer the C++ wrapper uses. ---- */

/* name: "cornell" (34 triangles, one rect light, Lambert walls),
 * "heightfield" (param = grid vertices per side, e.g. 1024 -> 2,093,058
 * triangles; GGX dielectric over Burley; HDR lat-long dome of
 * param_dome_width x param_dome_width/2 texels), "triangle" (the
 * reference's one-triangle test scene), "materials" (every closure lobe,
 * light type and texture path), "furnace", "mixed", "mixed_lights",
 * "instanced", "instanced_small", and "irradiance" (param = case: a white
 * Lambert receiver under one light whose irradiance has a closed form, a
 * glass slab, or a closed sphere; listed in procedural.cpp). */
int hdcx_scene_create_procedural(const char* name, uint64_t seed,
                                 uint32_t param, uint32_t param_dome_width,
                                 hdcx_scene** out_scene);
int hdcx_scene_view_get(const hdcx_scene* scene, hdcx_scene_view* out_view);
int hdcx_scene_camera_get(const hdcx_scene* scene, uint32_t width,
                          uint32_t height, hdcx_camera* out_camera);
/* A copy of a procedural scene with every instance written out as ordinary
 * triangles in primitive-id order (the scene the reference would flatten;
 * test input for the CPU oracle). Same camera. */
int hdcx_scene_flatten(const hdcx_scene* scene, hdcx_scene** out_scene);
int hdcx_scene_destroy(hdcx_scene* scene);

/* ---- BVH (no counterpart in the reference: its acceleration structure is
 * built inside the Vulkan driver, :3032-3125). ---- */

/* BVH builders. The reference leaves acceleration-structure builds to the
 * Vulkan driver (src/gpu/vulkan_path_tracer.cpp:3032-3125, PREFER_FAST_TRACE
 * BLAS + TLAS); all three produce the same four-wide layout, so traversal
 * and hit results do not depend on the choice. */
#define HDCX_BVH_HOST_SAH 0    /* binned SAH (16 bins) on the host */
#define HDCX_BVH_DEVICE_LBVH 1 /* on the GPU: Morton radix tree under binned-SAH top levels */
#define HDCX_BVH_SERIAL_LBVH 2 /* the device builder's steps run serially on the host */

/* Layout summary of a four-wide BVH: 128-byte nodes, 48-byte triangle
 * slots (three float4 vertices, primitive id in the first w). */
typedef struct hdcx_bvh_info {
    uint32_t node_count;
    uint32_t slot_count;
    uint32_t max_stack;    /* worst-case traversal stack entries */
    uint32_t depth;
    int32_t builder;       /* HDCX_BVH_* that produced it */
    uint32_t top_clusters; /* LBVH: subtrees under the SAH top levels (0: none) */
    uint32_t refits;       /* refits applied since the last build */
    uint32_t reserved;
    double build_ms;       /* wall time of the last build or refit */
} hdcx_bvh_info;

/* Builds the host (binned-SAH) BVH of a scene view and validates it. Call
 * with out_nodes = out_slots = NULL to get the sizes, then again with
 * buffers of node_count*128 bytes and slot_count*12 floats. */
int hdcx_bvh_build_host(const hdcx_scene_view* scene, hdcx_bvh_info* out_info,
                        void* out_nodes, float* out_slots);

/* As hdcx_bvh_build_host with a choice of builder (HDCX_BVH_*);
 * HDCX_BVH_DEVICE_LBVH runs on the current HIP device. top_clusters: LBVH
 * subtrees under the SAH top levels (0 = radix tree only). The result is
 * validated on the host (every triangle once, leaves in range, stack fits). */
int hdcx_bvh_build(const hdcx_scene_view* scene, int builder, uint32_t top_clusters,
                   hdcx_bvh_info* out_info, void* out_nodes, float* out_slots);

/* Tracer BVH policy. refit != 0: hdcx_set_scene with the same triangles
 * (vertex count and index buffer unchanged, e.g. skinned points) refits the
 * existing BVH on the device instead of rebuilding it. Defaults:
 * HDCX_BVH_DEVICE_LBVH, 2048 top clusters, refit on. */
typedef struct hdcx_bvh_options {
    int32_t builder;
    uint32_t top_clusters;
    uint32_t refit;
    uint32_t reserved;
} hdcx_bvh_options;

int hdcx_set_bvh_options(hdcx_tracer* tracer, const hdcx_bvh_options* options);
int hdcx_get_bvh_info(hdcx_tracer* tracer, hdcx_bvh_info* out_info);

/* Inspection of the current scene's light tables (no counterpart in the
 * reference, whose tables are host vectors). hdcx_read_emissive_table
 * copies the table the device samples (4 floats per triangle of the view,
 * count = the view's triangle_count). hdcx_read_dome_importance writes the
 * importance-table size of light `light` (0 x 0: sampled uniformly) and,
 * for each non-NULL pointer, its marginal CDF (height floats), conditional
 * CDFs (width x height) and pdf (width x height). */
int hdcx_read_emissive_table(hdcx_tracer* tracer, float* out, uint32_t count);
int hdcx_read_dome_importance(hdcx_tracer* tracer, uint32_t light, uint32_t* out_width, uint32_t* out_height,
                              float* out_marginal, float* out_conditional, float* out_pdf);

/* ---- Batched multi-scene rendering (no counterpart in the reference: one
 * VulkanPathTracer renders one scene per Render call). Many independent
 * scenes, each with its own BVH, camera, image size and sample count,
 * render together: their pixels are cut into tiles (all of a pixel's
 * samples in one tile) and packed into waves of at most wave_paths lane
 * paths; one wave runs raygen -> (extend -> shade -> shadow) x bounces ->
 * film as single launches over every scene at once (each path carries its
 * scene index), and the per-wave launch sequence is captured once in a
 * hipGraph and replayed for every wave (the wave index is a device counter).
 * Each scene's film follows hdcx_render's progressive-mean semantics. */
typedef struct hdcx_batch hdcx_batch;

typedef struct hdcx_batch_tuning {
    uint32_t wave_paths;  /* lane paths per wave (path-pool size) */
    uint32_t tile_pixels; /* pixels per tile (fewer when samples * 3 * tile > wave_paths) */
    uint32_t max_blocks;  /* workgroups per stage launch (128 threads each) */
    uint32_t use_graph;   /* nonzero: replay the captured per-wave graph */
} hdcx_batch_tuning;

typedef struct hdcx_batch_stats {
    uint64_t pixel_samples;  /* last hdcx_batch_render */
    uint64_t lane_paths;
    uint32_t waves;
    uint32_t graph_launches;
    double milliseconds;     /* wall time of the last hdcx_batch_render (device work incl.) */
} hdcx_batch_stats;

int hdcx_batch_create(int device_ordinal, hdcx_batch** out_batch);
int hdcx_batch_destroy(hdcx_batch* batch);
/* Uploads a scene (as hdcx_set_scene, with the BVH builder from options or
 * the default when NULL) and returns its index. */
int hdcx_batch_add_scene(hdcx_batch* batch, const hdcx_scene_view* scene, const hdcx_bvh_options* options,
                         uint32_t* out_index);
/* Camera, image size and samples per pixel per hdcx_batch_render call. */
int hdcx_batch_set_view(hdcx_batch* batch, uint32_t scene, const hdcx_camera* camera, uint32_t width,
                        uint32_t height, uint32_t samples);
int hdcx_batch_set_tuning(hdcx_batch* batch, const hdcx_batch_tuning* tuning);
int hdcx_batch_get_tuning(hdcx_batch* batch, hdcx_batch_tuning* out_tuning);
/* Times candidate tunings on one sample per pixel of every view and keeps
 * the fastest. With cache_path, a result stored there for the same device
 * and workload is reused, and a new one is appended. The films hold the
 * trial renders afterwards (render at sample_index 0 to restart them). */
int hdcx_batch_autotune(hdcx_batch* batch, uint32_t max_bounces, const char* cache_path,
                        hdcx_batch_tuning* out_tuning);
/* Renders samples [sample_index, sample_index + samples) of every view. */
int hdcx_batch_render(hdcx_batch* batch, uint32_t sample_index, uint32_t max_bounces);
/* Running mean of a scene's film: width * height RGBA32F, row 0 bottom. */
int hdcx_batch_read_film(hdcx_batch* batch, uint32_t scene, float* out_rgba);
int hdcx_batch_get_stats(hdcx_batch* batch, hdcx_batch_stats* out_stats);

/* ---- Radiance cache (no counterpart in the reference: its neural
 * rendering plan is unimplemented, docs/neural-rendering-plan.md:32-48).
 * An MLP (128 encoded inputs -> 5 x 64 ReLU -> 3) trained online on the
 * path tracer's own radiance estimates predicts the radiance scattered from
 * a surface point at the three wavelengths of a hero triple; paths
 * terminate into it at a fixed vertex. BF16 on MFMA, FP32 accumulation. ---- */

#define HDCX_RC_PARAMS 24768

/* One cache query or training sample (48 bytes). */
typedef struct hdcx_rc_query {
    float position[3];  /* normalised to the scene bounds, [0,1]^3 */
    float direction[3]; /* unit outgoing direction */
    float lambda0;      /* base wavelength of the hero triple (lane 0), nm */
    float roughness;
    float albedo;
    uint32_t lane;      /* 0..2: which output the sample trains / the path uses */
    uint32_t path;      /* path id (ignored by hdcx_rc_infer) */
    float target;       /* training target radiance of that lane */
} hdcx_rc_query;

typedef struct hdcx_radiance_cache_config {
    uint32_t enabled;
    uint32_t train_samples;    /* training samples gathered per frame (default 65536) */
    uint32_t terminate_bounce; /* 0-based path vertex answered by the cache (default 2) */
    uint32_t train_steps;      /* Adam steps (mini-batches) per frame (default 4) */
    float learning_rate;       /* default 1e-2 */
    float beta1;               /* default 0.9 */
    float beta2;               /* default 0.99 */
    float epsilon;             /* default 1e-8 */
    float ema_decay;           /* inference uses an exponential moving average of the
                                * weights with this decay per Adam step, in [0, 1),
                                * started from the initial weights; 0 = the current
                                * weights (default) */
    float spread_threshold;    /* a path stops into the cache at the first vertex from
                                * terminate_bounce on where its spread exceeds this
                                * times its primary footprint (Mueller et al. 2021);
                                * <= 0: at terminate_bounce unconditionally */
    uint64_t seed;             /* weight initialisation */
    float calibration_decay;   /* in (0, 1): cached radiance is scaled by the ratio of the
                                * training targets' sum to the network's outputs on the same
                                * samples, summed over frames with this decay per frame
                                * (the plain estimator calibrates the cache's brightness);
                                * 0 = off (default). hdcx_render only; hdcx_rc_infer
                                * returns the raw network outputs */
    uint32_t flags;            /* HDCX_RC_* bits */
} hdcx_radiance_cache_config;

/* Training pixel samples keep their plain radiance in the image (by default
 * they show the cached value at the terminate vertex, like every other
 * path, and their plain rest of the path only trains the cache). */
#define HDCX_RC_PLAIN_TRAINING_PATHS 1u

typedef struct hdcx_rc_stats {
    uint64_t frames;
    uint64_t queries;          /* cache queries answered */
    uint64_t train_samples;    /* samples trained on */
    uint64_t adam_steps;
    float last_loss;           /* mean relative L2 loss of the last step */
    float calibration_scale;   /* current brightness calibration factor (1 when off) */
    double infer_ms;           /* HIP-event time of inference launches (when timing is on) */
    double train_ms;
} hdcx_rc_stats;

/* Enables/disables the cache for hdcx_render and (re)initialises the
 * network when the config's seed changes. NULL disables. */
int hdcx_set_radiance_cache(hdcx_tracer* tracer, const hdcx_radiance_cache_config* config);
/* FP32 master weights, row-major W0 [64][128], W1..W4 [64][64], W5 [3][64]. */
int hdcx_rc_get_params(hdcx_tracer* tracer, float* out_params);
int hdcx_rc_set_params(hdcx_tracer* tracer, const float* params);
/* Network outputs (3 per query) for host-side queries, with the inference
 * weights (the moving average when ema_decay > 0). */
int hdcx_rc_infer(hdcx_tracer* tracer, const hdcx_rc_query* queries, uint32_t count, float* out);
/* Gradient of the mean relative L2 loss over the samples, without an
 * update; out_loss receives the mean loss. */
int hdcx_rc_gradient(hdcx_tracer* tracer, const hdcx_rc_query* samples, uint32_t count, float* out_grad,
                     float* out_loss);
/* hdcx_rc_gradient for callers that cannot allocate (code under HIP graph
 * capture, tight serving or training loops, harnesses that forbid
 * hipMalloc); no counterpart in the reference. samples, out_grad
 * (HDCX_RC_PARAMS floats) and out_loss (one float, may be NULL) are DEVICE
 * pointers of the tracer's device. The training workspace lives in a static
 * device array of about 8 MB that each GPU holds until the process ends, so
 * a call takes at most 4096 samples (more: HDCX_ERR_INVALID_ARGUMENT); the
 * cache must already have a network (hdcx_set_radiance_cache). Calls on one
 * GPU share that array: calls that overlap on one GPU (several threads or
 * tracers) race. Synchronous, like hdcx_rc_gradient. */
int hdcx_rc_gradient_noalloc(hdcx_tracer* tracer, const hdcx_rc_query* samples, uint32_t count, float* out_grad,
                             float* out_loss);
/* `steps` Adam steps over consecutive mini-batches of the samples. */
int hdcx_rc_train(hdcx_tracer* tracer, const hdcx_rc_query* samples, uint32_t count, uint32_t steps,
                  float* out_loss);
int hdcx_rc_get_stats(hdcx_tracer* tracer, hdcx_rc_stats* out_stats);

/* Radiance caches of a batch: one network per scene, as the tracer's (no
 * counterpart in the reference). config->train_samples counts samples per
 * scene and render; scene i's weights start from seed + i. The networks are
 * (re)initialised when scenes were added or the seed changed; each
 * hdcx_batch_render trains them, after its waves, on the samples it
 * gathered. NULL disables. */
int hdcx_batch_set_radiance_cache(hdcx_batch* batch, const hdcx_radiance_cache_config* config);
typedef struct hdcx_batch_rc_stats {
    uint64_t queries;        /* cache queries answered by the last render */
    uint64_t train_samples;  /* samples trained on, all renders */
    uint64_t adam_steps;     /* all renders, all scenes */
    double train_ms;         /* wall time of training, all renders */
    float mean_loss;         /* mean relative L2 loss of the last render's steps */
    uint32_t scenes_trained; /* scenes with samples in the last render */
} hdcx_batch_rc_stats;
int hdcx_batch_get_rc_stats(hdcx_batch* batch, hdcx_batch_rc_stats* out_stats);

/* ---- Measurement (no counterpart in the reference, which has no GPU
 * timing: docs/architecture.md:255-257). ---- */

#define HDCX_KERNEL_RAYGEN 0
#define HDCX_KERNEL_EXTEND 1
#define HDCX_KERNEL_SHADE 2
#define HDCX_KERNEL_SHADOW 3
#define HDCX_KERNEL_FILM 4
#define HDCX_KERNEL_CACHE 5 /* radiance-cache inference */
#define HDCX_KERNEL_WALK 6  /* subsurface random walks deferred by shade */
#define HDCX_KERNEL_PAIR 7  /* no launches: the lane pairing runs inside shade (slot kept for the layout) */
#define HDCX_KERNEL_MISS 8  /* environment at misses; queues the hits for shade */
#define HDCX_KERNEL_COUNT 9

typedef struct hdcx_kernel_stats {
    uint64_t launches[HDCX_KERNEL_COUNT];
    double milliseconds[HDCX_KERNEL_COUNT]; /* HIP-event time summed over launches */
    uint64_t rays[HDCX_KERNEL_COUNT];       /* work items processed: extend counts the
                                               closest-hit rays traced (lanes of a pixel
                                               sample with equal rays trace them once),
                                               miss the path vertices, shade the
                                               vertices on a surface */
    uint64_t nodes_visited;                 /* counted only when counting is on */
    uint64_t triangles_tested;
    uint64_t extend_rays_counted;
    uint64_t shadow_rays_counted;
    uint64_t shadow_nodes_visited;
    uint64_t shadow_triangles_tested;
    uint64_t node_bytes;                    /* bytes per BVH node record */
    uint64_t triangle_bytes;                /* bytes per leaf triangle record */
    uint64_t bvh_nodes;
    uint64_t bvh_triangles;
    uint64_t extend_hits_counted;           /* paths whose closest-hit query found a surface
                                               (shade vertices that are not misses); counted
                                               only when counting is on */
    uint64_t shadow_queued_paths;           /* paths that shade gave at least one shadow ray */
} hdcx_kernel_stats;

/* flags bit 0: time every kernel launch with HIP events on the launch
 * stream; bit 1: run the traversal kernels compiled with node/triangle
 * counters (a separate instantiation; the counting build is slower and is
 * used only to measure algorithmic bytes); bit 2: run subsurface walks
 * inside the shade kernel instead of deferring them to the walk kernel
 * (the same results; kept to check the split against); bit 3: every lane
 * traces its own rays (by default the lanes of a pixel sample trace an
 * equal ray once: the same results; kept to check the sharing against);
 * bit 4: shade scenes whose materials have no coat, sheen, transmission,
 * subsurface or translucency with the full shading kernel (by default a
 * variant without those lobes runs: the same results; kept to check the
 * variant against). */
int hdcx_set_instrumentation(hdcx_tracer* tracer, uint32_t flags);
int hdcx_get_kernel_stats(hdcx_tracer* tracer, hdcx_kernel_stats* out_stats);
int hdcx_reset_kernel_stats(hdcx_tracer* tracer);

#ifdef __cplusplus
}
#endif

#endif /* HDCX_H */
