/*
 * ornament_hip.h -- the thin C-ABI between the ornament host layer (scene,
 * camera, state, BVH build) and the MI355X (gfx950) HIP kernels.
 *
 * This is the boundary a Zig `HipMi355xPathTracer` binds with @cImport in
 * place of the reference's WebGPU driver (src/wgpu_backend/path_tracer.zig)
 * and HIP driver (src/hip_backend/path_tracer.zig). Plain C: opaque handles,
 * plain pointers and sizes, every function returns an orn_status.
 *
 * Status codes: 0 is success; runtime failures carry the hipError_t value
 * they came from (the numbering of CuintHipError, src/hip_backend/hip.zig:7-62,
 * e.g. 1 InvalidValue, 2 OutOfMemory, 100 NoDevice); codes >= 9000 are
 * ornament's own (an empty scene panics in the reference, bvh.zig:35-37).
 *
 * Threading: one handle may be used by one host thread at a time. Calls are
 * synchronous unless their name ends in _async, like the reference, which
 * blocks on wgpuDevicePoll(wait=true) after every dispatch
 * (wgpu_backend/path_tracer.zig:248-249).
 */
#ifndef ORNAMENT_HIP_H
#define ORNAMENT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes                                                        */
/* ------------------------------------------------------------------ */
typedef int orn_status;
#define ORN_OK 0
#define ORN_ERROR_INVALID_VALUE 1          /* hipErrorInvalidValue */
#define ORN_ERROR_OUT_OF_MEMORY 2          /* hipErrorOutOfMemory */
#define ORN_ERROR_NO_DEVICE 100            /* hipErrorNoDevice */
#define ORN_ERROR_INVALID_DEVICE 101       /* hipErrorInvalidDevice */
#define ORN_ERROR_EMPTY_SCENE 9001         /* bvh.zig:35-37 panics */
#define ORN_ERROR_EMPTY_BVH 9002           /* bvh.zig:222-223, 273-274 panic */
#define ORN_ERROR_INVALID_SCENE 9003       /* an index in the arrays is out of range */
#define ORN_ERROR_BVH_TOO_DEEP 9004        /* traversal stack > 64 (pathtracer.wgsl:20) */
#define ORN_ERROR_UNSUPPORTED_TEXTURE 9005 /* wgpu buffers.zig:316-336 asserts */
#define ORN_ERROR_INVALID_HANDLE 9006      /* null or destroyed handle */
#define ORN_ERROR_SIZE_MISMATCH 9007       /* destination smaller than the target */
#define ORN_ERROR_IO 9008                  /* a file could not be opened, read or written */
#define ORN_ERROR_INVALID_FILE 9009        /* a scene or image file is malformed */
#define ORN_ERROR_SCRATCH_TOO_SMALL 9010   /* image larger than orn_render_noalloc's scratch */

/* Human-readable name of a status ("InvalidValue", "EmptyScene", ...). */
const char *orn_status_string(orn_status status);

/* ------------------------------------------------------------------ */
/* Binary layouts: identical to the reference's extern structs          */
/* ------------------------------------------------------------------ */

/* BvhNode, gpu_structs.zig:22-31 (64 bytes). node_type: 0 internal,
 * 1 sphere, 2 mesh, 3 triangle (gpu_structs.zig:15-20). */
typedef struct orn_bvh_node {
    float lmin_or_v0[3];
    uint32_t left_or_custom_id;
    float lmax_or_v1[3];
    uint32_t right_or_material_index;
    float rmin_or_v2[3];
    uint32_t node_type;
    float rmax_or_v3[3];
    uint32_t transform_id;
} orn_bvh_node;

/* Material, gpu_structs.zig:33-59 (32 bytes). type: 0 Lambertian, 1 Metal,
 * 2 Dielectric, 3 DiffuseLight (material.zig:4-9). */
typedef struct orn_material {
    float albedo[3];
    uint32_t albedo_texture_index; /* 0xFFFFFFFF: plain colour */
    float fuzz;
    float ior;
    uint32_t type;
    uint32_t padding;
} orn_material;

/* Camera, gpu_structs.zig:88-117 (112 bytes) */
typedef struct orn_camera {
    float origin[3];
    float lens_radius;
    float lower_left_corner[3];
    uint32_t padding0;
    float horizontal[3];
    uint32_t padding1;
    float vertical[3];
    uint32_t padding2;
    float u[3];
    uint32_t padding3;
    float v[3];
    uint32_t padding4;
    float w[3];
    uint32_t padding5;
} orn_camera;

/* ConstantParams, gpu_structs.zig:61-86 (144 bytes) */
typedef struct orn_constant_params {
    orn_camera camera;
    uint32_t depth;
    uint32_t width;
    uint32_t height;
    uint32_t flip_y;
    float inverted_gamma;
    float ray_cast_epsilon;
    uint32_t textures_count;
    float current_iteration;
} orn_constant_params;

/* One texture as ornament.Texture holds it (texture.zig:4-38). Supported
 * formats (wgpu_backend/buffers.zig:316-336): 8-bit unorm with 1, 2 or 4
 * components, and 16-bit float (is_hdr) with 1, 2 or 4 components. Width and
 * height are at most ORN_MAX_TEXTURE_DIMENSION, WebGPU's default
 * maxTextureDimension2D (the reference requests its device without raised
 * limits, wgpu_backend/device_state.zig:78-82); larger textures are rejected
 * with ORN_ERROR_UNSUPPORTED_TEXTURE. */
#define ORN_MAX_TEXTURE_DIMENSION 8192u
typedef struct orn_texture_desc {
    uint32_t width;
    uint32_t height;
    uint32_t num_components;
    uint32_t bytes_per_component;
    uint32_t is_hdr;
    uint32_t padding;
    const uint8_t *data; /* width * height * num_components * bytes_per_component */
} orn_texture_desc;

/* The arrays Bvh.init produces (bvh.zig:22-30), as the WGSL megakernel binds
 * them (pathtracer.wgsl:7-17). The library copies everything; the caller
 * keeps ownership of the host arrays. */
typedef struct orn_scene_desc {
    const orn_bvh_node *tlas_nodes;
    size_t tlas_nodes_len;
    const orn_bvh_node *blas_nodes;
    size_t blas_nodes_len;
    const float *normals; /* 4 floats per normal (Normal, gpu_structs.zig:8) */
    size_t normals_len;
    const uint32_t *normal_indices;
    size_t normal_indices_len;
    const float *uvs; /* 2 floats per uv */
    size_t uvs_len;
    const uint32_t *uv_indices;
    size_t uv_indices_len;
    const float *transforms; /* 16 floats per matrix: [inverse, model] per leaf */
    size_t transforms_len;
    const orn_material *materials;
    size_t materials_len;
    const orn_texture_desc *textures;
    size_t textures_len;
} orn_scene_desc;

/* Closest hit of one ray, the outputs of bvh_hit (bvh.wgsl:14-21). */
typedef struct orn_hit {
    uint32_t hit;
    float t;
    uint32_t material_index;
    uint32_t node_type; /* 1 sphere, 2 mesh */
    uint32_t inverted_transform_id;
    uint32_t tri_id; /* 3 * global triangle index (mesh hits; 0 for spheres) */
    float uv[2];     /* barycentrics (mesh hits; 0 for spheres) */
} orn_hit;

typedef struct orn_ctx orn_ctx;
typedef struct orn_device_scene orn_device_scene;
typedef struct orn_target orn_target;

/* ------------------------------------------------------------------ */
/* Device                                                              */
/* ------------------------------------------------------------------ */

/* Number of visible HIP devices. */
orn_status orn_device_count(int *count);

/* Open a device and its stream. device < 0 selects the last device, as the
 * reference's HIP driver does (hip_backend/path_tracer.zig:40-41). Replaces
 * DeviceState.init (wgpu_backend/device_state.zig:18-92). */
orn_status orn_device_open(int device, orn_ctx **out);
orn_status orn_device_close(orn_ctx *ctx);
/* Device name and compute-unit count, for logs (hip.zig:239-252). */
orn_status orn_device_info(orn_ctx *ctx, char *name, size_t name_len, int *compute_units);

/* ------------------------------------------------------------------ */
/* Scene upload: replaces buffers.Storage/Textures.init                 */
/* (wgpu_backend/path_tracer.zig:55-66, buffers.zig:113-336)            */
/* ------------------------------------------------------------------ */

/* Validates every index (node links, leaf transform/material ids, normal and
 * uv indices, texture ids) so that no kernel can read out of bounds, then
 * copies the arrays to HBM. Empty arrays are allowed except the TLAS. */
orn_status orn_scene_upload(orn_ctx *ctx, const orn_scene_desc *desc, orn_device_scene **out);
/* The validation orn_scene_upload performs, without a device: returns the
 * status upload would return for these arrays and, on success, the
 * traversal stack slots the BVH needs. */
orn_status orn_scene_validate(const orn_scene_desc *desc, uint32_t *stack_slots);
orn_status orn_scene_destroy(orn_device_scene *scene);
/* Besides the reference's BVH2 (bvh.zig), the upload derives a BVH4 from it
 * (each node holds a BVH2 node's grandchildren, in the same order), which the
 * kernels traverse when asked to (orn_tuning.bvh_width, and
 * orn_scene_set_trace_width for ray queries).
 * Traversal stack slots the BVH of the current trace width needs. */
orn_status orn_scene_stack_depth(orn_device_scene *scene, uint32_t *depth);
/* Node width (2 or 4) orn_trace_rays traverses (default 2). */
orn_status orn_scene_set_trace_width(orn_device_scene *scene, uint32_t width);

/* ------------------------------------------------------------------ */
/* Render target: replaces buffers.Target (wgpu_backend/buffers.zig:11-98) */
/* ------------------------------------------------------------------ */

/* Framebuffer and accumulation (float4 per pixel, zero-initialised) and the
 * per-pixel RNG state, seeded index + seed * width * height (seed 0 is the
 * reference's seeding, buffers.zig:25-30). At most 2^28 pixels per target
 * (and per batch, all images together). */
orn_status orn_target_create(orn_device_scene *scene, uint32_t width, uint32_t height,
                             uint32_t seed, orn_target **out);
orn_status orn_target_destroy(orn_target *target);

/* One PathTracer.render() (wgpu_backend/path_tracer.zig:252-265): `params`
 * carries State.current_iteration as it is BEFORE the call; the kernel
 * advances it by max(iterations, 1), accumulates one sample per pixel per
 * iteration and writes the post-processed framebuffer once at the end.
 * Blocks until the GPU is done. */
orn_status orn_render(orn_target *target, const orn_constant_params *params, uint32_t iterations);
/* Same, without waiting; orn_target_wait blocks until it has finished. */
orn_status orn_render_async(orn_target *target, const orn_constant_params *params,
                            uint32_t iterations);
orn_status orn_target_wait(orn_target *target);
/* Launch tuning of a target's render kernel; any value gives the same image.
 * regroup_lanes (1..64): a wave interrupts BVH traversal for a shading round
 *   once this many of its 64 lanes have finished their ray (default 48).
 * workgroups_per_cu (0..64): persistent 256-lane workgroups launched per
 *   compute unit; 0 launches as many as can be resident at once.
 * bvh_width (2 or 4): traverse the reference's BVH2 (default) or the derived
 *   BVH4. The BVH4 visits a superset of the reference's nodes, so it returns
 *   the reference's closest hit except where two primitives tie within
 *   rounding at a box boundary (2 of 2,073,600 pixels of a 1080p 32-spp render
 *   of 1,024 instanced icospheres differ, by < 2e-7 relative).
 * leaf_regroup_lanes (1..64): lanes standing at a BVH leaf (sphere, triangle,
 *   mesh entry) wait until this many of the wave's lanes do, then test their
 *   leaves together (default 16).
 * split_percent (0..100), samples_per_job (0..65535): the last split_percent
 *   percent of each XCD's pixels are rendered as jobs of samples_per_job
 *   consecutive samples, handed from lane to lane through memory, so that no
 *   lane is left with a whole pixel's samples at the end of a render (default
 *   0 and 16; either 0: one job per pixel). Renders whose current_iteration
 *   is not a whole number below 2^24 - iterations run one job per pixel.
 * node_steps (0..16): internal BVH nodes a traversing lane visits in a row
 *   before its wave looks for leaf and shading rounds again; fewer checks
 *   cost fewer instructions, while lanes that reach a leaf or finish their
 *   ray early wait longer (0 takes the default, 4).
 * workgroup_pool (0 or 1): 1 pools the rays at internal nodes of a
 *   workgroup's 4 waves into as few waves as hold them before each run of
 *   node steps (state through LDS, two barriers per run), and shades with all
 *   4 waves at once, triggered when regroup_lanes x 4 of the workgroup's lanes
 *   wait. It applies to scenes of spheres only with the BVH2 and is ignored
 *   elsewhere (default 0).
 * cost_order (0 or 1): 1 queues each XCD's pixels most expensive first, by
 *   what they cost in the target's (or batch's) previous render with
 *   cost_order 1 (GPU time of their jobs, in quarter-octave classes; tile
 *   order within a class), so that the cheapest pixels are the last ones
 *   drawn and the end of the render, where lanes run out of work, is short
 *   (longest processing time first). Every such render records the costs
 *   for the next. A target's first one, with no costs recorded yet, renders
 *   its first eighth of the samples in tile order to record them and the
 *   rest queued by them (two launches with the same result as one; a batch's
 *   first replay, or a render whose current_iteration is not a whole number
 *   below 2^24 - iterations, queues in tile order). orn_render_noalloc
 *   ignores it (default 0). */
typedef struct orn_tuning {
    uint32_t regroup_lanes;
    uint32_t workgroups_per_cu;
    uint32_t bvh_width;
    uint32_t leaf_regroup_lanes;
    uint32_t samples_per_job;
    uint32_t split_percent;
    uint32_t node_steps;
    uint32_t workgroup_pool;
    uint32_t cost_order;
} orn_tuning;
orn_status orn_target_set_tuning(orn_target *target, const orn_tuning *tuning);
orn_status orn_target_get_tuning(orn_target *target, orn_tuning *tuning);

/* A render that allocates nothing. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). It does what
 * orn_render_async does for a target, on image state the caller owns, all
 * device pointers on the scene's device in the layouts orn_get_* return:
 * `accumulation` (width * height float4, read and written), `rng`
 * (width * height uint32, read and written) and `framebuffer`
 * (width * height float4, written); width and height come from `params`.
 * The launch's scratch (the job-queue counters and one stage word per pixel
 * for split pixels) is a library-owned device array, one per device, sized
 * for ORN_RENDER_SCRATCH_PIXELS pixels (8 MiB) and held until the process
 * ends; larger images return ORN_ERROR_SCRATCH_TOO_SMALL before anything is
 * enqueued. `tuning` may be NULL (the default tuning). `stream` is a
 * hipStream_t of the scene's device, or NULL for the device context's own
 * stream. The call returns once the work is enqueued. Every call on a device
 * uses the same scratch: calls whose work overlaps in time on one device (on
 * different streams) race; calls on one stream are ordered and do not. */
#define ORN_RENDER_SCRATCH_PIXELS (1u << 21)
orn_status orn_render_noalloc(orn_device_scene *scene, const orn_constant_params *params, uint32_t iterations,
                              const orn_tuning *tuning, float *accumulation, uint32_t *rng, float *framebuffer,
                              void *stream);

/* GPU time of the most recent render's kernels (HIP events on the target's
 * stream), in milliseconds. */
orn_status orn_target_last_render_ms(orn_target *target, float *ms);

/* Readback (Target.getFrameBuffer, wgpu_backend/buffers.zig:67-91):
 * `pixels` is the number of float4 entries dst can hold. */
orn_status orn_get_framebuffer(orn_target *target, float *dst, size_t pixels);
orn_status orn_get_accumulation(orn_target *target, float *dst, size_t pixels);
orn_status orn_get_rng_state(orn_target *target, uint32_t *dst, size_t pixels);
/* Per-pixel cost recorded by the target's latest render with tuning
 * cost_order 1: the GPU time of the pixel's jobs, in ticks of the 100 MHz
 * device clock (width * height entries, row-major like the accumulation).
 * No counterpart in the reference; it is what orders the next render's
 * queues. ORN_ERROR_INVALID_VALUE if no render has recorded costs. */
orn_status orn_get_pixel_costs(orn_target *target, uint32_t *dst, size_t pixels);
/* Restore a saved accumulation + RNG state (checkpoint / resume of a
 * progressive render; either pointer may be NULL to keep the current one). */
orn_status orn_set_state(orn_target *target, const float *accumulation, const uint32_t *rng,
                         size_t pixels);

/* ------------------------------------------------------------------ */
/* Batches: many images of many scenes in one persistent launch         */
/* ------------------------------------------------------------------ */

/* A batch renders n images, each of its own scene, resolution and RNG seed,
 * in one persistent kernel launch replayed from a captured HIP graph. Each
 * image's result is the one a separate target would give for the same
 * scene, seed, params and iterations (bit for bit): a batch is the
 * reference's render() (wgpu_backend/path_tracer.zig:252-265) run for n
 * PathTracers at once, and fills the GPU where one small image cannot. */
typedef struct orn_batch orn_batch;
typedef struct orn_batch_item {
    const orn_scene_desc *scene; /* copied at create time */
    uint32_t width;
    uint32_t height;
    uint32_t seed; /* as in orn_target_create */
    uint32_t reserved;
} orn_batch_item;
/* Validates every scene as orn_scene_upload does (the first failing status
 * is returned) and uploads them as one device scene. */
orn_status orn_batch_create(orn_ctx *ctx, const orn_batch_item *items, size_t n, orn_batch **out);
orn_status orn_batch_destroy(orn_batch *batch);
/* One render() of every image: params[i] and iterations[i] mean what they
 * mean for orn_render of image i (params[i].width/height must match it). */
orn_status orn_batch_render(orn_batch *batch, const orn_constant_params *params,
                            const uint32_t *iterations);
orn_status orn_batch_render_async(orn_batch *batch, const orn_constant_params *params,
                                  const uint32_t *iterations);
orn_status orn_batch_wait(orn_batch *batch);
orn_status orn_batch_set_tuning(orn_batch *batch, const orn_tuning *tuning);
orn_status orn_batch_get_tuning(orn_batch *batch, orn_tuning *tuning);
orn_status orn_batch_last_render_ms(orn_batch *batch, float *ms);
orn_status orn_batch_get_framebuffer(orn_batch *batch, size_t item, float *dst, size_t pixels);
orn_status orn_batch_get_accumulation(orn_batch *batch, size_t item, float *dst, size_t pixels);
orn_status orn_batch_get_rng_state(orn_batch *batch, size_t item, uint32_t *dst, size_t pixels);
orn_status orn_batch_set_state(orn_batch *batch, size_t item, const float *accumulation,
                               const uint32_t *rng, size_t pixels);

/* ------------------------------------------------------------------ */
/* Autotuning (no counterpart in the reference, which has no launch     */
/* parameters to tune)                                                   */
/* ------------------------------------------------------------------ */

/* Let the autotuner consider the BVH4 (bvh_width 4), whose closest hits may
 * differ from the reference's in rare rounding ties (see orn_tuning). */
#define ORN_TUNE_ALLOW_BVH4 1u
/* Re-measure even when the cache holds an entry for this scene class. */
#define ORN_TUNE_IGNORE_CACHE 2u

/* Times renders of `iterations` samples per pixel (a few suffice) under
 * candidate tunings (coordinate descent over regroup_lanes,
 * leaf_regroup_lanes, node_steps, workgroups_per_cu, split_percent,
 * samples_per_job, cost_order and, with ORN_TUNE_ALLOW_BVH4, bvh_width;
 * then the neighbours of the best regroup_lanes,
 * leaf_regroup_lanes and node_steps),
 * sets the fastest on the target and returns it with its time.
 * The target's accumulation, RNG state and framebuffer are left as they
 * were. With a cache_path (a JSON file, created if absent), a tuning stored
 * for the same device architecture and scene class (scene features, BVH
 * sizes, pixel count, samples and depth, each bucketed) is applied without
 * measuring, and a new measurement is stored. Not thread-safe across
 * processes sharing one cache file. */
orn_status orn_target_autotune(orn_target *target, const orn_constant_params *params,
                               uint32_t iterations, uint32_t flags, const char *cache_path,
                               orn_tuning *chosen, float *ms);
/* The same for a batch; params and iterations as in orn_batch_render. */
orn_status orn_batch_autotune(orn_batch *batch, const orn_constant_params *params,
                              const uint32_t *iterations, uint32_t flags, const char *cache_path,
                              orn_tuning *chosen, float *ms);

/* ------------------------------------------------------------------ */
/* Neural radiance cache (no counterpart in the reference: an addition  */
/* that changes the method; the plain tracer above stays the reference) */
/* ------------------------------------------------------------------ */

/* A multiresolution hash grid (16 levels x 2 FP16 features, 2^log2_table
 * entries per level) over the scene box feeds a 5 x 64 ReLU MLP (BF16
 * weights and activations, FP32 accumulation) that predicts the value the
 * reference estimator (ray_color, pathtracer.wgsl:95-161) would return for
 * the rest of a path from a shading point. orn_nrc_render traces each
 * sample with the reference arithmetic up to the hit of bounce
 * terminate_bounce and, on a scattering surface, replaces the rest of the
 * path with throughput x prediction. orn_nrc_train traces train_paths paths
 * with their own RNG streams to full depth, records the inputs and the
 * estimator's value at that vertex, and takes Adam steps (relative L2 loss;
 * hash-grid parameters whose gradient in a step is exactly 0 skip that step,
 * keeping value and moments, as in Instant-NGP). */
typedef struct orn_nrc orn_nrc;
typedef struct orn_nrc_config {
    uint32_t log2_table;       /* default 19 */
    uint32_t base_resolution;  /* coarsest grid, cells per axis (default 16) */
    uint32_t max_resolution;   /* finest grid (default 2048) */
    uint32_t terminate_bounce; /* bounce whose hit is queried (default 1; 0: primary hit) */
    uint32_t train_paths;      /* training paths per orn_nrc_train (default 65536) */
    uint32_t train_batch;      /* samples per Adam step, a multiple of 64 (default 16384); a step
                                * of a frame with fewer training records takes each record once */
    float learning_rate;       /* default 1e-2 */
    float beta1, beta2, epsilon; /* default 0.9, 0.99, 1e-15 */
    uint32_t seed;             /* initial parameters (default 1) */
    uint32_t reserved[5];
} orn_nrc_config;
/* One query or training sample: the cache's inputs at a shading point and
 * the path throughput up to it (queries) or the target value (training). */
typedef struct orn_nrc_record {
    float pos[3];
    uint32_t pixel;
    float dir[3];      /* towards the viewer: -normalize(ray direction) */
    float roughness;   /* metal fuzz, else 0 */
    float normal[3];   /* front-facing shading normal */
    uint32_t type;     /* material type 0..3 */
    float albedo[3];   /* albedo or texel; 1 for dielectrics */
    uint32_t flags;
    float weight[3];
    uint32_t pad;
} orn_nrc_record;
typedef struct orn_nrc_info_t {
    uint32_t levels;
    uint32_t log2_table;
    uint32_t resolution[16]; /* cells per axis of each level */
    uint32_t dense[16];      /* 1: level indexed directly, 0: hashed */
    float bounds_lo[3];      /* position -> [0,1]^3: (p - lo) * inv_extent, clamped */
    float inv_extent[3];
    uint32_t mlp_params;     /* 21,504: W0..W4 64 x 64 then W5 64 x 16, each [in][out] */
    uint64_t grid_params;    /* levels x 2^log2_table x 2 */
    uint32_t adam_steps;
    uint32_t reserved[7];
} orn_nrc_info_t;
orn_status orn_nrc_default_config(orn_nrc_config *config);
orn_status orn_nrc_create(orn_device_scene *scene, const orn_nrc_config *config, orn_nrc **out);
orn_status orn_nrc_destroy(orn_nrc *nrc);
orn_status orn_nrc_get_info(orn_nrc *nrc, orn_nrc_info_t *info);
/* One render() (params and iterations as for orn_render on the same target;
 * the target must belong to the cache's scene) with cached termination. */
orn_status orn_nrc_render(orn_nrc *nrc, orn_target *target, const orn_constant_params *params,
                          uint32_t iterations);
/* Training paths of frame `frame` (pixels and RNG streams hashed from it)
 * and `steps` Adam steps of train_batch samples; loss (optional) returns the
 * mean loss of the steps. */
orn_status orn_nrc_train(orn_nrc *nrc, const orn_constant_params *params, uint32_t frame,
                         uint32_t steps, float *loss);
/* One frame: the training paths of `frame` on a second stream while the
 * render (as orn_nrc_render) runs, then `steps` Adam steps; the render uses
 * the cache as it was before this frame's steps. */
orn_status orn_nrc_frame(orn_nrc *nrc, orn_target *target, const orn_constant_params *params,
                         uint32_t iterations, uint32_t frame, uint32_t steps, float *loss);
/* GPU milliseconds of the latest orn_nrc_frame, from its start to its end. */
orn_status orn_nrc_frame_ms(orn_nrc *nrc, float *ms);
/* GPU milliseconds of the latest orn_nrc_render and orn_nrc_train. */
orn_status orn_nrc_last_ms(orn_nrc *nrc, float *render_ms, float *train_ms);
/* Parameters (FP32 masters; the kernels read BF16 / FP16 copies made from
 * them). mlp: 21,504 floats; grid: [level][entry][feature]. */
orn_status orn_nrc_get_params(orn_nrc *nrc, float *mlp, size_t mlp_len, float *grid, size_t grid_len);
orn_status orn_nrc_set_params(orn_nrc *nrc, const float *mlp, size_t mlp_len, const float *grid,
                              size_t grid_len);
/* Raw predictions (3 floats per record) and, optionally, the network's 64
 * inputs per record before BF16 rounding. */
orn_status orn_nrc_query(orn_nrc *nrc, const orn_nrc_record *records, size_t n, float *rgb,
                         float *inputs);
/* Loss and gradients over host records (weight = target), without an
 * update: mlp_grad 21,504 floats, grid_grad (optional) grid_params floats. */
orn_status orn_nrc_gradients(orn_nrc *nrc, const orn_nrc_record *records, size_t n, float *loss,
                             float *mlp_grad, float *grid_grad, size_t grid_len);
/* One Adam step on host records. */
orn_status orn_nrc_step(orn_nrc *nrc, const orn_nrc_record *records, size_t n, float *loss);

/* ------------------------------------------------------------------ */
/* Ray queries                                                         */
/* ------------------------------------------------------------------ */

/* Closest hit of n world-space rays (6 floats each: origin, direction) with
 * t_min = ray_cast_epsilon, by the same traversal the render kernel runs
 * (bvh_hit, bvh.wgsl:14-137). Host arrays in and out; blocks. */
orn_status orn_trace_rays(orn_device_scene *scene, float t_min, const float *rays, size_t n,
                          orn_hit *out);
/* The same closest hits (bit for bit, BVH2) with each wave's 64 consecutive
 * rays traced together: the wave visits the union of their nodes in the
 * reference's order, each ray testing only the nodes its own traversal
 * visits. Fast when consecutive rays are coherent (a camera tile), slower
 * for scattered rays. No counterpart in the reference. */
orn_status orn_trace_rays_packet(orn_device_scene *scene, float t_min, const float *rays, size_t n,
                                 orn_hit *out);

/* Evaluates the device's FP32 transcendental functions (the kernels' stand-ins
 * for the WGSL builtins pow, sin, cos, acos, atan2) over host arrays:
 * op 0 sin(a), 1 cos(a), 2 pow(a, b), 3 acos(a), 4 atan2(a, b); and op 5 the
 * radiance cache's hardware bf16 rounding of an activation, returned as the
 * float the rounded value stands for. */
orn_status orn_eval_math(orn_ctx *ctx, int op, const float *a, const float *b, float *out,
                         size_t n);

/* Measurement aid: copies `bytes` (a multiple of 16) from one device buffer to
 * another with a plain 16-byte-per-lane copy kernel and reports its GPU time.
 * Profilers use it to scale their traffic counters against a known size, and
 * it gives the achievable HBM bandwidth of the device. */
orn_status orn_calibration_copy(orn_ctx *ctx, size_t bytes, float *ms);

#ifdef __cplusplus
}
#endif

#endif
